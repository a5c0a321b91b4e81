"""ctypes declarations for libkrt.so (include/krt.h)."""
from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libkrt.so"

KRT_OK = 0
KRT_ERR_INVALID = 1
KRT_ERR_HIP = 2
KRT_ERR_BUDGET = 3
KRT_ERR_STATE = 4

OPT_NONE, OPT_INT, OPT_FLOAT, OPT_BOOL, OPT_STRING = 0, 1, 2, 3, 4
LAYOUT_SAH, LAYOUT_REFERENCE = 0, 1

fp = C.POINTER(C.c_float)
u32p = C.POINTER(C.c_uint32)
i32p = C.POINTER(C.c_int32)
vp = C.c_void_p
handle = C.c_uint64


class InstancedDesc(C.Structure):
    _fields_ = [
        ("positions", C.POINTER(C.c_float)), ("normals", C.POINTER(C.c_float)),
        ("mesh_nv", C.POINTER(C.c_uint32)), ("indices", C.POINTER(C.c_uint32)),
        ("mesh_nt", C.POINTER(C.c_uint32)), ("nmesh", C.c_uint32),
        ("instance_mesh", C.POINTER(C.c_uint32)), ("transforms", C.POINTER(C.c_float)), ("ninst", C.c_uint32),
        ("tri_material", C.POINTER(C.c_uint32)), ("materials", C.POINTER(C.c_float)), ("nm", C.c_uint32),
    ]


class SceneDesc(C.Structure):
    _fields_ = [
        ("positions", fp), ("normals", fp), ("nv", C.c_uint32),
        ("indices", u32p), ("tri_material", u32p), ("nt", C.c_uint32),
        ("materials", fp), ("nm", C.c_uint32),
    ]


class SceneTextures(C.Structure):
    _fields_ = [
        ("uvs", fp), ("tangents", fp), ("nv", C.c_uint32),
        ("texels", u32p), ("dims", u32p), ("ntex", C.c_uint32),
        ("material_textures", C.POINTER(C.c_int32)), ("nm", C.c_uint32),
    ]


class BvhStats(C.Structure):
    _fields_ = [
        ("triangles", C.c_uint32), ("serialized_nodes", C.c_uint32), ("traversal_nodes", C.c_uint32),
        ("depth_bound", C.c_uint32), ("build_levels", C.c_uint32), ("build_ms", C.c_float),
        ("layout", C.c_int), ("binary_nodes", C.c_uint32), ("node_bytes", C.c_uint32),
        ("refit_ms", C.c_float),
    ]


class Camera(C.Structure):
    _fields_ = [("center", C.c_float * 3), ("U", C.c_float * 3), ("V", C.c_float * 3), ("W", C.c_float * 3),
                ("pv", C.c_float * 16)]


class Option(C.Structure):
    _fields_ = [("kind", C.c_int), ("i", C.c_int), ("f", C.c_float), ("b", C.c_int), ("s", C.c_char_p)]


class Buffers(C.Structure):
    _fields_ = [("color", vp), ("normal", vp), ("albedo", vp), ("position", vp)]


class MlpConfig(C.Structure):
    _fields_ = [("lr", C.c_float), ("beta1", C.c_float), ("beta2", C.c_float), ("eps", C.c_float),
                ("loss_scale", C.c_float), ("growth_interval", C.c_int32), ("pos_min", C.c_float * 3),
                ("pos_scale", C.c_float * 3)]


class MlpState(C.Structure):
    _fields_ = [("loss_scale", C.c_float), ("beta1_t", C.c_float), ("beta2_t", C.c_float),
                ("last_loss", C.c_float), ("step", C.c_int32), ("good_steps", C.c_int32),
                ("skipped_steps", C.c_int32), ("pending_nonfinite", C.c_int32)]


MLP_RAW, MLP_OUT, MLP_PARAMS = 13, 3, 20867


class ServerConfig(C.Structure):
    _fields_ = [("hbm_budget", C.c_uint64), ("max_active", C.c_uint32), ("use_graph", C.c_int32),
                ("max_queue_ms", C.c_double), ("step_pixels", C.c_uint64),
                ("wave_pixels", C.c_uint64)]


class RestirOptions(C.Structure):
    _fields_ = [("candidates", C.c_int32), ("m_cap", C.c_int32), ("spatial_passes", C.c_int32),
                ("neighbours", C.c_int32), ("radius", C.c_int32)]


KRT_REQUEST_PATH_TRACER = 0
KRT_REQUEST_RESTIR = 1


class Request(C.Structure):
    _fields_ = [("scene", C.c_uint64), ("camera", Camera), ("width", C.c_uint32), ("height", C.c_uint32),
                ("spp", C.c_uint32), ("max_depth", C.c_int32), ("seed", C.c_uint32), ("kind", C.c_int32),
                ("restir", RestirOptions)]


class ServerStats(C.Structure):
    _fields_ = [("submitted", C.c_uint64), ("completed", C.c_uint64), ("rejected", C.c_uint64),
                ("steps", C.c_uint64), ("active", C.c_uint32), ("queued", C.c_uint32),
                ("max_concurrent", C.c_uint32), ("reserved", C.c_uint32), ("p50_ms", C.c_double),
                ("p99_ms", C.c_double), ("step_p50_ms", C.c_double), ("step_p99_ms", C.c_double),
                ("samples", C.c_double), ("busy_ms", C.c_double), ("wall_ms", C.c_double),
                ("hbm_scratch", C.c_uint64), ("hbm_in_use", C.c_uint64), ("hbm_peak", C.c_uint64)]


class FrameStats(C.Structure):
    _fields_ = [("closest_rays", C.c_uint64), ("shadow_rays", C.c_uint64), ("samples", C.c_uint64),
                ("gpu_ms", C.c_float), ("kernel_ms", C.c_float * 8), ("pass_rays", C.c_uint64 * 4),
                ("node_visits", C.c_uint64), ("tri_tests", C.c_uint64), ("pass_nodes", C.c_uint64 * 4),
                ("pass_tris", C.c_uint64 * 4)]


class AttachmentRet(C.Structure):
    _fields_ = [("name", C.c_char_p), ("ptr", vp)]


_SIGS = {
    "krt_last_error": (C.c_char_p, []),
    "krt_create": (C.c_int, [C.c_int, C.POINTER(vp)]),
    "krt_destroy": (None, [vp]),
    "krt_scene_upload": (C.c_int, [vp, C.POINTER(SceneDesc), C.POINTER(handle)]),
    "krt_scene_upload_textured": (C.c_int, [vp, C.POINTER(SceneDesc), C.POINTER(SceneTextures), C.POINTER(handle)]),
    "krt_scene_free": (C.c_int, [vp, handle]),
    "krt_scene_light_count": (C.c_int, [vp, handle, u32p]),
    "krt_scene_lights": (C.c_int, [vp, handle, fp, C.c_size_t]),
    "krt_accel_build": (C.c_int, [vp, handle, C.c_int]),
    "krt_accel_build_layout": (C.c_int, [vp, handle, C.c_int, C.c_int]),
    "krt_accel_stats": (C.c_int, [vp, handle, C.POINTER(BvhStats)]),
    "krt_scene_update_vertices": (C.c_int, [vp, handle, C.c_uint32, C.c_uint32, fp, fp]),
    "krt_accel_refit": (C.c_int, [vp, handle]),
    "krt_accel_serialized": (C.c_int, [vp, handle, fp, C.c_size_t]),
    "krt_scene_upload_instanced": (C.c_int, [vp, C.POINTER(InstancedDesc), C.POINTER(handle)]),
    "krt_accel_mesh_serialized": (C.c_int, [vp, handle, C.c_uint32, fp, C.c_size_t, u32p]),
    "krt_accel_instance_tree": (C.c_int, [vp, handle, fp, C.c_size_t, u32p]),
    "krt_accel_leaf_order": (C.c_int, [vp, handle, i32p, C.c_size_t]),
    "krt_trace_closest": (C.c_int, [vp, handle, fp, C.c_uint32, fp]),
    "krt_trace_occluded": (C.c_int, [vp, handle, fp, C.c_uint32, i32p]),
    "krt_trace_closest_device": (C.c_int, [vp, handle, vp, C.c_uint32, vp]),
    "krt_trace_closest_device_noalloc": (C.c_int, [vp, handle, vp, C.c_uint32, vp]),
    "krt_write_png": (C.c_int, [C.c_char_p, C.POINTER(C.c_uint32), C.c_uint32, C.c_uint32]),
    "krt_read_png": (C.c_int, [C.c_char_p, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32),
                               C.c_size_t]),
    "krt_armada_write_png": (C.c_int, [vp, C.c_int, C.c_char_p]),
    "krt_camera_rays": (C.c_int, [vp, C.POINTER(Camera), C.c_uint32, C.c_uint32, C.c_int, C.c_int, C.c_uint32,
                                  fp]),
    "krt_copy_probe": (C.c_int, [vp, C.c_size_t, C.c_int, C.POINTER(C.c_float)]),
    "krt_timer_begin": (C.c_int, [vp]),
    "krt_timer_end": (C.c_int, [vp, C.POINTER(C.c_float)]),
    "krt_sync": (C.c_int, [vp]),
    "krt_buffer_create": (C.c_int, [vp, C.c_size_t, C.POINTER(vp)]),
    "krt_buffer_destroy": (C.c_int, [vp, vp]),
    "krt_buffer_upload": (C.c_int, [vp, vp, vp, C.c_size_t]),
    "krt_buffer_download": (C.c_int, [vp, vp, vp, C.c_size_t]),
    "krt_armada_create": (C.c_int, [vp, C.c_uint32, C.c_uint32, C.POINTER(vp)]),
    "krt_armada_destroy": (None, [vp]),
    "krt_attachment_create": (C.c_int, [vp, C.c_char_p, C.POINTER(vp)]),
    "krt_attachment_destroy": (None, [vp]),
    "krt_armada_attach": (C.c_int, [vp, C.c_char_p, vp]),
    "krt_armada_activate": (C.c_int, [vp, C.c_char_p]),
    "krt_armada_set_option": (C.c_int, [vp, C.c_char_p, C.POINTER(Option)]),
    "krt_armada_get_option": (C.c_int, [vp, C.c_char_p, C.POINTER(Option)]),
    "krt_armada_set_depth": (C.c_int, [vp, C.c_int]),
    "krt_armada_set_envmap": (C.c_int, [vp, fp, C.c_uint32, C.c_uint32]),
    "krt_armada_set_envmap_enabled": (C.c_int, [vp, C.c_int]),
    "krt_armada_render": (C.c_int, [vp, handle, C.POINTER(Camera), C.c_int]),
    "krt_armada_buffers": (C.c_int, [vp, C.POINTER(Buffers)]),
    "krt_armada_readback": (C.c_int, [vp, C.c_int, fp]),
    "krt_armada_upload": (C.c_int, [vp, C.c_int, fp]),
    "krt_armada_stats": (C.c_int, [vp, C.POINTER(FrameStats)]),
    "krt_armada_sync": (C.c_int, [vp]),
    "krt_armada_set_f64_accumulation": (C.c_int, [vp, C.c_int]),
    "krt_armada_read_reservoirs": (C.c_int, [vp, C.c_int, fp]),
    "krt_armada_read_f64": (C.c_int, [vp, C.POINTER(C.c_double), C.POINTER(C.c_uint64)]),
    "krt_armada_read_f64_squares": (C.c_int, [vp, C.POINTER(C.c_double)]),
    "krt_hdr_to_ldr": (C.c_int, [vp, vp, vp, C.c_uint32, C.c_uint32, C.c_int]),
    "krt_armada_read_ldr": (C.c_int, [vp, C.c_int, C.POINTER(C.c_uint32)]),
    "krt_load_attachment": (AttachmentRet, [vp, C.c_char_p]),
    "krt_mlp_default_config": (None, [C.POINTER(MlpConfig)]),
    "krt_mlp_create": (C.c_int, [vp, C.POINTER(MlpConfig), C.c_uint64, C.POINTER(vp)]),
    "krt_mlp_destroy": (None, [vp]),
    "krt_mlp_get_params": (C.c_int, [vp, fp]),
    "krt_mlp_set_params": (C.c_int, [vp, fp]),
    "krt_mlp_set_encoding": (C.c_int, [vp, fp, fp]),
    "krt_mlp_infer": (C.c_int, [vp, fp, C.c_uint32, fp]),
    "krt_mlp_infer_device": (C.c_int, [vp, vp, C.c_uint32, vp]),
    "krt_mlp_train_step": (C.c_int, [vp, fp, fp, C.c_uint32, fp]),
    "krt_mlp_train_step_device": (C.c_int, [vp, vp, vp, C.c_uint32]),
    "krt_mlp_gradient": (C.c_int, [vp, fp, fp, C.c_uint32, fp, fp]),
    "krt_mlp_apply_gradient": (C.c_int, [vp, fp]),
    "krt_mlp_get_state": (C.c_int, [vp, C.POINTER(MlpState)]),
    "krt_attachment_cache": (C.c_int, [vp, C.POINTER(vp)]),
    "krt_server_default_config": (None, [C.POINTER(ServerConfig)]),
    "krt_server_create": (C.c_int, [vp, C.POINTER(ServerConfig), C.POINTER(vp)]),
    "krt_server_destroy": (None, [vp]),
    "krt_server_submit": (C.c_int, [vp, C.POINTER(Request), C.POINTER(C.c_uint64)]),
    "krt_server_wait": (C.c_int, [vp, C.c_uint64, fp, C.POINTER(C.c_double)]),
    "krt_server_drain": (C.c_int, [vp]),
    "krt_server_get_stats": (C.c_int, [vp, C.POINTER(ServerStats)]),
}

EXPORTED = tuple(_SIGS)

_lib = None


class KrtError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"krt error {code}: {msg}")
        self.code = code


def lib():
    """Load libkrt.so; raises if it has not been built (no fallback path)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} is not built; run `python -m artemis_amd.build` first")
        L = C.CDLL(str(LIB_PATH))
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(code: int) -> None:
    if code != KRT_OK:
        raise KrtError(code, lib().krt_last_error().decode())
