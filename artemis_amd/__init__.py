"""artemis_amd — MI355X-native ray-traced rendering path of the Kobra renderer.

Python mirror of the reference's ArmadaRTX / AttachmentRTX / Accelerator API
(include/amadeus/armada.cuh:107-336, include/amadeus/accelerator.cuh:133-277)
over the C ABI of include/krt.h (libkrt.so, HIP for gfx950).  Every method is a
thin call into the library; there is no CPU fallback: a missing or failing
library raises.
"""
from __future__ import annotations

import ctypes as C
import weakref

import numpy as np

from . import _lib as L
from .camera import CameraFrame, look_at
from .scenes import SceneData, pack_textures

__all__ = ["Context", "Scene", "ArmadaRTX", "PathTracer", "ReSTIR", "ReSTIRGI", "Mamba", "CameraFrame", "look_at", "SceneData",
           "load_attachment", "RadianceField", "Server", "DeviceArray"]


def _f(a):
    return a.ctypes.data_as(L.fp)


class Context:
    """krt_create / krt_destroy: one HIP device, one stream.  Objects created
    on a context are released before it (close() releases them first)."""

    def __init__(self, device: int = 0):
        self._lib = L.lib()
        h = C.c_void_p()
        L.check(self._lib.krt_create(device, C.byref(h)))
        self.h = h
        self._children = weakref.WeakSet()

    def _adopt(self, obj):
        self._children.add(obj)

    def close(self):
        if self.h:
            for ch in list(self._children):
                ch._release()
            self._lib.krt_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def copy_probe(self, nbytes: int, reps: int = 10) -> float:
        """Average ms of a device-to-device streaming copy kernel of nbytes."""
        ms = C.c_float()
        L.check(self._lib.krt_copy_probe(self.h, nbytes, reps, C.byref(ms)))
        return ms.value

    def timer_begin(self) -> None:
        """Start a device-time measurement on the context's stream."""
        L.check(self._lib.krt_timer_begin(self.h))

    def timer_end(self) -> float:
        """Milliseconds of device time since timer_begin (HIP events)."""
        ms = C.c_float()
        L.check(self._lib.krt_timer_end(self.h, C.byref(ms)))
        return ms.value

    def sync(self) -> None:
        L.check(self._lib.krt_sync(self.h))

    def hdr_to_ldr(self, pixels_ptr: int, target_ptr: int, width: int, height: int, tonemapping: int = 0) -> None:
        """hdr_to_ldr (include/cuda/color.cuh:82-121) on device buffers: width x
        height float4 -> packed RGBA8 uint32, rows flipped; 0 = ACES, 1 = sRGB.
        Queued on the context stream."""
        L.check(self._lib.krt_hdr_to_ldr(self.h, pixels_ptr, target_ptr, width, height, tonemapping))

    def device_array(self, data=None, shape=None, dtype=np.float32) -> "DeviceArray":
        """A device buffer, filled from `data` or of `shape` x `dtype`."""
        return DeviceArray(self, data, shape, dtype)

    def camera_rays(self, camera: CameraFrame, width: int, height: int, jitter: bool = False, sample: int = 0,
                    user_seed: int = 0) -> np.ndarray:
        """make_ray on the device for every pixel (8 floats per ray)."""
        out = np.empty((width * height, 8), np.float32)
        cam = camera.to_ctypes()
        L.check(self._lib.krt_camera_rays(self.h, C.byref(cam), width, height, int(jitter), sample, user_seed,
                                          _f(out)))
        return out

    def upload(self, scene, leaf_max: int = 4, build: bool = True, layout: str = "sah") -> "Scene":
        """A SceneData (world-space soup) or an InstancedSceneData (two-level)."""
        return Scene(self, scene, leaf_max=leaf_max, build=build, layout=layout)


class DeviceArray:
    """Device memory owned through krt_buffer_*; .ptr for *_device calls."""

    def __init__(self, ctx: Context, data=None, shape=None, dtype=np.float32):
        if data is not None:
            data = np.ascontiguousarray(data)
            shape, dtype = data.shape, data.dtype
        self.ctx, self.shape, self.dtype = ctx, tuple(shape), np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        h = C.c_void_p()
        L.check(ctx._lib.krt_buffer_create(ctx.h, max(self.nbytes, 1), C.byref(h)))
        self.h = h
        ctx._adopt(self)
        if data is not None:
            self.upload(data)

    @property
    def ptr(self) -> int:
        return self.h.value

    def upload(self, data) -> None:
        data = np.ascontiguousarray(data, self.dtype)
        assert data.nbytes == self.nbytes
        L.check(self.ctx._lib.krt_buffer_upload(self.ctx.h, self.h, data.ctypes.data_as(C.c_void_p), self.nbytes))

    def download(self) -> np.ndarray:
        out = np.empty(self.shape, self.dtype)
        L.check(self.ctx._lib.krt_buffer_download(self.ctx.h, out.ctypes.data_as(C.c_void_p), self.h, self.nbytes))
        return out

    def _release(self):
        if self.h is not None and self.ctx.h:
            self.ctx._lib.krt_buffer_destroy(self.ctx.h, self.h)
        self.h = None

    def __del__(self):
        try:
            self._release()
        except Exception:
            pass


class Scene:
    """A resident scene and its acceleration structure (the "System" layer)."""

    def __init__(self, ctx: Context, data: SceneData, leaf_max: int = 4, build: bool = True, layout: str = "sah"):
        self.ctx = ctx
        self.data = data
        lib = ctx._lib
        self.instanced = hasattr(data, "inst_mesh")
        if self.instanced:
            self._upload_instanced(data)
            ctx._adopt(self)
            if build:
                self.build(leaf_max, layout)
            return
        self._keep = [np.ascontiguousarray(data.positions, np.float32),
                      np.ascontiguousarray(data.normals, np.float32),
                      np.ascontiguousarray(data.indices, np.uint32),
                      np.ascontiguousarray(data.tri_material, np.uint32),
                      np.ascontiguousarray(data.materials, np.float32)]
        p, n, i, m, mats = self._keep
        d = L.SceneDesc(_f(p), _f(n), len(p), i.ctypes.data_as(L.u32p), m.ctypes.data_as(L.u32p), len(i),
                        _f(mats), len(mats))
        h = L.handle()
        if getattr(data, "textures", None) is not None:
            # krt_scene_upload_textured: tex_coords, tangent frames, RGBA8 images
            uv = np.ascontiguousarray(data.uvs, np.float32)
            tan = None if data.tangents is None else np.ascontiguousarray(data.tangents, np.float32)
            texels, dims = pack_textures(data.textures)
            mt = np.ascontiguousarray(data.material_textures, np.int32)
            self._keep += [uv, tan, texels, dims, mt]
            t = L.SceneTextures(_f(uv), None if tan is None else _f(tan), len(uv), texels.ctypes.data_as(L.u32p),
                                dims.ctypes.data_as(L.u32p), len(dims), mt.ctypes.data_as(C.POINTER(C.c_int32)),
                                len(mats))
            L.check(lib.krt_scene_upload_textured(ctx.h, C.byref(d), C.byref(t), C.byref(h)))
        else:
            L.check(lib.krt_scene_upload(ctx.h, C.byref(d), C.byref(h)))
        self.h = h.value
        ctx._adopt(self)
        if build:
            self.build(leaf_max, layout)

    def _upload_instanced(self, d) -> None:
        """krt_scene_upload_instanced (Accelerator::update, accelerator.cuh:145-190)."""
        u32 = lambda a: a.ctypes.data_as(L.u32p)
        pos = np.ascontiguousarray(np.concatenate([m[0] for m in d.meshes]), np.float32)
        nrm = np.ascontiguousarray(np.concatenate([m[1] for m in d.meshes]), np.float32)
        idx = np.ascontiguousarray(np.concatenate([m[2] for m in d.meshes]), np.uint32)
        nv = np.array([len(m[0]) for m in d.meshes], np.uint32)
        nt = np.array([len(m[2]) for m in d.meshes], np.uint32)
        im = np.ascontiguousarray(d.inst_mesh, np.uint32)
        xf = np.ascontiguousarray(d.transforms, np.float32).reshape(-1, 12)
        tm = np.ascontiguousarray(d.tri_material, np.uint32)
        mats = np.ascontiguousarray(d.materials, np.float32)
        self._keep = [pos, nrm, idx, nv, nt, im, xf, tm, mats]
        desc = L.InstancedDesc(_f(pos), _f(nrm), u32(nv), u32(idx), u32(nt), len(nv), u32(im), _f(xf), len(im),
                               u32(tm), _f(mats), len(mats))
        h = L.handle()
        L.check(self.ctx._lib.krt_scene_upload_instanced(self.ctx.h, C.byref(desc), C.byref(h)))
        self.h = h.value

    def mesh_serialized(self, mesh: int) -> np.ndarray:
        """Instanced scenes: the bvh.cpp tree of one base mesh (12 floats per node)."""
        n = C.c_uint32()
        L.check(self.ctx._lib.krt_accel_mesh_serialized(self.ctx.h, self.h, mesh, None, 0, C.byref(n)))
        out = np.empty((n.value, 12), np.float32)
        L.check(self.ctx._lib.krt_accel_mesh_serialized(self.ctx.h, self.h, mesh, _f(out), out.size, C.byref(n)))
        return out

    def instance_tree(self) -> np.ndarray:
        """Instanced scenes: the bvh.cpp tree over the instances' world boxes."""
        n = C.c_uint32()
        L.check(self.ctx._lib.krt_accel_instance_tree(self.ctx.h, self.h, None, 0, C.byref(n)))
        out = np.empty((n.value, 12), np.float32)
        L.check(self.ctx._lib.krt_accel_instance_tree(self.ctx.h, self.h, _f(out), out.size, C.byref(n)))
        return out

    def build(self, leaf_max: int = 4, layout: str = "sah") -> None:
        """Accelerator::update + build_tlas: GPU build of the bvh.cpp tree plus
        the traversal tree ("sah", or "reference" = collapse of the bvh.cpp tree)."""
        code = {"sah": L.LAYOUT_SAH, "reference": L.LAYOUT_REFERENCE}[layout]
        L.check(self.ctx._lib.krt_accel_build_layout(self.ctx.h, self.h, leaf_max, code))

    def update_vertices(self, first: int, positions: np.ndarray, normals: np.ndarray | None = None) -> None:
        """Moves vertices [first, first + len(positions)) (world space); the
        scene then needs refit() or build() before it is rendered again."""
        p = np.ascontiguousarray(positions, np.float32).reshape(-1, 3)
        n = None if normals is None else np.ascontiguousarray(normals, np.float32).reshape(-1, 3)
        if n is not None and len(n) != len(p):
            raise ValueError("positions and normals differ in length")
        L.check(self.ctx._lib.krt_scene_update_vertices(self.ctx.h, self.h, first, len(p), _f(p),
                                                        None if n is None else _f(n)))

    def refit(self) -> None:
        """Recomputes the trees' boxes, triangle records and lights for moved
        vertices, keeping the trees' shapes."""
        L.check(self.ctx._lib.krt_accel_refit(self.ctx.h, self.h))

    def stats(self) -> dict:
        s = L.BvhStats()
        L.check(self.ctx._lib.krt_accel_stats(self.ctx.h, self.h, C.byref(s)))
        return {k: getattr(s, k) for k, _ in L.BvhStats._fields_}

    def serialized(self) -> np.ndarray:
        st = self.stats()
        out = np.empty((st["serialized_nodes"], 12), np.float32)
        L.check(self.ctx._lib.krt_accel_serialized(self.ctx.h, self.h, _f(out), out.size))
        return out

    def leaf_order(self) -> np.ndarray:
        out = np.empty(len(self.data.indices), np.int32)
        L.check(self.ctx._lib.krt_accel_leaf_order(self.ctx.h, self.h, out.ctypes.data_as(L.i32p), out.size))
        return out

    def lights(self) -> np.ndarray:
        n = C.c_uint32()
        L.check(self.ctx._lib.krt_scene_light_count(self.ctx.h, self.h, C.byref(n)))
        out = np.empty((n.value, 12), np.float32)
        if n.value:
            L.check(self.ctx._lib.krt_scene_lights(self.ctx.h, self.h, _f(out), out.size))
        return out

    def trace_closest(self, rays: np.ndarray) -> np.ndarray:
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        hits = np.empty((len(rays), 4), np.float32)
        L.check(self.ctx._lib.krt_trace_closest(self.ctx.h, self.h, _f(rays), len(rays), _f(hits)))
        return hits

    def trace_occluded(self, rays: np.ndarray) -> np.ndarray:
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        out = np.empty(len(rays), np.int32)
        L.check(self.ctx._lib.krt_trace_occluded(self.ctx.h, self.h, _f(rays), len(rays),
                                                 out.ctypes.data_as(L.i32p)))
        return out

    def free(self):
        if self.h and self.ctx.h:
            L.check(self.ctx._lib.krt_scene_free(self.ctx.h, self.h))
        self.h = 0

    def _release(self):
        self.h = 0  # freed with the context


class _Attachment:
    kind = ""

    def __init__(self, ctx: Context):
        self.ctx = ctx
        h = C.c_void_p()
        L.check(ctx._lib.krt_attachment_create(ctx.h, self.kind.encode(), C.byref(h)))
        self.h = h
        self._users = weakref.WeakSet()
        ctx._adopt(self)

    def _release(self):
        if self.h:
            for u in list(self._users):
                u._release()
            self.ctx._lib.krt_attachment_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self._release()
        except Exception:
            pass

    def radiance_cache(self):
        """The radiance-cache network of this attachment (a RadianceField
        view), or None before the first cached frame."""
        h = C.c_void_p()
        L.check(self.ctx._lib.krt_attachment_cache(self.h, C.byref(h)))
        return RadianceField._borrow(self, h) if h.value else None


class PathTracer(_Attachment):
    """include/amadeus/path_tracer.cuh:17-167.  Options beyond the reference:
    radiance_cache (bool), cache_depth, train_samples, cache_training,
    cache_emission."""
    kind = "path_tracer"


class ReSTIR(_Attachment):
    """include/amadeus/restir.cuh:35-305.  With the option radiance_cache the
    indirect path stops at the first bounce and asks the cache (same options as
    PathTracer)."""
    kind = "restir"


class ReSTIRGI(_Attachment):
    """ReSTIR GI (no reference counterpart; Ouyang et al. 2021)."""
    kind = "restir_gi"


class Mamba(_Attachment):
    """The editor's G-buffer ReSTIR DI (editor/optix/mamba_shader.cu:35-505):
    8 light candidates, temporal reuse (M-cap 200), spatial reuse of 2
    neighbours in 20 px; options temporal, spatial (bool).  The colour buffer
    holds the accumulated direct lighting; reservoirs() gives 16 floats per
    pixel {light position, normal, emission, area, w, W, M, seed}."""
    kind = "mamba"


def write_png(path, rgba: np.ndarray) -> None:
    """(H, W) uint32 packed RGBA8 pixels (hdr_to_ldr's output) -> PNG file,
    rows in memory order."""
    rgba = np.ascontiguousarray(rgba, np.uint32)
    h, w = rgba.shape
    L.check(L.lib().krt_write_png(str(path).encode(), rgba.ctypes.data_as(C.POINTER(C.c_uint32)), w, h))


def read_png(path) -> np.ndarray:
    """PNG file -> (H, W, 4) uint8 RGBA, rows in file order (krt_read_png:
    every colour type and bit depth, interlacing, tRNS)."""
    w, h = C.c_uint32(), C.c_uint32()
    L.check(L.lib().krt_read_png(str(path).encode(), C.byref(w), C.byref(h), None, 0))
    out = np.empty((h.value, w.value), np.uint32)
    L.check(L.lib().krt_read_png(str(path).encode(), C.byref(w), C.byref(h),
                                 out.ctypes.data_as(C.POINTER(C.c_uint32)), out.size))
    return out.view(np.uint8).reshape(h.value, w.value, 4)


def load_attachment(ctx: Context, kind: str):
    """Plugin entry (source/amadeus/naive_path_tracer.cu:280-344)."""
    return {"path_tracer": PathTracer, "restir": ReSTIR, "restir_gi": ReSTIRGI, "mamba": Mamba}[kind](ctx)


class RadianceField:
    """Online-trained radiance-cache MLP over krt_mlp_* (the reference's
    RadianceField, editor/mlp.cuh:71-154, trained here).  Queries are
    (n, 13) float32 {pos, dir, normal, albedo, roughness}; outputs and
    targets (n, 3) float32."""

    RAW, OUT, PARAMS = L.MLP_RAW, L.MLP_OUT, L.MLP_PARAMS

    def __init__(self, ctx: Context, seed: int = 0, lr: float | None = None, loss_scale: float | None = None,
                 growth_interval: int | None = None, pos_min=(0.0, 0.0, 0.0), pos_scale=(1.0, 1.0, 1.0),
                 beta1: float | None = None, beta2: float | None = None, eps: float | None = None):
        self.ctx = ctx
        cfg = L.MlpConfig()
        ctx._lib.krt_mlp_default_config(C.byref(cfg))
        for name, val in (("lr", lr), ("loss_scale", loss_scale), ("growth_interval", growth_interval),
                          ("beta1", beta1), ("beta2", beta2), ("eps", eps)):
            if val is not None:
                setattr(cfg, name, val)
        cfg.pos_min[:] = [float(x) for x in pos_min]
        cfg.pos_scale[:] = [float(x) for x in pos_scale]
        self.config = cfg
        h = C.c_void_p()
        L.check(ctx._lib.krt_mlp_create(ctx.h, C.byref(cfg), seed, C.byref(h)))
        self.h = h
        ctx._adopt(self)

    @classmethod
    def _borrow(cls, owner, handle):
        """View of a network owned by an attachment (not destroyed here)."""
        obj = cls.__new__(cls)
        obj.ctx = owner.ctx
        obj.h = handle
        obj._owner = owner
        obj.config = None
        return obj

    def _release(self):
        if self.h and getattr(self, "_owner", None) is None:
            self.ctx._lib.krt_mlp_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self._release()
        except Exception:
            pass

    @staticmethod
    def _rows(a, width):
        return np.ascontiguousarray(a, np.float32).reshape(-1, width)

    def params(self) -> np.ndarray:
        out = np.empty(self.PARAMS, np.float32)
        L.check(self.ctx._lib.krt_mlp_get_params(self.h, _f(out)))
        return out

    def set_params(self, params) -> None:
        p = np.ascontiguousarray(params, np.float32)
        assert p.shape == (self.PARAMS,)
        L.check(self.ctx._lib.krt_mlp_set_params(self.h, _f(p)))

    def set_encoding(self, pos_min, pos_scale) -> None:
        a = np.ascontiguousarray(pos_min, np.float32)
        b = np.ascontiguousarray(pos_scale, np.float32)
        L.check(self.ctx._lib.krt_mlp_set_encoding(self.h, _f(a), _f(b)))

    def infer(self, raw) -> np.ndarray:
        raw = self._rows(raw, self.RAW)
        out = np.empty((len(raw), self.OUT), np.float32)
        L.check(self.ctx._lib.krt_mlp_infer(self.h, _f(raw), len(raw), _f(out)))
        return out

    def infer_device(self, raw_ptr: int, n: int, out_ptr: int) -> None:
        L.check(self.ctx._lib.krt_mlp_infer_device(self.h, C.c_void_p(raw_ptr), n, C.c_void_p(out_ptr)))

    def train_step(self, raw, target) -> float:
        raw, target = self._rows(raw, self.RAW), self._rows(target, self.OUT)
        assert len(raw) == len(target)
        loss = C.c_float()
        L.check(self.ctx._lib.krt_mlp_train_step(self.h, _f(raw), _f(target), len(raw), C.byref(loss)))
        return loss.value

    def train_step_device(self, raw_ptr: int, target_ptr: int, n: int) -> None:
        L.check(self.ctx._lib.krt_mlp_train_step_device(self.h, C.c_void_p(raw_ptr), C.c_void_p(target_ptr), n))

    def gradient(self, raw, target):
        """(loss, gradient) of one batch without an optimiser step."""
        raw, target = self._rows(raw, self.RAW), self._rows(target, self.OUT)
        g = np.empty(self.PARAMS, np.float32)
        loss = C.c_float()
        L.check(self.ctx._lib.krt_mlp_gradient(self.h, _f(raw), _f(target), len(raw), _f(g), C.byref(loss)))
        return loss.value, g

    def apply_gradient(self, grad) -> None:
        g = np.ascontiguousarray(grad, np.float32)
        assert g.shape == (self.PARAMS,)
        L.check(self.ctx._lib.krt_mlp_apply_gradient(self.h, _f(g)))

    def state(self) -> dict:
        st = L.MlpState()
        L.check(self.ctx._lib.krt_mlp_get_state(self.h, C.byref(st)))
        return {k: getattr(st, k) for k, _ in st._fields_}


class Server:
    """Render-request server over krt_server_* (continuous batching of
    path-tracer and ReSTIR DI requests on resident scenes, reservoirs
    resident per request, admission by HBM budget)."""

    def __init__(self, ctx: Context, hbm_budget: int | None = None, max_active: int | None = None,
                 use_graph: bool = True, max_queue_ms: float = 0.0, step_pixels: int | None = None,
                 wave_pixels: int | None = None):
        self.ctx = ctx
        cfg = L.ServerConfig()
        ctx._lib.krt_server_default_config(C.byref(cfg))
        if hbm_budget is not None:
            cfg.hbm_budget = hbm_budget
        if max_active is not None:
            cfg.max_active = max_active
        cfg.use_graph = int(use_graph)
        cfg.max_queue_ms = max_queue_ms
        if step_pixels is not None:
            cfg.step_pixels = step_pixels
        if wave_pixels is not None:
            cfg.wave_pixels = wave_pixels
        h = C.c_void_p()
        L.check(ctx._lib.krt_server_create(ctx.h, C.byref(cfg), C.byref(h)))
        self.h = h
        self._shape = {}
        ctx._adopt(self)

    def _release(self):
        if self.h:
            self.ctx._lib.krt_server_destroy(self.h)
            self.h = None

    close = _release

    def __del__(self):
        try:
            self._release()
        except Exception:
            pass

    def submit(self, scene: Scene, camera: CameraFrame, width: int, height: int, spp: int = 1,
               max_depth: int = 5, seed: int = 0, kind: str = "path_tracer", candidates: int = 32,
               m_cap: int = 200, spatial_passes: int = 2, neighbours: int = 5, radius: int = 30) -> int:
        """Queue a request: `spp` accumulated frames of the path tracer
        (kind "path_tracer") or of ReSTIR DI (kind "restir", with the restir
        attachment's options)."""
        r = L.Request()
        r.scene = scene.h
        r.camera = camera.to_ctypes()
        r.width, r.height, r.spp, r.max_depth, r.seed = width, height, spp, max_depth, seed
        if kind not in ("path_tracer", "restir"):
            raise ValueError(f"unknown request kind {kind!r}")
        r.kind = L.KRT_REQUEST_RESTIR if kind == "restir" else L.KRT_REQUEST_PATH_TRACER
        r.restir = L.RestirOptions(candidates, m_cap, spatial_passes, neighbours, radius)
        rid = C.c_uint64()
        L.check(self.ctx._lib.krt_server_submit(self.h, C.byref(r), C.byref(rid)))
        self._shape[rid.value] = (height, width)
        return rid.value

    def wait(self, rid: int):
        """(colour (H, W, 4) float32, latency ms), or (None, latency ms) if
        the request was rejected by the queueing deadline."""
        h, w = self._shape.pop(rid)
        out = np.empty((h, w, 4), np.float32)
        lat = C.c_double()
        rc = self.ctx._lib.krt_server_wait(self.h, rid, _f(out), C.byref(lat))
        if rc == L.KRT_ERR_BUDGET:
            return None, lat.value
        L.check(rc)
        return out, lat.value

    def drain(self) -> None:
        L.check(self.ctx._lib.krt_server_drain(self.h))

    def stats(self) -> dict:
        st = L.ServerStats()
        L.check(self.ctx._lib.krt_server_get_stats(self.h, C.byref(st)))
        return {k: getattr(st, k) for k, _ in st._fields_}


class ArmadaRTX:
    """include/amadeus/armada.cuh:142-336 over krt_armada_*."""

    def __init__(self, ctx: Context, width: int, height: int):
        self.ctx = ctx
        self.width, self.height = width, height
        h = C.c_void_p()
        L.check(ctx._lib.krt_armada_create(ctx.h, width, height, C.byref(h)))
        self.h = h
        self._attached = {}
        ctx._adopt(self)

    def _release(self):
        if self.h:
            self.ctx._lib.krt_armada_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self._release()
        except Exception:
            pass

    def size(self) -> int:
        return self.width * self.height

    def extent(self):
        return (self.width, self.height)

    def attach(self, name: str, attachment: _Attachment) -> None:
        L.check(self.ctx._lib.krt_armada_attach(self.h, name.encode(), attachment.h))
        self._attached[name] = attachment
        attachment._users.add(self)

    def attachments(self):
        return sorted(self._attached)

    def activate(self, name: str) -> None:
        L.check(self.ctx._lib.krt_armada_activate(self.h, name.encode()))

    def set_option(self, field: str, value) -> None:
        o = L.Option()
        if isinstance(value, bool):
            o.kind, o.b = L.OPT_BOOL, int(value)
        elif isinstance(value, int):
            o.kind, o.i = L.OPT_INT, value
        elif isinstance(value, float):
            o.kind, o.f = L.OPT_FLOAT, value
        else:
            o.kind, o.s = L.OPT_STRING, str(value).encode()
        L.check(self.ctx._lib.krt_armada_set_option(self.h, field.encode(), C.byref(o)))

    def get_option(self, field: str):
        o = L.Option()
        L.check(self.ctx._lib.krt_armada_get_option(self.h, field.encode(), C.byref(o)))
        return {L.OPT_INT: o.i, L.OPT_FLOAT: o.f, L.OPT_BOOL: bool(o.b)}.get(o.kind)

    def set_depth(self, depth: int) -> None:
        L.check(self.ctx._lib.krt_armada_set_depth(self.h, depth))

    def set_envmap(self, rgba: np.ndarray | None) -> None:
        """ArmadaRTX::set_envmap: an (H, W, 4) float32 environment map (row 0
        is v = 0), or None to remove it."""
        if rgba is None:
            L.check(self.ctx._lib.krt_armada_set_envmap(self.h, None, 0, 0))
            return
        a = np.ascontiguousarray(rgba, np.float32)
        if a.ndim != 3 or a.shape[2] != 4:
            raise ValueError("environment map must be (H, W, 4)")
        self._env = a
        L.check(self.ctx._lib.krt_armada_set_envmap(self.h, _f(a), a.shape[1], a.shape[0]))

    def set_envmap_enabled(self, enabled: bool) -> None:
        L.check(self.ctx._lib.krt_armada_set_envmap_enabled(self.h, int(bool(enabled))))

    def render(self, scene: Scene, camera: CameraFrame, accumulate: bool = True) -> None:
        cam = camera.to_ctypes()
        L.check(self.ctx._lib.krt_armada_render(self.h, scene.h, C.byref(cam), int(accumulate)))

    def _read(self, which: int) -> np.ndarray:
        out = np.empty((self.height, self.width, 4), np.float32)
        L.check(self.ctx._lib.krt_armada_readback(self.h, which, _f(out)))
        return out

    def color_buffer(self):
        return self._read(0)

    def normal_buffer(self):
        return self._read(1)

    def albedo_buffer(self):
        return self._read(2)

    def position_buffer(self):
        return self._read(3)

    def ldr_buffer(self, tonemapping: int = 0) -> np.ndarray:
        """The colour buffer through hdr_to_ldr: (H, W) uint32 RGBA8, rows
        flipped (ACES = 0, sRGB = 1)."""
        out = np.empty((self.height, self.width), np.uint32)
        L.check(self.ctx._lib.krt_armada_read_ldr(self.h, tonemapping, out.ctypes.data_as(C.POINTER(C.c_uint32))))
        return out

    def write_png(self, path, tonemapping: int = 0) -> None:
        """The colour buffer through hdr_to_ldr into a PNG file (the
        reference's still-image capture, bench.cu:226-233)."""
        L.check(self.ctx._lib.krt_armada_write_png(self.h, tonemapping, str(path).encode()))

    def upload_buffer(self, which: int, data: np.ndarray) -> None:
        data = np.ascontiguousarray(data, np.float32)
        assert data.size == self.width * self.height * 4
        L.check(self.ctx._lib.krt_armada_upload(self.h, which, _f(data)))

    def device_buffers(self) -> dict:
        b = L.Buffers()
        L.check(self.ctx._lib.krt_armada_buffers(self.h, C.byref(b)))
        return {"color": b.color, "normal": b.normal, "albedo": b.albedo, "position": b.position}

    def stats(self) -> dict:
        s = L.FrameStats()
        L.check(self.ctx._lib.krt_armada_stats(self.h, C.byref(s)))
        return {"closest_rays": s.closest_rays, "shadow_rays": s.shadow_rays, "samples": s.samples,
                "gpu_ms": s.gpu_ms, "kernel_ms": list(s.kernel_ms), "pass_rays": list(s.pass_rays),
                "node_visits": s.node_visits, "tri_tests": s.tri_tests, "pass_nodes": list(s.pass_nodes),
                "pass_tris": list(s.pass_tris)}

    def sync(self) -> None:
        L.check(self.ctx._lib.krt_armada_sync(self.h))

    def reservoirs(self, which: int = 0, gi: bool = False) -> np.ndarray:
        """ReSTIR reservoirs, 16 (DI) or 20 (GI) floats per pixel (krt_armada_read_reservoirs)."""
        out = np.empty((self.height, self.width, 20 if gi else 16), np.float32)
        L.check(self.ctx._lib.krt_armada_read_reservoirs(self.h, which, _f(out)))
        return out

    def set_f64_accumulation(self, enable: bool) -> None:
        """FP64 per-pixel sums of path-tracer samples (converged references)."""
        L.check(self.ctx._lib.krt_armada_set_f64_accumulation(self.h, int(enable)))

    def read_f64(self):
        out = np.empty((self.height, self.width, 3), np.float64)
        n = C.c_uint64()
        L.check(self.ctx._lib.krt_armada_read_f64(self.h, out.ctypes.data_as(C.POINTER(C.c_double)), C.byref(n)))
        return out, n.value

    def read_f64_squares(self) -> np.ndarray:
        """FP64 per-pixel sums of squared path-tracer samples (H, W, 3)."""
        out = np.empty((self.height, self.width, 3), np.float64)
        L.check(self.ctx._lib.krt_armada_read_f64_squares(self.h, out.ctypes.data_as(C.POINTER(C.c_double))))
        return out
