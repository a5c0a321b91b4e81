"""CPU oracle bindings — TEST INFRASTRUCTURE ONLY.

ctypes front end of liboracle.so (oracle.h).  Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg import this module,
and only as the checker; the product (artemis_amd/) never does.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"

fp = C.POINTER(C.c_float)
dp = C.POINTER(C.c_double)
u32p = C.POINTER(C.c_uint32)
i32p = C.POINTER(C.c_int32)
u64p = C.POINTER(C.c_uint64)
vp = C.c_void_p


class OrcCamera(C.Structure):
    _fields_ = [("center", C.c_float * 3), ("U", C.c_float * 3), ("V", C.c_float * 3), ("W", C.c_float * 3),
                ("pv", C.c_float * 16)]


class OrcParams(C.Structure):
    _fields_ = [("width", C.c_int), ("height", C.c_int), ("max_depth", C.c_int), ("samples", C.c_int),
                ("accumulate", C.c_int), ("user_seed", C.c_uint32), ("candidates", C.c_int), ("m_cap", C.c_int),
                ("spatial_passes", C.c_int), ("neighbours", C.c_int), ("radius", C.c_int),
                ("fixed_offset", C.c_int)]


_SIGS = {
    "orc_scene_create": (vp, [fp, fp, C.c_int, u32p, u32p, C.c_int, fp, C.c_int, fp, C.c_int]),
    "orc_scene_destroy": (None, [vp]),
    "orc_scene_set_envmap": (None, [vp, fp, C.c_uint32, C.c_uint32]),
    "orc_scene_set_textures": (None, [vp, fp, fp, u32p, u32p, C.c_int, i32p]),
    "orc_tex_lookup": (None, [u32p, C.c_uint32, C.c_uint32, fp, C.c_int, fp]),
    "orc_env_math": (None, [fp, C.c_int, C.c_int, fp]),
    "orc_env_lookup": (None, [vp, fp, C.c_int, C.c_int, fp]),
    "orc_env_sample_dirs": (None, [C.c_uint32, C.c_int, fp]),
    "orc_scene_light_count": (C.c_int, [vp]),
    "orc_scene_lights": (None, [vp, fp]),
    "orc_pcg3d": (None, [u32p, u32p, C.c_int]),
    "orc_pcg3f": (None, [fp, C.c_int]),
    "orc_sincos_turn": (None, [fp, fp, fp, C.c_int]),
    "orc_tri_boxes": (None, [fp, u32p, C.c_int, fp]),
    "orc_bvh_build": (C.c_int, [fp, C.c_int, fp]),
    "orc_bvh_leaf_order": (C.c_int, [fp, C.c_int, i32p]),
    "orc_bvh_depth": (C.c_int, [fp, C.c_int]),
    "orc_trace": (None, [vp, fp, C.c_int, C.c_int, fp, u64p]),
    "orc_occluded": (None, [vp, fp, C.c_int, i32p]),
    "orc_ggx_pdf": (C.c_float, [fp, fp, fp, fp, fp, C.c_float]),
    "orc_ggx_brdf": (None, [fp, fp, fp, fp, fp, C.c_float, fp]),
    "orc_ggx_sample": (None, [fp, fp, fp, fp, C.c_float, fp, fp]),
    "orc_light_sample": (None, [fp, fp, fp]),
    "orc_ggx_sample_batch": (None, [fp, fp, fp, fp, C.c_float, C.c_int, C.c_uint32, fp]),
    "orc_glass_sample": (None, [fp, fp, C.c_float, C.c_int, fp]),
    "orc_render_pt": (None, [vp, C.POINTER(OrcCamera), C.POINTER(OrcParams), fp, fp, fp, fp, C.c_int, u64p]),
    "orc_render_pt_pixels": (None, [vp, C.POINTER(OrcCamera), C.POINTER(OrcParams), i32p, C.c_int, fp]),
    "orc_accumulate_pt_f64": (None, [vp, C.POINTER(OrcCamera), C.POINTER(OrcParams), C.c_int, C.c_int, dp,
                                     C.c_int]),
    "orc_restir_create": (vp, [C.c_int, C.c_int, C.c_uint32]),
    "orc_restir_destroy": (None, [vp]),
    "orc_restir_frame": (None, [vp, vp, C.POINTER(OrcCamera), C.POINTER(OrcParams), fp, fp, fp, fp, C.c_int,
                                u64p]),
    "orc_restir_reservoirs": (None, [vp, C.c_int, fp]),
    "orc_gi_create": (vp, [C.c_int, C.c_int, C.c_uint32]),
    "orc_gi_destroy": (None, [vp]),
    "orc_gi_frame": (None, [vp, vp, C.POINTER(OrcCamera), C.POINTER(OrcParams), fp, C.c_int, u64p]),
    "orc_gi_reservoirs": (None, [vp, C.c_int, fp]),
    "orc_gi_jacobian": (C.c_float, [fp, fp, fp, fp]),
    "orc_camera_rays": (None, [C.POINTER(OrcCamera), C.POINTER(OrcParams), C.c_int, fp]),
    "orc_mlp_param_count": (C.c_int, []),
    "orc_mlp_encode": (None, [fp, C.c_int, fp, fp, fp]),
    "orc_mlp_forward": (None, [dp, fp, C.c_int, fp, fp, C.c_int, dp]),
    "orc_mlp_loss": (C.c_double, [dp, fp, fp, C.c_int, fp, fp, dp]),
    "orc_mlp_grad": (C.c_double, [dp, fp, fp, C.c_int, fp, fp, C.c_int, C.c_float, dp]),
    "orc_mlp_adam": (None, [fp, fp, fp, fp, C.c_int, C.c_float, C.c_float, C.c_float, C.c_float, fp, fp]),
    "orc_hdr_to_ldr": (None, [fp, u32p, C.c_int, C.c_int, C.c_int]),
    "orc_mamba_create": (vp, [vp, C.c_int, C.c_int, C.c_uint32]),
    "orc_mamba_destroy": (None, [vp]),
    "orc_mamba_light_count": (C.c_int, [vp]),
    "orc_mamba_frame": (None, [vp, vp, C.POINTER(OrcCamera), C.c_int, C.c_int, C.c_int, fp, C.c_int]),
    "orc_mamba_reservoirs": (None, [vp, C.c_int, fp]),
    "orc_scene_create_instanced": (vp, [C.c_int, i32p, i32p, fp, fp, u32p, C.c_int, i32p, fp, u32p, fp, C.c_int]),
    "orc_scene_tlas": (C.c_int, [vp, fp]),
    "orc_affine_inverse": (None, [fp, fp]),
    "orc_instance_box": (None, [fp, fp, fp, fp]),
}

_lib = None


def build() -> Path:
    """Compile liboracle.so with the committed Makefile."""
    r = subprocess.run(["make", "-s", "-C", str(HERE)], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"oracle build failed:\n{r.stdout}\n{r.stderr}")
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        L = C.CDLL(str(LIB_PATH))
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _f(a):
    return a.ctypes.data_as(fp)


def _u(a):
    return a.ctypes.data_as(u32p)


def pcg3d(words: np.ndarray) -> np.ndarray:
    w = np.ascontiguousarray(words, np.uint32).reshape(-1, 3)
    out = np.empty_like(w)
    lib().orc_pcg3d(_u(w), _u(out), len(w))
    return out


def pcg3f(seeds: np.ndarray) -> np.ndarray:
    s = np.ascontiguousarray(seeds, np.float32).reshape(-1, 3).copy()
    lib().orc_pcg3f(_f(s), len(s))
    return s


def sincos_turn(u: np.ndarray):
    u = np.ascontiguousarray(u, np.float32)
    c = np.empty_like(u)
    s = np.empty_like(u)
    lib().orc_sincos_turn(_f(u), _f(c), _f(s), len(u))
    return c, s


def hdr_to_ldr(pixels: np.ndarray, tonemapping: int = 0) -> np.ndarray:
    """(H, W, 4) float32 -> (H, W) uint32 RGBA8, rows flipped (color.cuh:82-121)."""
    p = np.ascontiguousarray(pixels, np.float32)
    H, W = p.shape[:2]
    out = np.empty((H, W), np.uint32)
    lib().orc_hdr_to_ldr(_f(p), _u(out), W, H, tonemapping)
    return out


def affine_inverse(M: np.ndarray) -> np.ndarray:
    m = np.ascontiguousarray(M, np.float32).reshape(12)
    out = np.empty(12, np.float32)
    lib().orc_affine_inverse(_f(m), _f(out))
    return out.reshape(3, 4)


def tri_boxes(pos: np.ndarray, idx: np.ndarray) -> np.ndarray:
    pos = np.ascontiguousarray(pos, np.float32)
    idx = np.ascontiguousarray(idx, np.uint32)
    out = np.empty((len(idx), 6), np.float32)
    lib().orc_tri_boxes(_f(pos), _u(idx), len(idx), _f(out))
    return out


def bvh_build(boxes: np.ndarray) -> np.ndarray:
    boxes = np.ascontiguousarray(boxes, np.float32).reshape(-1, 6)
    n = len(boxes)
    out = np.zeros((max(2 * n - 1, 1), 12), np.float32)
    k = lib().orc_bvh_build(_f(boxes), n, _f(out))
    return out[:k]


def bvh_leaf_order(nodes: np.ndarray) -> np.ndarray:
    nodes = np.ascontiguousarray(nodes, np.float32)
    out = np.empty(len(nodes), np.int32)
    k = lib().orc_bvh_leaf_order(_f(nodes), len(nodes), out.ctypes.data_as(i32p))
    return out[:k]


def bvh_depth(nodes: np.ndarray) -> int:
    nodes = np.ascontiguousarray(nodes, np.float32)
    return lib().orc_bvh_depth(_f(nodes), len(nodes))


def camera_struct(frame) -> OrcCamera:
    c = OrcCamera()
    center, U, V, W, pv = frame.as_arrays()
    c.center[:] = [float(x) for x in center]
    c.U[:] = [float(x) for x in U]
    c.V[:] = [float(x) for x in V]
    c.W[:] = [float(x) for x in W]
    c.pv[:] = [float(x) for x in pv]
    return c


def params(width, height, max_depth=5, samples=0, accumulate=1, user_seed=0, candidates=32, m_cap=200,
           spatial_passes=2, neighbours=5, radius=30, fixed_offset=0) -> OrcParams:
    return OrcParams(width, height, max_depth, samples, accumulate, user_seed, candidates, m_cap, spatial_passes,
                     neighbours, radius, fixed_offset)


def pack_textures(textures):
    """RGBA8 images ((H, W, 4) uint8 each) as one uint32 texel array (r in
    the low byte, row-major, back to back) and their (width, height) pairs."""
    parts, dims = [], []
    for t in textures:
        a = np.ascontiguousarray(t, np.uint8)
        assert a.ndim == 3 and a.shape[2] == 4, "textures are (H, W, 4) uint8"
        parts.append(a.reshape(-1, 4).view("<u4").reshape(-1))
        dims.append((a.shape[1], a.shape[0]))
    texels = np.ascontiguousarray(np.concatenate(parts) if parts else np.zeros(1, np.uint32), np.uint32)
    return texels, np.ascontiguousarray(np.array(dims, np.uint32).reshape(-1, 2))


def tex_lookup(texture, uv) -> np.ndarray:
    """The texture fetch of one (H, W, 4) uint8 image at (n, 2) coordinates."""
    texels, dims = pack_textures([texture])
    uv = np.ascontiguousarray(uv, np.float32).reshape(-1, 2)
    out = np.empty((len(uv), 4), np.float32)
    lib().orc_tex_lookup(_u(texels), int(dims[0, 0]), int(dims[0, 1]), _f(uv), len(uv), _f(out))
    return out


class OracleScene:
    """Oracle view of a scene (positions, normals, indices, materials, BVH)."""

    def __init__(self, data, nodes: np.ndarray | None = None):
        self.data = data
        if hasattr(data, "inst_mesh"):  # InstancedSceneData: two-level scene
            self._init_instanced(data)
            return
        self._keep = [np.ascontiguousarray(data.positions, np.float32),
                      np.ascontiguousarray(data.normals, np.float32),
                      np.ascontiguousarray(data.indices, np.uint32),
                      np.ascontiguousarray(data.tri_material, np.uint32),
                      np.ascontiguousarray(data.materials, np.float32)]
        if nodes is None:
            nodes = bvh_build(tri_boxes(self._keep[0], self._keep[2]))
        self.nodes = np.ascontiguousarray(nodes, np.float32)
        p, n, i, m, mats = self._keep
        self.h = lib().orc_scene_create(_f(p), _f(n), len(p), _u(i), _u(m), len(i), _f(mats), len(mats),
                                        _f(self.nodes), len(self.nodes))
        if getattr(data, "textures", None) is not None:
            self._set_textures(data)

    def _set_textures(self, d):
        """Material textures (SceneData.uvs / tangents / textures /
        material_textures): calculate_material / calculate_normal."""
        uv = np.ascontiguousarray(d.uvs, np.float32)
        tan = None if d.tangents is None else np.ascontiguousarray(d.tangents, np.float32)
        texels, dims = pack_textures(d.textures)
        mt = np.ascontiguousarray(d.material_textures, np.int32)
        self._keep += [uv, tan, texels, dims, mt]
        lib().orc_scene_set_textures(self.h, _f(uv), None if tan is None else _f(tan), _u(texels), _u(dims),
                                     len(dims), mt.ctypes.data_as(i32p))

    def _init_instanced(self, d):
        pos = np.ascontiguousarray(np.concatenate([m[0] for m in d.meshes]), np.float32)
        nrm = np.ascontiguousarray(np.concatenate([m[1] for m in d.meshes]), np.float32)
        idx = np.ascontiguousarray(np.concatenate([m[2] for m in d.meshes]), np.uint32)
        nv = np.array([len(m[0]) for m in d.meshes], np.int32)
        nt = np.array([len(m[2]) for m in d.meshes], np.int32)
        im = np.ascontiguousarray(d.inst_mesh, np.int32)
        xf = np.ascontiguousarray(d.transforms, np.float32).reshape(-1, 12)
        tm = np.ascontiguousarray(d.tri_material, np.uint32)
        mats = np.ascontiguousarray(d.materials, np.float32)
        self._keep = [pos, nrm, idx, nv, nt, im, xf, tm, mats]
        i32 = lambda a: a.ctypes.data_as(i32p)
        self.h = lib().orc_scene_create_instanced(len(nv), i32(nv), i32(nt), _f(pos), _f(nrm), _u(idx), len(im),
                                                  i32(im), _f(xf), _u(tm), _f(mats), len(mats))
        self.nodes = None

    def tlas(self) -> np.ndarray:
        """The bvh.cpp tree over the instances (12 floats per node)."""
        n = lib().orc_scene_tlas(self.h, None)
        out = np.empty((n, 12), np.float32)
        lib().orc_scene_tlas(self.h, _f(out))
        return out

    def __del__(self):
        try:
            if self.h:
                lib().orc_scene_destroy(self.h)
        except Exception:
            pass

    def set_envmap(self, rgba) -> None:
        """The armada's environment map: (H, W, 4) float32, or None."""
        if rgba is None:
            lib().orc_scene_set_envmap(self.h, None, 0, 0)
            return
        a = np.ascontiguousarray(rgba, np.float32)
        lib().orc_scene_set_envmap(self.h, _f(a), a.shape[1], a.shape[0])

    def env_lookup(self, dirs, flip_v=False) -> np.ndarray:
        d = np.ascontiguousarray(dirs, np.float32).reshape(-1, 3)
        out = np.empty((len(d), 3), np.float32)
        lib().orc_env_lookup(self.h, _f(d), len(d), int(flip_v), _f(out))
        return out

    def lights(self) -> np.ndarray:
        n = lib().orc_scene_light_count(self.h)
        out = np.empty((n, 12), np.float32)
        if n:
            lib().orc_scene_lights(self.h, _f(out))
        return out

    def trace(self, rays: np.ndarray, mode: int = 0):
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        hits = np.empty((len(rays), 4), np.float32)
        ctr = np.zeros(2, np.uint64)
        lib().orc_trace(self.h, _f(rays), len(rays), mode, _f(hits), ctr.ctypes.data_as(u64p))
        return hits, ctr

    def occluded(self, rays: np.ndarray) -> np.ndarray:
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        out = np.empty(len(rays), np.int32)
        lib().orc_occluded(self.h, _f(rays), len(rays), out.ctypes.data_as(i32p))
        return out

    def render_pt(self, frame, width, height, frames=1, max_depth=5, user_seed=0, nthreads=8, buffers=None,
                  first_sample=0):
        """Accumulate `frames` frames like ArmadaRTX (samples counter from first_sample)."""
        if buffers is None:
            buffers = [np.zeros((height, width, 4), np.float32) for _ in range(4)]
        cam = camera_struct(frame)
        ctr = np.zeros(5, np.uint64)
        for f in range(frames):
            p = params(width, height, max_depth=max_depth, samples=first_sample + f, user_seed=user_seed)
            lib().orc_render_pt(self.h, C.byref(cam), C.byref(p), *[_f(b) for b in buffers], nthreads,
                                ctr.ctypes.data_as(u64p))
        return buffers, ctr

    def render_pt_pixels(self, frame, width, height, pixels, sample=0, max_depth=5, user_seed=0):
        pixels = np.ascontiguousarray(pixels, np.int32)
        out = np.empty((len(pixels), 4), np.float32)
        cam = camera_struct(frame)
        p = params(width, height, max_depth=max_depth, samples=sample, user_seed=user_seed)
        lib().orc_render_pt_pixels(self.h, C.byref(cam), C.byref(p), pixels.ctypes.data_as(i32p), len(pixels),
                                   _f(out))
        return out

    def accumulate_pt_f64(self, frame, width, height, first, count, max_depth=5, user_seed=0, nthreads=8, acc=None,
                          fixed_offset=False):
        if acc is None:
            acc = np.zeros((height, width, 3), np.float64)
        cam = camera_struct(frame)
        p = params(width, height, max_depth=max_depth, user_seed=user_seed, fixed_offset=int(fixed_offset))
        lib().orc_accumulate_pt_f64(self.h, C.byref(cam), C.byref(p), first, count, acc.ctypes.data_as(dp),
                                    nthreads)
        return acc


class OracleReSTIR:
    def __init__(self, width, height, user_seed=0):
        self.w, self.h_ = width, height
        self.h = lib().orc_restir_create(width, height, user_seed)
        self.user_seed = user_seed
        self.buffers = [np.zeros((height, width, 4), np.float32) for _ in range(4)]
        self.samples = 0

    def __del__(self):
        try:
            if self.h:
                lib().orc_restir_destroy(self.h)
        except Exception:
            pass

    def frame(self, scene: OracleScene, frame, max_depth=2, candidates=32, m_cap=200, spatial_passes=2,
              neighbours=5, radius=30, nthreads=8, reset=False):
        if reset:
            self.samples = 0
        cam = camera_struct(frame)
        p = params(self.w, self.h_, max_depth=max_depth, samples=self.samples, user_seed=self.user_seed,
                   candidates=candidates, m_cap=m_cap, spatial_passes=spatial_passes, neighbours=neighbours,
                   radius=radius)
        ctr = np.zeros(5, np.uint64)
        lib().orc_restir_frame(self.h, scene.h, C.byref(cam), C.byref(p), *[_f(b) for b in self.buffers], nthreads,
                               ctr.ctypes.data_as(u64p))
        self.samples += 1
        return ctr

    def reservoirs(self, which=0) -> np.ndarray:
        out = np.empty((self.h_, self.w, 16), np.float32)
        lib().orc_restir_reservoirs(self.h, which, _f(out))
        return out


class OracleGI:
    """ReSTIR GI restatement (our algorithm; oracle.cpp)."""

    def __init__(self, width, height, user_seed=0):
        self.w, self.h_ = width, height
        self.h = lib().orc_gi_create(width, height, user_seed)
        self.user_seed = user_seed
        self.color = np.zeros((height, width, 4), np.float32)
        self.samples = 0

    def __del__(self):
        try:
            if self.h:
                lib().orc_gi_destroy(self.h)
        except Exception:
            pass

    def frame(self, scene: OracleScene, frame, max_depth=3, m_cap=30, spatial_passes=2, neighbours=5, radius=30,
              nthreads=8, accumulate=True):
        if not accumulate:
            self.samples = 0
        cam = camera_struct(frame)
        p = params(self.w, self.h_, max_depth=max_depth, samples=self.samples, user_seed=self.user_seed,
                   m_cap=m_cap, spatial_passes=spatial_passes, neighbours=neighbours, radius=radius)
        ctr = np.zeros(5, np.uint64)
        lib().orc_gi_frame(self.h, scene.h, C.byref(cam), C.byref(p), _f(self.color), nthreads,
                           ctr.ctypes.data_as(u64p))
        self.samples += 1
        return ctr

    def reservoirs(self, which=0) -> np.ndarray:
        out = np.empty((self.h_, self.w, 20), np.float32)
        lib().orc_gi_reservoirs(self.h, which, _f(out))
        return out


class OracleMamba:
    """The editor's G-buffer ReSTIR DI (editor/optix/mamba_shader.cu), our restatement."""

    def __init__(self, scene: OracleScene, width, height, user_seed=0):
        self.scene = scene
        self.w, self.h_ = width, height
        self.h = lib().orc_mamba_create(scene.h, width, height, user_seed)
        self.color = np.zeros((height, width, 4), np.float32)

    def __del__(self):
        try:
            if self.h:
                lib().orc_mamba_destroy(self.h)
        except Exception:
            pass

    def light_count(self) -> int:
        return int(lib().orc_mamba_light_count(self.h))

    def frame(self, frame, temporal=True, spatial=True, accumulate=True, nthreads=8):
        cam = camera_struct(frame)
        lib().orc_mamba_frame(self.h, self.scene.h, C.byref(cam), int(temporal), int(spatial), int(accumulate),
                              _f(self.color), nthreads)
        return self.color

    def reservoirs(self, which=0) -> np.ndarray:
        out = np.empty((self.h_, self.w, 16), np.float32)
        lib().orc_mamba_reservoirs(self.h, which, _f(out))
        return out


def env_atan2(y, x) -> np.ndarray:
    a = np.ascontiguousarray(np.stack([np.asarray(y, np.float32), np.asarray(x, np.float32)], -1), np.float32)
    out = np.empty(len(a), np.float32)
    lib().orc_env_math(_f(a), len(a), 0, _f(out))
    return out


def env_asin(x) -> np.ndarray:
    a = np.zeros((len(x), 2), np.float32)
    a[:, 0] = x
    out = np.empty(len(a), np.float32)
    lib().orc_env_math(_f(a), len(a), 1, _f(out))
    return out


def env_sample_dirs(seed_word: int, n: int) -> np.ndarray:
    out = np.empty((n, 3), np.float32)
    lib().orc_env_sample_dirs(seed_word, n, _f(out))
    return out


def gi_jacobian(xq, xr, xs, ns) -> float:
    a = [np.ascontiguousarray(v, np.float32) for v in (xq, xr, xs, ns)]
    return float(lib().orc_gi_jacobian(*[_f(v) for v in a]))


def camera_rays(frame, width, height, jitter=False, sample=0, user_seed=0) -> np.ndarray:
    out = np.empty((height * width, 8), np.float32)
    cam = camera_struct(frame)
    p = params(width, height, samples=sample, user_seed=user_seed)
    lib().orc_camera_rays(C.byref(cam), C.byref(p), int(jitter), _f(out))
    return out


# ---------------------------------------------------------------------------
# Radiance-cache MLP (mlp.cpp)
# ---------------------------------------------------------------------------
MLP_RAW = 13
MLP_FEATURES = 64


def _d(a):
    return a.ctypes.data_as(dp)


def _mlp_args(raw, pmin, pscale):
    raw = np.ascontiguousarray(raw, np.float32).reshape(-1, MLP_RAW)
    pmin = np.ascontiguousarray(pmin, np.float32).reshape(3)
    pscale = np.ascontiguousarray(pscale, np.float32).reshape(3)
    return raw, pmin, pscale


def mlp_param_count() -> int:
    return int(lib().orc_mlp_param_count())


def mlp_encode(raw, pmin=(0, 0, 0), pscale=(1, 1, 1)) -> np.ndarray:
    raw, pmin, pscale = _mlp_args(raw, pmin, pscale)
    out = np.empty((len(raw), MLP_FEATURES), np.float32)
    lib().orc_mlp_encode(_f(raw), len(raw), _f(pmin), _f(pscale), _f(out))
    return out


def mlp_forward(params, raw, pmin=(0, 0, 0), pscale=(1, 1, 1), bf16=False) -> np.ndarray:
    params = np.ascontiguousarray(params, np.float64)
    raw, pmin, pscale = _mlp_args(raw, pmin, pscale)
    out = np.empty((len(raw), 3), np.float64)
    lib().orc_mlp_forward(_d(params), _f(raw), len(raw), _f(pmin), _f(pscale), int(bf16), _d(out))
    return out


def mlp_loss(params, raw, target, pmin=(0, 0, 0), pscale=(1, 1, 1), den=None) -> float:
    params = np.ascontiguousarray(params, np.float64)
    raw, pmin, pscale = _mlp_args(raw, pmin, pscale)
    target = np.ascontiguousarray(target, np.float32).reshape(-1, 3)
    dptr = None
    if den is not None:
        den = np.ascontiguousarray(den, np.float64).reshape(-1, 3)
        dptr = _d(den)
    return float(lib().orc_mlp_loss(_d(params), _f(raw), _f(target), len(raw), _f(pmin), _f(pscale), dptr))


def mlp_grad(params, raw, target, pmin=(0, 0, 0), pscale=(1, 1, 1), bf16=False, loss_scale=1.0):
    """(loss, gradient) of the mean relative L2 loss, gradient in canonical order."""
    params = np.ascontiguousarray(params, np.float64)
    raw, pmin, pscale = _mlp_args(raw, pmin, pscale)
    target = np.ascontiguousarray(target, np.float32).reshape(-1, 3)
    g = np.empty(mlp_param_count(), np.float64)
    loss = lib().orc_mlp_grad(_d(params), _f(raw), _f(target), len(raw), _f(pmin), _f(pscale), int(bf16),
                              float(loss_scale), _d(g))
    return float(loss), g


def mlp_adam(params, m, v, grad, lr, b1, b2, eps, b1t, b2t):
    """One FP32 Adam step in place; returns the advanced (b1t, b2t)."""
    for a in (params, m, v):
        assert a.dtype == np.float32 and a.flags.c_contiguous
    grad = np.ascontiguousarray(grad, np.float32)
    t1 = np.array([b1t], np.float32)
    t2 = np.array([b2t], np.float32)
    lib().orc_mlp_adam(_f(params), _f(m), _f(v), _f(grad), len(params), lr, b1, b2, eps, _f(t1), _f(t2))
    return float(t1[0]), float(t2[0])
