"""Seeded procedural scenes (world-space triangle soups).

* cornell(): room 5 faces x 16x16 quads, two boxes x 10x10 quads per face, a
  level-4 icosphere and one ceiling quad light (2 triangles, emission 15):
  10,082 triangles.  Per object albedo U[0.2,0.8]^3, GGX roughness
  U[0.05,1], specular 0 (include/material.hpp:22), type eDiffuse.
* spheres(): `instances` copies of 16 seeded base meshes built like
  Submesh::sphere(slices, stacks) (source/mesh.cpp:189-290, 2*slices*(stacks-1)
  triangles) with radial noise, random rotation/scale, positions uniform in a
  `box` m cube; `emissive` triangles chosen uniformly get their own material
  with log-uniform emission in [0.1, 10].
  sphere(64, 33) x 256 = 1,048,576 triangles; sphere(64, 65) x 256 = 2,097,152.

Random numbers come from SplitMix64 so the scenes are identical on every host.
Materials are rows of 12 floats {diffuse, specular, emission, roughness,
refraction, type}.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

E_DIFFUSE = 4.0
_M64 = (1 << 64) - 1


class SplitMix64:
    def __init__(self, seed: int):
        self.s = seed & _M64

    def next(self) -> int:
        self.s = (self.s + 0x9E3779B97F4A7C15) & _M64
        z = self.s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _M64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _M64
        return z ^ (z >> 31)

    def uniform(self, lo=0.0, hi=1.0, n=None):
        if n is None:
            return lo + (hi - lo) * ((self.next() >> 11) * (1.0 / (1 << 53)))
        return np.array([self.uniform(lo, hi) for _ in range(n)])

    def uniform_array(self, n: int) -> np.ndarray:
        """n uniforms in [0,1): vectorised SplitMix64 stream seeded from this one."""
        base = np.uint64(self.next())
        with np.errstate(over="ignore"):
            k = np.arange(1, n + 1, dtype=np.uint64)
            z = base + k * np.uint64(0x9E3779B97F4A7C15)
            z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
            z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
            z = z ^ (z >> np.uint64(31))
        return (z >> np.uint64(11)).astype(np.float64) * (1.0 / (1 << 53))


@dataclass
class SceneData:
    positions: np.ndarray     # (nv, 3) float32
    normals: np.ndarray       # (nv, 3) float32
    indices: np.ndarray       # (nt, 3) uint32
    tri_material: np.ndarray  # (nt,) uint32
    materials: np.ndarray     # (nm, 12) float32
    camera: dict = field(default_factory=dict)
    name: str = ""
    # material textures (krt_scene_upload_textured): Vertex::tex_coords,
    # {tangent, bitangent}, RGBA8 images and per material the texture of
    # {diffuse, normal, specular, emission, roughness} (-1: none)
    uvs: np.ndarray | None = None                # (nv, 2) float32
    tangents: np.ndarray | None = None           # (nv, 6) float32
    textures: list | None = None                 # [(H, W, 4) uint8]
    material_textures: np.ndarray | None = None  # (nm, 5) int32

    @property
    def triangles(self) -> int:
        return len(self.indices)


def pack_textures(textures):
    """RGBA8 images ((H, W, 4) uint8 each) as one uint32 texel array (r in
    the low byte, row-major, back to back) and their (width, height) pairs."""
    parts, dims = [], []
    for t in textures:
        a = np.ascontiguousarray(t, np.uint8)
        if a.ndim != 3 or a.shape[2] != 4:
            raise ValueError("textures are (H, W, 4) uint8 images")
        parts.append(a.reshape(-1, 4).view("<u4").reshape(-1))
        dims.append((a.shape[1], a.shape[0]))
    texels = np.ascontiguousarray(np.concatenate(parts) if parts else np.zeros(1, np.uint32), np.uint32)
    return texels, np.ascontiguousarray(np.array(dims, np.uint32).reshape(-1, 2))


def xform_points(M: np.ndarray, p: np.ndarray, w: float = 1.0) -> np.ndarray:
    """Row-major 3x4 transform of float32 points (w = 1) or directions (w = 0)
    with glm's pairing (c0 x + c1 y) + (c2 z + c3 w), in float32."""
    M = np.asarray(M, np.float32)
    p = np.asarray(p, np.float32)
    w = np.float32(w)
    out = np.empty_like(p)
    for r in range(3):
        out[:, r] = (M[r, 0] * p[:, 0] + M[r, 1] * p[:, 1]) + (M[r, 2] * p[:, 2] + M[r, 3] * w)
    return out


@dataclass
class InstancedSceneData:
    """Base meshes in their own space and instances {mesh, row-major 3x4
    transform} (Accelerator::update / build_tlas,
    include/amadeus/accelerator.cuh:145-277).  Triangle ids are global:
    instance i's triangles follow those of instances 0..i-1; tri_material
    holds the material of every global triangle."""
    meshes: list              # [(positions (nv, 3) f32, normals (nv, 3) f32, indices (nt, 3) u32)]
    inst_mesh: np.ndarray     # (ni,) int32
    transforms: np.ndarray    # (ni, 3, 4) float32
    tri_material: np.ndarray  # (sum of the instances' triangles,) uint32
    materials: np.ndarray     # (nm, 12) float32
    camera: dict = field(default_factory=dict)
    name: str = ""
    # material textures (as SceneData's): per mesh (nv, 2) tex_coords and
    # (nv, 6) {tangent, bitangent} in its own space, RGBA8 images, 5 texture
    # slots per material
    mesh_uvs: list | None = None
    mesh_tangents: list | None = None
    textures: list | None = None
    material_textures: np.ndarray | None = None

    @property
    def triangles(self) -> int:
        return int(sum(len(self.meshes[m][2]) for m in self.inst_mesh))

    def flatten(self) -> SceneData:
        """The same scene as one world-space soup (vertices through the
        transforms in float32, normals through the 3x3 part).  Tangent
        frames go through the 3x3 part twice: calculate_normal applies the
        model matrix to the interpolated tangent and bitangent and then again
        (amadeus_common.cuh:125-136), which a world-space soup reproduces
        with the product taken before the interpolation."""
        P, N, I, UV, T = [], [], [], [], []
        nv = 0
        for i, m in enumerate(self.inst_mesh):
            pos, nrm, idx = self.meshes[m]
            M = self.transforms[i]
            P.append(xform_points(M, pos, 1.0))
            N.append(xform_points(M, nrm, 0.0))
            I.append(idx + nv)
            if self.textures is not None:
                UV.append(np.asarray(self.mesh_uvs[m], np.float32))
                t = np.asarray(self.mesh_tangents[m], np.float32)
                T.append(np.concatenate([xform_points(M, xform_points(M, t[:, 0:3], 0.0), 0.0),
                                         xform_points(M, xform_points(M, t[:, 3:6], 0.0), 0.0)], axis=1))
            nv += len(pos)
        d = SceneData(np.concatenate(P), np.concatenate(N), np.concatenate(I).astype(np.uint32),
                      self.tri_material, self.materials, camera=self.camera, name=self.name + "_flat")
        if self.textures is not None:
            d.uvs, d.tangents = np.concatenate(UV), np.concatenate(T)
            d.textures, d.material_textures = list(self.textures), np.asarray(self.material_textures, np.int32)
        return d


def material(diffuse, roughness, emission=(0.0, 0.0, 0.0), specular=(0.0, 0.0, 0.0), refraction=1.0,
             type_=E_DIFFUSE):
    return [*diffuse, *specular, *emission, roughness, refraction, type_]


class _Builder:
    def __init__(self):
        self.p, self.n, self.i, self.m, self.uv, self.tan = [], [], [], [], [], []
        self.nv = 0

    def add(self, pos, nrm, idx, mat_id, uv=None, tan=None):
        pos = np.asarray(pos, np.float64)
        nrm = np.asarray(nrm, np.float64)
        idx = np.asarray(idx, np.int64).reshape(-1, 3)
        self.p.append(pos)
        self.n.append(nrm)
        self.i.append(idx + self.nv)
        self.m.append(np.full(len(idx), mat_id, np.uint32) if np.isscalar(mat_id) else np.asarray(mat_id, np.uint32))
        self.uv.append(np.zeros((len(pos), 2)) if uv is None else np.asarray(uv, np.float64))
        self.tan.append(np.zeros((len(pos), 6)) if tan is None else np.asarray(tan, np.float64))
        self.nv += len(pos)

    def build(self, mats, textures=None, material_textures=None, **kw) -> SceneData:
        d = SceneData(np.concatenate(self.p).astype(np.float32), np.concatenate(self.n).astype(np.float32),
                      np.concatenate(self.i).astype(np.uint32), np.concatenate(self.m),
                      np.asarray(mats, np.float32), **kw)
        if textures is not None:
            d.uvs = np.concatenate(self.uv).astype(np.float32)
            d.tangents = np.concatenate(self.tan).astype(np.float32)
            d.textures = list(textures)
            d.material_textures = np.asarray(material_textures, np.int32).reshape(-1, 5)
        return d


def grid_quad(origin, du, dv, n: int, normal):
    """n x n subdivided parallelogram origin + s*du + t*dv (2 n^2 triangles)."""
    origin, du, dv = (np.asarray(x, np.float64) for x in (origin, du, dv))
    s = np.linspace(0.0, 1.0, n + 1)
    S, T = np.meshgrid(s, s, indexing="ij")
    pos = origin + S.reshape(-1, 1) * du + T.reshape(-1, 1) * dv
    nrm = np.tile(np.asarray(normal, np.float64), (len(pos), 1))
    idx = []
    for a in range(n):
        for b in range(n):
            v00 = a * (n + 1) + b
            v10 = (a + 1) * (n + 1) + b
            v01 = a * (n + 1) + b + 1
            v11 = (a + 1) * (n + 1) + b + 1
            idx += [(v00, v10, v11), (v00, v11, v01)]
    return pos, nrm, np.array(idx)


def grid_frame(du, dv, n: int, scale=(1.0, 1.0), offset=(0.0, 0.0)):
    """Texture coordinates of grid_quad's vertices ((s, t) * scale + offset)
    and its tangent frame {du / |du|, dv / |dv|}."""
    s = np.linspace(0.0, 1.0, n + 1)
    S, T = np.meshgrid(s, s, indexing="ij")
    uv = np.stack([S.reshape(-1) * scale[0] + offset[0], T.reshape(-1) * scale[1] + offset[1]], axis=1)
    du, dv = np.asarray(du, np.float64), np.asarray(dv, np.float64)
    tan = np.concatenate([du / np.linalg.norm(du), dv / np.linalg.norm(dv)])
    return uv, np.tile(tan, (len(uv), 1))


def box_mesh(center, half, n: int):
    c, h = np.asarray(center, np.float64), np.asarray(half, np.float64)
    faces = []
    for axis in range(3):
        for sgn in (-1.0, 1.0):
            a1, a2 = (axis + 1) % 3, (axis + 2) % 3
            o = c.copy()
            o[axis] += sgn * h[axis]
            o[a1] -= h[a1]
            o[a2] -= h[a2]
            du = np.zeros(3)
            du[a1] = 2 * h[a1]
            dv = np.zeros(3)
            dv[a2] = 2 * h[a2]
            nrm = np.zeros(3)
            nrm[axis] = sgn
            faces.append(grid_quad(o, du, dv, n, nrm))
    return faces


def icosphere(level: int):
    t = (1.0 + 5.0 ** 0.5) / 2.0
    verts = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
             (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    verts = [np.array(v, np.float64) / np.linalg.norm(v) for v in verts]
    faces = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
             (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11),
             (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    for _ in range(level):
        cache = {}

        def mid(a, b):
            key = (min(a, b), max(a, b))
            if key not in cache:
                m = verts[a] + verts[b]
                verts.append(m / np.linalg.norm(m))
                cache[key] = len(verts) - 1
            return cache[key]

        nf = []
        for a, b, c in faces:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = nf
    return np.array(verts), np.array(faces)


def cornell(seed: int = 1, light_half: float = 0.25, light_emission: float = 15.0) -> SceneData:
    rng = SplitMix64(seed)

    def rand_mat():
        albedo = [rng.uniform(0.2, 0.8) for _ in range(3)]
        return material(albedo, rng.uniform(0.05, 1.0))

    mats, b = [], _Builder()
    # room [-1,1] x [0,2] x [-1,1]: floor, ceiling, back, left, right (inward normals)
    walls = [((-1, 0, -1), (2, 0, 0), (0, 0, 2), (0, 1, 0)),
             ((-1, 2, -1), (0, 0, 2), (2, 0, 0), (0, -1, 0)),
             ((-1, 0, -1), (0, 2, 0), (2, 0, 0), (0, 0, 1)),
             ((-1, 0, -1), (0, 0, 2), (0, 2, 0), (1, 0, 0)),
             ((1, 0, -1), (0, 2, 0), (0, 0, 2), (-1, 0, 0))]
    for o, du, dv, nrm in walls:
        mats.append(rand_mat())
        b.add(*grid_quad(o, du, dv, 16, nrm), len(mats) - 1)
    for center, half in (((-0.4, 0.6, -0.35), (0.3, 0.6, 0.3)), ((0.45, 0.3, 0.3), (0.3, 0.3, 0.3))):
        mats.append(rand_mat())
        for face in box_mesh(center, half, 10):
            b.add(*face, len(mats) - 1)
    mats.append(rand_mat())
    v, f = icosphere(4)
    r, c = 0.28, np.array([0.35, 0.88, 0.25])
    b.add(c + r * v, v, f, len(mats) - 1)
    mats.append(material((0.8, 0.8, 0.8), 1.0, emission=(light_emission,) * 3))
    h = light_half
    b.add(np.array([(-h, 1.99, -h), (h, 1.99, -h), (h, 1.99, h), (-h, 1.99, h)]),
          np.tile((0.0, -1.0, 0.0), (4, 1)), np.array([(0, 2, 1), (0, 3, 2)]), len(mats) - 1)
    cam = dict(eye=(0.0, 1.0, 3.5), lookat=(0.0, 1.0, 0.0), up=(0.0, 1.0, 0.0), fov=45.0)
    return b.build(mats, camera=cam, name="cornell")


def cornell_glossy(seed: int = 5) -> SceneData:
    """The Cornell box with every branch of the BSDF (brdf.cuh) in use: walls
    and the tall box with specular U[0.05, 0.9]^3 (the GGX lobe is sampled
    with probability Ks / (Kd + Ks)), the short box a near-mirror (Kd 0.05,
    Ks 0.95, roughness 0.05), the icosphere glass (type eTransmission,
    refraction 1.5, specular 0.9: Glass::sample, brdf.cuh:125-176)."""
    base = cornell(seed=seed)
    rng = SplitMix64(seed + 7)
    mats = np.array(base.materials, np.float32)
    for m in range(len(mats) - 1):  # the last material is the light
        mats[m, 3:6] = [rng.uniform(0.05, 0.9) for _ in range(3)]
    short_box, sphere = 6, 7
    mats[short_box] = material((0.05, 0.05, 0.05), 0.05, specular=(0.95, 0.95, 0.95))
    mats[sphere] = material((0.0, 0.0, 0.0), 0.1, specular=(0.9, 0.9, 0.9), refraction=1.5, type_=2.0)
    return SceneData(base.positions, base.normals, base.indices, base.tri_material, mats, camera=base.camera,
                     name="cornell_glossy")


def sphere_mesh(slices: int, stacks: int):
    """Submesh::sphere (source/mesh.cpp:189-290): unit sphere, 2*slices*(stacks-1) triangles."""
    pos = [(0.0, 1.0, 0.0)]
    for i in range(stacks - 1):
        phi = np.pi * (i + 1) / stacks
        for j in range(slices):
            th = 2.0 * np.pi * j / slices
            pos.append((np.sin(phi) * np.cos(th), np.cos(phi), np.sin(phi) * np.sin(th)))
    pos.append((0.0, -1.0, 0.0))
    pos = np.array(pos)
    idx = []
    last = len(pos) - 1
    for i in range(slices):
        idx.append((0, (i + 1) % slices + 1, i + 1))
        idx.append((last, i + slices * (stacks - 2) + 1, (i + 1) % slices + slices * (stacks - 2) + 1))
    for i in range(stacks - 2):
        for j in range(slices):
            i0 = i * slices + j + 1
            i1 = i * slices + (j + 1) % slices + 1
            i2 = (i + 1) * slices + (j + 1) % slices + 1
            i3 = (i + 1) * slices + j + 1
            idx += [(i0, i1, i2), (i0, i2, i3)]
    return pos, np.array(idx)


def _rotation(rng: SplitMix64) -> np.ndarray:
    u1, u2, u3 = rng.uniform(), rng.uniform(), rng.uniform()
    q = np.array([np.sqrt(1 - u1) * np.sin(2 * np.pi * u2), np.sqrt(1 - u1) * np.cos(2 * np.pi * u2),
                  np.sqrt(u1) * np.sin(2 * np.pi * u3), np.sqrt(u1) * np.cos(2 * np.pi * u3)])
    x, y, z, w = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def spheres(instances: int = 256, slices: int = 64, stacks: int = 33, emissive: int = 4096, box: float = 20.0,
            seed: int = 2, base_meshes: int = 16, instanced: bool = False):
    """The instanced-sphere scene; instanced=True returns it as base meshes +
    transforms (InstancedSceneData), same random stream, same materials."""
    rng = SplitMix64(seed)
    unit, idx = sphere_mesh(slices, stacks)
    bases = []
    for _ in range(base_meshes):
        noise = 1.0 + 0.15 * (rng.uniform_array(len(unit)) - 0.5)
        bases.append(unit * noise[:, None])
    ntb = len(idx)
    nt = instances * ntb
    mats = []
    P, N, I, M, X = [], [], [], [], []
    for k in range(instances):
        base = bases[k % base_meshes]
        R = _rotation(rng)
        s = rng.uniform(0.4, 1.2)
        t = np.array([rng.uniform(-box / 2, box / 2) for _ in range(3)])
        mats.append(material([rng.uniform(0.2, 0.8) for _ in range(3)], rng.uniform(0.05, 1.0)))
        P.append((base * s) @ R.T + t)
        N.append(unit @ R.T)
        I.append(idx + k * len(unit))
        M.append(np.full(ntb, len(mats) - 1, np.uint32))
        X.append(np.concatenate([s * R, t[:, None]], axis=1))
    tri_mat = np.concatenate(M)
    if emissive > 0:
        picks = np.unique((rng.uniform_array(emissive * 2) * nt).astype(np.int64))[:emissive]
        k = len(picks)
        while k < emissive:  # top up (duplicates are rare)
            extra = (rng.uniform_array(emissive) * nt).astype(np.int64)
            picks = np.unique(np.concatenate([picks, extra]))[:emissive]
            k = len(picks)
        levels = np.exp(np.log(0.1) + (np.log(10.0) - np.log(0.1)) * rng.uniform_array(emissive))
        for tri, e in zip(picks, levels):
            base = mats[tri_mat[tri]]
            mats.append(material(base[0:3], base[9], emission=(e, e, e)))
            tri_mat[tri] = len(mats) - 1
    d = box * 0.9
    cam = dict(eye=(0.0, 0.0, d), lookat=(0.0, 0.0, 0.0), up=(0.0, 1.0, 0.0), fov=60.0)
    if instanced:
        meshes = [(b.astype(np.float32), unit.astype(np.float32), idx.astype(np.uint32)) for b in bases]
        return InstancedSceneData(meshes, np.array([k % base_meshes for k in range(instances)], np.int32),
                                  np.asarray(X, np.float32), tri_mat, np.asarray(mats, np.float32), camera=cam,
                                  name=f"spheres{instances}x{ntb}_instanced")
    return SceneData(np.concatenate(P).astype(np.float32), np.concatenate(N).astype(np.float32),
                     np.concatenate(I).astype(np.uint32), tri_mat, np.asarray(mats, np.float32), camera=cam,
                     name=f"spheres{instances}x{ntb}")


def spheres_lamps(instances: int = 256, slices: int = 64, stacks: int = 33, lamps: int = 4096, box: float = 20.0,
                  seed: int = 2, lamp_size: float = 0.15, clearance: float = 0.3) -> SceneData:
    """The instanced-sphere scene with its 4,096 emitters as separate lamps:
    small black equilateral triangles (side `lamp_size`, random orientation,
    log-uniform emission in [0.1, 10]) at least `clearance` metres from every
    sphere's bounding sphere and from each other, so that next-event
    estimation never sees an emitter closer than about the clearance (in
    spheres() the emitters are triangles of the sphere meshes)."""
    base = spheres(instances=instances, slices=slices, stacks=stacks, emissive=0, box=box, seed=seed)
    rng = SplitMix64(seed + 1000)
    nv = len(base.positions)
    # bounding spheres of the instances (vertex ranges are consecutive per instance)
    per = nv // instances
    P = base.positions.reshape(instances, per, 3).astype(np.float64)
    centers = P.mean(axis=1)
    radii = np.linalg.norm(P - centers[:, None, :], axis=2).max(axis=1)
    mats = [list(m) for m in base.materials]
    pos, nrm, idx, tmat = [], [], [], []
    placed = 0
    lamp_c = np.zeros((lamps, 3))
    while placed < lamps:
        c = np.array([rng.uniform(-box / 2, box / 2) for _ in range(3)])
        if np.any(np.linalg.norm(centers - c, axis=1) < radii + clearance + lamp_size):
            continue
        if placed and np.any(np.linalg.norm(lamp_c[:placed] - c, axis=1) < clearance + lamp_size):
            continue
        lamp_c[placed] = c
        R = _rotation(rng)
        tri = np.array([[0.0, 1.0, 0.0], [-0.866, -0.5, 0.0], [0.866, -0.5, 0.0]]) * (lamp_size / 1.732)
        v = tri @ R.T + c
        n = np.cross(v[1] - v[0], v[2] - v[0])
        n /= np.linalg.norm(n)
        e = float(np.exp(np.log(0.1) + (np.log(10.0) - np.log(0.1)) * rng.uniform()))
        mats.append(material((0.0, 0.0, 0.0), 1.0, emission=(e, e, e)))
        pos.append(v)
        nrm.append(np.tile(n, (3, 1)))
        idx.append(np.arange(3) + nv + 3 * placed)
        tmat.append(len(mats) - 1)
        placed += 1
    return SceneData(np.concatenate([base.positions, np.concatenate(pos).astype(np.float32)]),
                     np.concatenate([base.normals, np.concatenate(nrm).astype(np.float32)]),
                     np.concatenate([base.indices, np.array(idx, np.uint32)]).astype(np.uint32),
                     np.concatenate([base.tri_material, np.array(tmat, np.uint32)]),
                     np.asarray(mats, np.float32), camera=base.camera, name=f"spheres_lamps{instances}")


def random_soup(n: int, seed: int = 3, extent: float = 1.0, size: float = 0.2) -> SceneData:
    """n independent random triangles (parity tests of the build and traversal)."""
    rng = SplitMix64(seed)
    c = (rng.uniform_array(3 * n).reshape(n, 3) - 0.5) * 2 * extent
    off = (rng.uniform_array(9 * n).reshape(n, 3, 3) - 0.5) * size
    pos = (c[:, None, :] + off).reshape(-1, 3)
    e1 = off[:, 1] - off[:, 0]
    e2 = off[:, 2] - off[:, 0]
    nrm = np.cross(e1, e2)
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True) + 1e-30
    nrm = np.repeat(nrm, 3, axis=0)
    idx = np.arange(3 * n).reshape(n, 3)
    mats = [material((0.5, 0.5, 0.5), 0.5)]
    return SceneData(pos.astype(np.float32), nrm.astype(np.float32), idx.astype(np.uint32),
                     np.zeros(n, np.uint32), np.asarray(mats, np.float32), name=f"soup{n}")


def sky_room(seed: int = 4) -> SceneData:
    """Low-variance test scene: a 10 m floor, two boxes and a large emissive
    ceiling 4 m up (no light closer than ~2 m to any lit surface)."""
    rng = SplitMix64(seed)
    mats, b = [], _Builder()
    mats.append(material((0.6, 0.6, 0.6), 0.8))
    b.add(*grid_quad((-5, 0, -5), (10, 0, 0), (0, 0, 10), 4, (0, 1, 0)), 0)
    for center, half in (((-1.0, 0.7, -0.5), (0.6, 0.7, 0.6)), ((1.2, 0.5, 0.6), (0.5, 0.5, 0.5))):
        mats.append(material([rng.uniform(0.3, 0.8) for _ in range(3)], rng.uniform(0.3, 1.0)))
        for face in box_mesh(center, half, 2):
            b.add(*face, len(mats) - 1)
    mats.append(material((0.0, 0.0, 0.0), 1.0, emission=(1.0, 1.0, 1.0)))
    b.add(np.array([(-5, 4, -5), (5, 4, -5), (5, 4, 5), (-5, 4, 5)]), np.tile((0.0, -1.0, 0.0), (4, 1)),
          np.array([(0, 2, 1), (0, 3, 2)]), len(mats) - 1)
    cam = dict(eye=(0.0, 3.0, 4.0), lookat=(0.0, 0.5, 0.0), up=(0.0, 1.0, 0.0), fov=50.0)
    return b.build(mats, camera=cam, name="sky_room")


def _texture(rng: SplitMix64, w: int, h: int, fn) -> np.ndarray:
    img = np.zeros((h, w, 4), np.uint8)
    for y in range(h):
        for x in range(w):
            img[y, x] = np.clip(np.round(np.asarray(fn(x, y, rng), np.float64) * 255.0), 0, 255)
    return img


def textured_room(seed: int = 6, coarse: bool = False) -> SceneData:
    """The Cornell room with every texture kind of calculate_material /
    calculate_normal in use (RGBA8 images, 5 per material: diffuse, normal,
    specular, emission, roughness):
    * floor: checker diffuse texture tiled 3 x 3 (wrap addressing) and a
      roughness texture;
    * back wall: noise diffuse, normal map, specular texture (the GGX lobe);
    * left wall: the normal map alone; the tall box: the noise diffuse at
      negative texture coordinates; the icosphere: checker diffuse and the
      normal map over a spherical (u, v) and tangent frame;
    * the ceiling light: an emission texture beside its constant emission
      15; a small wall quad with an emission texture and zero constant
      emission (its triangles are lights of zero intensity);
    * right wall, ceiling and the short box untextured.
    coarse=True tessellates it lightly (580 triangles instead of 10,084)."""
    gw, gb, lvl = (4, 2, 2) if coarse else (16, 10, 4)
    rng = SplitMix64(seed)
    checker = _texture(rng, 16, 16, lambda x, y, r: (0.85, 0.8, 0.7, 1.0) if (x // 4 + y // 4) % 2 else
                       (0.15, 0.25, 0.45, 1.0))
    noise = _texture(rng, 32, 32, lambda x, y, r: (r.uniform(0.1, 0.9), r.uniform(0.1, 0.9), r.uniform(0.1, 0.9), 1.0))

    def bump(x, y, r):
        t = np.array([r.uniform(-0.5, 0.5), r.uniform(-0.5, 0.5), 1.0])
        t /= np.linalg.norm(t)
        return (*(0.5 * t + 0.5), 1.0)

    normal = _texture(rng, 32, 32, bump)
    spec = _texture(rng, 8, 8, lambda x, y, r: (r.uniform(0.0, 0.6),) * 3 + (1.0,))
    rough = _texture(rng, 16, 16, lambda x, y, r: (r.uniform(0.05, 1.0), 0.0, 0.0, 1.0))
    glow = _texture(rng, 4, 4, lambda x, y, r: (r.uniform(0.5, 1.0), r.uniform(0.5, 1.0), r.uniform(0.5, 1.0), 1.0))
    CHECKER, NOISE, NORMAL, SPEC, ROUGH, GLOW = range(6)
    mats, mtex, b = [], [], _Builder()

    def mat(m, diffuse=-1, nmap=-1, specular=-1, emission=-1, roughness=-1):
        mats.append(m)
        mtex.append((diffuse, nmap, specular, emission, roughness))
        return len(mats) - 1

    walls = [((-1, 0, -1), (2, 0, 0), (0, 0, 2), (0, 1, 0), dict(diffuse=CHECKER, roughness=ROUGH), (3.0, 3.0)),
             ((-1, 2, -1), (0, 0, 2), (2, 0, 0), (0, -1, 0), {}, (1.0, 1.0)),
             ((-1, 0, -1), (0, 2, 0), (2, 0, 0), (0, 0, 1), dict(diffuse=NOISE, nmap=NORMAL, specular=SPEC),
              (1.0, 1.0)),
             ((-1, 0, -1), (0, 0, 2), (0, 2, 0), (1, 0, 0), dict(nmap=NORMAL), (2.0, 2.0)),
             ((1, 0, -1), (0, 2, 0), (0, 0, 2), (-1, 0, 0), {}, (1.0, 1.0))]
    for o, du, dv, nrm, tex, scale in walls:
        m = mat(material([rng.uniform(0.2, 0.8) for _ in range(3)], rng.uniform(0.2, 1.0),
                         specular=(0.2, 0.2, 0.2) if "specular" in tex else (0.0, 0.0, 0.0)), **tex)
        uv, tan = grid_frame(du, dv, gw, scale)
        b.add(*grid_quad(o, du, dv, gw, nrm), m, uv, tan)
    for center, half, tex in (((-0.4, 0.6, -0.35), (0.3, 0.6, 0.3), dict(diffuse=NOISE)),
                              ((0.45, 0.3, 0.3), (0.3, 0.3, 0.3), {})):
        m = mat(material([rng.uniform(0.2, 0.8) for _ in range(3)], rng.uniform(0.05, 1.0)), **tex)
        for face in box_mesh(center, half, gb):
            pos = face[0]
            du = pos[gb + 1] - pos[0]  # grid_quad: vertex (a=1, b=0) minus (0, 0), times n
            dv = pos[1] - pos[0]
            uv, tan = grid_frame(du * gb, dv * gb, gb, (2.0, 2.0), (-1.0, -1.5))
            b.add(*face, m, uv, tan)
    m = mat(material((0.7, 0.7, 0.7), 0.4), diffuse=CHECKER, nmap=NORMAL)
    v, f = icosphere(lvl)
    r, c = 0.28, np.array([0.35, 0.88, 0.25])
    uv = np.stack([np.arctan2(v[:, 0], v[:, 2]) / (2 * np.pi) + 0.5, np.arccos(np.clip(v[:, 1], -1, 1)) / np.pi],
                  axis=1) * 2.0
    t = np.stack([v[:, 2], np.zeros(len(v)), -v[:, 0]], axis=1)
    t[np.linalg.norm(t, axis=1) < 1e-6] = (1.0, 0.0, 0.0)
    t /= np.linalg.norm(t, axis=1, keepdims=True)
    bt = np.cross(v, t)
    b.add(c + r * v, v, f, m, uv, np.concatenate([t, bt], axis=1))
    m = mat(material((0.8, 0.8, 0.8), 1.0, emission=(15.0,) * 3), emission=GLOW)
    h = 0.25
    quad_uv = np.array([(0, 0), (1, 0), (1, 1), (0, 1)], np.float64)
    quad_tan = np.tile((1.0, 0.0, 0.0, 0.0, 0.0, 1.0), (4, 1))
    b.add(np.array([(-h, 1.99, -h), (h, 1.99, -h), (h, 1.99, h), (-h, 1.99, h)]),
          np.tile((0.0, -1.0, 0.0), (4, 1)), np.array([(0, 2, 1), (0, 3, 2)]), m, quad_uv, quad_tan)
    m = mat(material((0.5, 0.5, 0.5), 1.0), emission=GLOW)
    b.add(np.array([(0.99, 0.9, -0.2), (0.99, 0.9, 0.2), (0.99, 1.3, 0.2), (0.99, 1.3, -0.2)]),
          np.tile((-1.0, 0.0, 0.0), (4, 1)), np.array([(0, 2, 1), (0, 3, 2)]), m, quad_uv, quad_tan)
    cam = dict(eye=(0.0, 1.0, 3.5), lookat=(0.0, 1.0, 0.0), up=(0.0, 1.0, 0.0), fov=45.0)
    return b.build(mats, textures=[checker, noise, normal, spec, rough, glow], material_textures=mtex, camera=cam,
                   name="textured_room")
