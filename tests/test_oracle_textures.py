"""The oracle's material textures and normal maps (calculate_material /
calculate_normal, source/optix/amadeus_common.cuh:74-187; texture
descriptor include/cuda/interop.cuh:11-51) pinned by independent numpy
restatements and by equivalences the reference's rules fix.

* the texture fetch (RGBA8 read as byte / 255, bilinear, wrap addressing)
  against numpy in FP64 at coordinates inside, outside and far outside
  [0, 1], negative ones included;
* the primary hit's albedo and normal AOVs of the path tracer on a scene
  with every texture kind: (u, v) = interpolated tex_coords with v flipped
  (LOAD_INTERSECTION_DATA, amadeus_common.cuh:498-502), the diffuse texture
  replacing the diffuse colour, the normal map n = normalize(T n_x + B n_y +
  N n_z) with n = 2 rgb - 1.  The plausible mistakes (no v flip, T and B
  swapped, the map's rgb used without 2x - 1, the geometric instead of the
  shading normal in the frame) each move the AOVs far outside the tolerance;
* an emission texture: its triangles are triangle lights with the
  material's constant emission (armada.cu:117-150 with the emissive test of
  :264-265); with zero constant emission a quad seen at depth 0 shows the
  texel's rgb as its radiance;
* specular and roughness textures: a 1x1 texture renders as the constant
  material of the texel's value (the roughness from the r channel, the other
  channels set to values that would show if read instead).
"""
import numpy as np
import pytest

import oracle as O
from artemis_amd import look_at, scenes


def np_bilinear(img, u, v):
    """tex2D <float4>: normalised coordinates, bilinear, wrap, byte / 255."""
    H, W = img.shape[:2]
    t = img.astype(np.float64) / 255.0
    x = u.astype(np.float64) * W - 0.5
    y = v.astype(np.float64) * H - 0.5
    x0, y0 = np.floor(x), np.floor(y)
    a, b = (x - x0)[:, None], (y - y0)[:, None]
    i0, j0 = x0.astype(np.int64) % W, y0.astype(np.int64) % H
    i1, j1 = (i0 + 1) % W, (j0 + 1) % H
    return (1 - a) * (1 - b) * t[j0, i0] + a * (1 - b) * t[j0, i1] + (1 - a) * b * t[j1, i0] + a * b * t[j1, i1]


def test_fetch_matches_numpy_bilinear_wrap():
    rng = np.random.default_rng(1)
    img = rng.integers(0, 256, (7, 5, 4), dtype=np.uint8)
    uv = np.concatenate([rng.uniform(0, 1, (3000, 2)), rng.uniform(-3, 4, (3000, 2)),
                         [[0, 0], [1, 1], [0.5, 0.5], [-1e-7, 1 + 1e-7]], rng.uniform(-200, 200, (500, 2))])
    uv = uv.astype(np.float32)
    got = O.tex_lookup(img, uv)
    want = np_bilinear(img, uv[:, 0], uv[:, 1])
    # the filter weights carry the float32 rounding of u * W - 0.5, which
    # grows with |u|
    np.testing.assert_allclose(got[:6004], want[:6004], atol=2e-6)
    np.testing.assert_allclose(got[6004:], want[6004:], atol=2e-4)
    # a texel centre reads the texel exactly
    c = np.array([[(2 + 0.5) / 5, (3 + 0.5) / 7]], np.float32)
    assert np.array_equal(O.tex_lookup(img, c)[0], img[3, 2].astype(np.float32) / np.float32(255.0))
    # NaN coordinates read zero
    assert np.array_equal(O.tex_lookup(img, np.array([[np.nan, 0.5]], np.float32))[0], np.zeros(4, np.float32))


def primary_hits(osc, cam, W, H):
    rays = O.camera_rays(cam, W, H, jitter=True, sample=0, user_seed=0)
    hits, _ = osc.trace(rays)
    return rays, hits


def np_surface(d, rays, hits, variant=""):
    """FP64 restatement of the primary hit's (u, v), textured albedo and
    normal-mapped shading normal."""
    prim = hits[:, 3].view(np.int32)
    ok = prim >= 0
    prim = np.where(ok, prim, 0)
    b1, b2 = hits[:, 1].astype(np.float64)[:, None], hits[:, 2].astype(np.float64)[:, None]
    w = 1.0 - b1 - b2
    tri = d.indices[prim].astype(np.int64)
    P = d.positions.astype(np.float64)
    lerp = lambda A: w * A[tri[:, 0]] + b1 * A[tri[:, 1]] + b2 * A[tri[:, 2]]
    uv = lerp(d.uvs.astype(np.float64))
    if variant != "no_flip":
        uv[:, 1] = 1.0 - uv[:, 1]
    m = d.tri_material[prim]
    mt = d.material_textures[m]
    albedo = d.materials[m, 0:3].astype(np.float64)
    for k, img in enumerate(d.textures):
        sel = ok & (mt[:, 0] == k)
        if sel.any():
            albedo[sel] = np_bilinear(img, uv[sel, 0], uv[sel, 1])[:, :3]
    ng = np.cross(P[tri[:, 1]] - P[tri[:, 0]], P[tri[:, 2]] - P[tri[:, 0]])
    dirs = rays[:, 4:7].astype(np.float64)
    ng = np.where((np.sum(ng * dirs, 1) > 0)[:, None], -ng, ng)
    N = lerp(d.normals.astype(np.float64))
    N = np.where((np.sum(N * ng, 1) < 0)[:, None], -N, N)
    N /= np.linalg.norm(N, axis=1, keepdims=True)
    if variant == "geometric_frame":
        N = ng / np.linalg.norm(ng, axis=1, keepdims=True)
    n = N.copy()
    Tn = lerp(d.tangents[:, 0:3].astype(np.float64))
    Bn = lerp(d.tangents[:, 3:6].astype(np.float64))
    Tn /= np.maximum(np.linalg.norm(Tn, axis=1, keepdims=True), 1e-300)
    Bn /= np.maximum(np.linalg.norm(Bn, axis=1, keepdims=True), 1e-300)
    if variant == "swap_tb":
        Tn, Bn = Bn, Tn
    for k, img in enumerate(d.textures):
        sel = ok & (mt[:, 1] == k)
        if sel.any():
            t = np_bilinear(img, uv[sel, 0], uv[sel, 1])[:, :3]
            if variant != "no_remap":
                t = 2.0 * t - 1.0
            v = Tn[sel] * t[:, 0:1] + Bn[sel] * t[:, 1:2] + N[sel] * t[:, 2:3]
            n[sel] = v / np.linalg.norm(v, axis=1, keepdims=True)
    return ok, albedo, n, mt


def test_primary_aovs_follow_the_textures():
    d = scenes.textured_room()
    osc = O.OracleScene(d)
    W, H = 96, 72
    cam = look_at(**d.camera, aspect=W / H)
    rays, hits = primary_hits(osc, cam, W, H)
    (col, nrm, alb, pos), _ = osc.render_pt(cam, W, H, frames=1, max_depth=0)
    nrm, alb = nrm.reshape(-1, 4)[:, :3], alb.reshape(-1, 4)[:, :3]
    ok, albedo, n, mt = np_surface(d, rays, hits)
    textured = ok & (mt[:, 0] >= 0)
    mapped = ok & (mt[:, 1] >= 0)
    assert textured.sum() > 1000 and mapped.sum() > 1000
    # float32 interpolation of (u, v) against FP64: a few 1e-6 of a texel step
    np.testing.assert_allclose(alb[ok], albedo[ok], atol=1e-5)
    np.testing.assert_allclose(nrm[ok], n[ok], atol=1e-4)
    # each plausible mistake moves the AOVs far outside the tolerances
    for variant, which in (("no_flip", "albedo"), ("swap_tb", "normal"), ("no_remap", "normal"),
                           ("geometric_frame", "normal")):
        _, a2, n2, _ = np_surface(d, rays, hits, variant)
        err = np.abs(alb[ok] - a2[ok]).max() if which == "albedo" else np.abs(nrm[ok] - n2[ok]).max()
        assert err > 1e-2, variant


def emissive_quad_scene(const_emission):
    pos = np.array([[-1, -1, 0], [1, -1, 0], [1, 1, 0], [-1, 1, 0]], np.float32)
    nrm = np.tile(np.array([0, 0, 1], np.float32), (4, 1))
    idx = np.array([[0, 1, 2], [0, 2, 3]], np.uint32)
    mats = np.asarray([scenes.material((0.5, 0.5, 0.5), 1.0, emission=(const_emission,) * 3)], np.float32)
    d = scenes.SceneData(pos, nrm, idx, np.zeros(2, np.uint32), mats)
    rng = np.random.default_rng(7)
    d.uvs = np.array([[0, 0], [1, 0], [1, 1], [0, 1]], np.float32)
    d.tangents = np.tile(np.array([1, 0, 0, 0, 1, 0], np.float32), (4, 1))
    d.textures = [rng.integers(0, 256, (6, 9, 4), dtype=np.uint8)]
    d.material_textures = np.array([[-1, -1, -1, 0, -1]], np.int32)
    return d


@pytest.mark.parametrize("const_emission", [0.0, 3.0])
def test_emission_texture_radiance_and_lights(const_emission):
    d = emissive_quad_scene(const_emission)
    osc = O.OracleScene(d)
    L = osc.lights()
    assert len(L) == 2  # emissive by texture even with zero constant emission
    np.testing.assert_array_equal(L[:, 9:12], np.full((2, 3), const_emission, np.float32))
    if const_emission:
        return  # next-event estimation from the offset hit point would add light from the quad itself
    W = H = 16
    cam = look_at((0.0, 0.0, 2.0), (0.0, 0.0, 0.0), fov=40.0)
    rays, hits = primary_hits(osc, cam, W, H)
    (col, *_), _ = osc.render_pt(cam, W, H, frames=1, max_depth=0)
    ok, *_ = np_surface(d, rays, hits)
    assert ok.all()
    prim = hits[:, 3].view(np.int32)
    b1, b2 = hits[:, 1].astype(np.float64), hits[:, 2].astype(np.float64)
    tri = d.indices[prim].astype(np.int64)
    w = 1 - b1 - b2
    u = w * d.uvs[tri[:, 0], 0] + b1 * d.uvs[tri[:, 1], 0] + b2 * d.uvs[tri[:, 2], 0]
    v = 1 - (w * d.uvs[tri[:, 0], 1] + b1 * d.uvs[tri[:, 1], 1] + b2 * d.uvs[tri[:, 2], 1])
    want = np_bilinear(d.textures[0], u, v)[:, :3]
    # the quad's lights have zero intensity: the pixel is the texel's radiance
    np.testing.assert_allclose(col.reshape(-1, 4)[:, :3], want, atol=3e-6)


def test_specular_and_roughness_textures_equal_constant_materials():
    base = scenes.cornell_glossy()
    rng = np.random.default_rng(3)
    mats = np.array(base.materials, np.float32)
    m_tex = 2  # the back wall
    spec = np.array([200, 90, 30], np.uint8)
    rough = np.uint8(40)
    tex_spec = np.array([[[*spec, 255]]], np.uint8)
    tex_rough = np.array([[[rough, 250, 7, 128]]], np.uint8)  # g, b, a would show if read
    d = scenes.SceneData(base.positions, base.normals, base.indices, base.tri_material, mats.copy(),
                         camera=base.camera)
    d.uvs = rng.uniform(-2, 3, (len(base.positions), 2)).astype(np.float32)
    d.tangents = np.tile(np.array([1, 0, 0, 0, 1, 0], np.float32), (len(base.positions), 1))
    d.textures = [tex_spec, tex_rough]
    mt = np.full((len(mats), 5), -1, np.int32)
    mt[m_tex, 2] = 0
    mt[m_tex, 4] = 1
    d.material_textures = mt
    const = np.array(mats)
    const[m_tex, 3:6] = spec.astype(np.float32) / np.float32(255.0)
    const[m_tex, 9] = np.float32(rough) / np.float32(255.0)
    e = scenes.SceneData(base.positions, base.normals, base.indices, base.tri_material, const, camera=base.camera)
    W, H = 48, 40
    cam = look_at(**base.camera, aspect=W / H)
    a, _ = O.OracleScene(d).render_pt(cam, W, H, frames=4, max_depth=3)
    b, _ = O.OracleScene(e).render_pt(cam, W, H, frames=4, max_depth=3)
    c, _ = O.OracleScene(base).render_pt(cam, W, H, frames=4, max_depth=3)
    ca, cb, cc = a[0][..., :3], b[0][..., :3], c[0][..., :3]
    # a 1x1 texture's bilinear sum can differ from the texel in the last bit,
    # which may turn a lobe choice: nearly every pixel agrees, the mean closely
    close = np.abs(ca - cb) <= 1e-4 * (1 + np.abs(cb))
    assert close.mean() > 0.97
    assert ca.mean() == pytest.approx(cb.mean(), rel=5e-3)
    assert np.abs(ca - cc).max() > 1e-2  # the textures change the image


def test_editor_restir_reads_only_the_diffuse_texture():
    """The editor's ReSTIR takes its material through convert_material
    (editor/optix/common.cuh:34-47: the diffuse texture only) and its surface
    from the ray-cast G-buffer (gbuffer_rtx_shader.cu:95-160: no normal
    map): removing every other texture leaves its frames bit for bit, while
    removing the diffuse textures changes them."""
    d = scenes.textured_room()
    W, H = 32, 24
    cams = [look_at((0.2 * f - 0.2, 1.0, 3.5), (0.0, 0.9, 0.0), aspect=W / H) for f in range(2)]

    def frames(mt):
        e = scenes.SceneData(d.positions, d.normals, d.indices, d.tri_material, d.materials, camera=d.camera)
        e.uvs, e.tangents, e.textures, e.material_textures = d.uvs, d.tangents, d.textures, mt
        om = O.OracleMamba(O.OracleScene(e), W, H)
        for c in cams:
            om.frame(c, temporal=True, spatial=True)
        return om.color.copy(), om.reservoirs(0).copy()

    full = frames(d.material_textures)
    only = d.material_textures.copy()
    only[:, 1:] = -1
    none = np.full_like(d.material_textures, -1)
    a, b, c = frames(d.material_textures), frames(only), frames(none)
    assert np.array_equal(a[0].view(np.uint32), b[0].view(np.uint32))
    assert np.array_equal(a[1].view(np.uint32), b[1].view(np.uint32))
    assert not np.array_equal(a[0], c[0])
    assert np.array_equal(full[0].view(np.uint32), a[0].view(np.uint32))  # deterministic
