"""The oracle's environment map (ArmadaRTX::set_envmap, armada.cu:86-96) pinned
by what the mathematics fixes.

* atan2 / asin (Maclaurin series after argument reduction, used because
  CUDA's atan2f / asinf are not restated) against numpy in FP64: within a few
  units in the last place, with the signed zeros and the axes;
* the texture lookup (u = atan2(x, z)/2 pi + 1/2, v = asin(y)/pi + 1/2,
  bilinear with wrap, include/cuda/interop.cuh:96-135) against an independent
  numpy restatement;
* Ld_Environment's directions (amadeus_common.cuh:221-233): theta =
  acos(sqrt(1 - u)) about world +z, so z^2 is uniform on [0, 1] and phi is
  uniform;
* rays that miss: the path tracer reads (u, 1 - v) (amadeus_path_tracer.cu:
  204-220), ReSTIR (u, v) (amadeus_restir.cu:566-577);
* a closed form for next-event estimation of a constant environment L over
  a diffuse floor (normal +y, albedo rho, roughness 1e4 so that f = rho/pi),
  with the reference's quirks kept: directions are cosine-distributed about
  world +z, not about the normal, while the pdf used is 1/(4 pi), and the
  floor blocks the half of them that points down.  The expectation is
      4 rho L / pi * int_{z > 0, y > 0} w_y w_z dw = 8 rho L / (3 pi);
  sampling about the normal would give 8 rho L / 3 (pi times more), the
  cosine pdf instead of 1/(4 pi) 2 rho L / 3, no shadowing by the floor
  twice the value: each far outside the 1 % tolerance.  ReSTIR DI's RIS
  with the same candidates estimates the same integral (its weights use the
  same pdf), so it converges there too (within 1.5 %: temporal reuse
  correlates its frames).
"""
import numpy as np
import pytest

import oracle as O
from artemis_amd import look_at, scenes

BIG_ROUGH = 1.0e4


def ulp_err(got, want):
    want32 = want.astype(np.float32)
    spacing = np.spacing(np.abs(want32)).astype(np.float64)
    return np.abs(got.astype(np.float64) - want) / spacing


def test_atan2_matches_numpy():
    rng = np.random.default_rng(1)
    y = np.concatenate([rng.normal(size=20000), rng.normal(size=2000) * 1e-6, [0.0, -0.0, 0.0, -0.0, 1, -1, 0, 0]])
    x = np.concatenate([rng.normal(size=20000), rng.normal(size=2000), [0.0, 0.0, -0.0, -0.0, 0, 0, 1, -1]])
    y, x = y.astype(np.float32), x.astype(np.float32)
    got = O.env_atan2(y, x)
    want = np.arctan2(y.astype(np.float64), x.astype(np.float64))
    assert np.array_equal(np.signbit(got[-8:-4]), np.signbit(want[-8:-4]))  # signed zeros and +-pi
    err = ulp_err(got, want)
    assert err.max() <= 4.0, err.max()


def test_asin_matches_numpy():
    rng = np.random.default_rng(2)
    x = np.concatenate([rng.uniform(-1, 1, 20000), [0.0, -0.0, 0.5, -0.5, 1.0, -1.0, 1 - 2**-24, 2**-20]])
    x = x.astype(np.float32)
    got = O.env_asin(x)
    want = np.arcsin(x.astype(np.float64))
    err = ulp_err(got, want)
    assert err.max() <= 4.0, err.max()
    assert O.env_asin(np.array([1.0000001], np.float32))[0] == np.float32(np.pi / 2)  # clamped


def numpy_lookup(env, d, flip_v):
    """Independent FP64 restatement of the lookup: tex2D, bilinear, wrap."""
    H, W = env.shape[:2]
    d = d.astype(np.float64)
    u = np.arctan2(d[:, 0], d[:, 2]) / (2 * np.pi) + 0.5
    v = np.arcsin(np.clip(d[:, 1], -1, 1)) / np.pi + 0.5
    if flip_v:
        v = 1.0 - v
    x, y = u * W - 0.5, v * H - 0.5
    x0, y0 = np.floor(x), np.floor(y)
    a, b = (x - x0)[:, None], (y - y0)[:, None]
    i0, j0 = x0.astype(int) % W, y0.astype(int) % H
    i1, j1 = (i0 + 1) % W, (j0 + 1) % H
    t = env[..., :3].astype(np.float64)
    return ((1 - a) * (1 - b) * t[j0, i0] + a * (1 - b) * t[j0, i1] + (1 - a) * b * t[j1, i0] + a * b * t[j1, i1])


def far_triangle_scene(rho=0.5):
    pos = np.array([[100, 100, 100], [100.1, 100, 100], [100, 100.1, 100]], np.float32)
    nrm = np.tile(np.array([0, 0, 1], np.float32), (3, 1))
    mats = np.asarray([scenes.material((rho,) * 3, 0.5)], np.float32)
    return scenes.SceneData(pos, nrm, np.array([[0, 1, 2]], np.uint32), np.zeros(1, np.uint32), mats)


@pytest.mark.parametrize("flip_v", [False, True])
def test_lookup_matches_numpy_bilinear(flip_v):
    rng = np.random.default_rng(3)
    env = rng.uniform(0, 4, (5, 7, 4)).astype(np.float32)
    osc = O.OracleScene(far_triangle_scene())
    osc.set_envmap(env)
    d = rng.normal(size=(5000, 3))
    d = np.concatenate([d, np.eye(3), -np.eye(3)])
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    got = osc.env_lookup(d, flip_v)
    want = numpy_lookup(env, d, flip_v)
    np.testing.assert_allclose(got, want, rtol=2e-5, atol=2e-5)


def test_sample_directions_cosine_about_z():
    d = O.env_sample_dirs(7, 200000).astype(np.float64)
    np.testing.assert_allclose(np.linalg.norm(d, axis=1), 1.0, atol=1e-6)
    assert d[:, 2].min() >= 0.0
    z2 = d[:, 2] ** 2  # uniform on [0, 1]
    hist, _ = np.histogram(z2, bins=20, range=(0, 1))
    assert np.abs(hist / len(z2) * 20 - 1).max() < 0.03
    phi = np.arctan2(d[:, 1], d[:, 0])
    hist, _ = np.histogram(phi, bins=16, range=(-np.pi, np.pi))
    assert np.abs(hist / len(phi) * 16 - 1).max() < 0.03


def test_misses_read_the_map():
    rng = np.random.default_rng(4)
    env = rng.uniform(0, 2, (8, 16, 4)).astype(np.float32)
    osc = O.OracleScene(far_triangle_scene())
    osc.set_envmap(env)
    W, H = 12, 8
    cam = look_at((0.0, 0.0, 0.0), (0.0, 0.3, -1.0), aspect=W / H)
    d = O.camera_rays(cam, W, H)[:, 4:7]
    acc = osc.accumulate_pt_f64(cam, W, H, 0, 1, max_depth=3, fixed_offset=True)
    np.testing.assert_allclose(acc.reshape(-1, 3), numpy_lookup(env, d, True), rtol=2e-5, atol=2e-5)
    rs = O.OracleReSTIR(W, H)
    rs.frame(osc, cam, max_depth=3)
    np.testing.assert_allclose(rs.buffers[0][..., :3].reshape(-1, 3), numpy_lookup(env, d, False), rtol=2e-5,
                               atol=2e-5)


def floor_scene(rho):
    s = 1000.0
    pos = np.array([[-s, 0, -s], [s, 0, -s], [s, 0, s], [-s, 0, s]], np.float32)
    nrm = np.tile(np.array([0, 1, 0], np.float32), (4, 1))
    idx = np.array([[0, 2, 1], [0, 3, 2]], np.uint32)
    mats = np.asarray([scenes.material((rho,) * 3, BIG_ROUGH)], np.float32)
    return scenes.SceneData(pos, nrm, idx, np.zeros(2, np.uint32), mats)


def test_floor_under_constant_sky_closed_form():
    rho, L = 0.6, 2.0
    osc = O.OracleScene(floor_scene(rho))
    osc.set_envmap(np.full((4, 8, 4), L, np.float32))
    cam = look_at((0.0, 2.0, 0.0), (0.0, 0.0, 0.0), up=(0.0, 0.0, -1.0), fov=20.0)
    expect = 8.0 * rho * L / (3.0 * np.pi)
    wrong = {"about the normal": 8.0 * rho * L / 3.0, "cosine pdf": 4.0 * rho * L / (3.0 * np.pi) * np.pi / 2,
             "no floor shadow": 2 * expect}
    for name, w in wrong.items():
        assert abs(w / expect - 1) > 0.2, name
    W = H = 4
    frames = 4096
    acc = osc.accumulate_pt_f64(cam, W, H, 0, frames, max_depth=0, nthreads=8)
    got = acc.reshape(-1, 3).mean(axis=0) / frames
    np.testing.assert_allclose(got, got[0], rtol=1e-6)
    print(f"path tracer NEE: {got[0]:.5f}, closed form {expect:.5f}")
    assert got[0] == pytest.approx(expect, rel=0.01)
    # temporal reuse correlates the frames: more of them, and a wider band
    # (five seeds of this run land within +-1 %)
    rs = O.OracleReSTIR(8, 8, user_seed=5)
    for _ in range(1024):
        rs.frame(osc, cam, max_depth=0, candidates=32, m_cap=20, spatial_passes=1, neighbours=0, nthreads=8)
    col = rs.buffers[0][..., :3].astype(np.float64).mean(axis=(0, 1))
    print(f"ReSTIR DI: {col[0]:.5f}, closed form {expect:.5f}")
    assert col[0] == pytest.approx(expect, rel=0.015)
