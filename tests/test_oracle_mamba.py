"""The oracle's restatement of the editor's G-buffer ReSTIR DI
(editor/optix/mamba_shader.cu:35-505, driven by editor/optix/mamba.cu)
pinned by what the mathematics fixes.

Scene: two small triangle lights (Le = 10) over a diffuse floor (Kd = 0.5);
one 1x1 camera looks straight down at the floor point x = 0, so the pixel's
expected direct lighting is the quadrature of test_oracle_estimators.py:
  sum_lights  Kd/pi * Le * int_A cos(th_x) cos(th_l) / r^2 dA.

* Initial RIS only (8 candidates, W = w / (M p_hat)): unbiased, so the
  accumulated colour converges to that integral.
* Temporal + spatial reuse with one surface: every merge keeps the expected
  value (the same target at every pixel), so it converges there too.
* The light pick: sample_light (mamba_shader.cu:85-124) chooses an area light
  (an emissive submesh) uniformly, then a triangle of it uniformly, and
  returns 1 / (area * lights), leaving out the 1/triangles factor of the
  second choice.  With both light triangles in ONE material (one area light
  of two triangles) the estimate is therefore exactly half the integral; with
  one material per triangle it is the integral.
* M: 8 candidates; the temporal merge caps the previous M at 200
  (:352), so M saturates at 208; the spatial merge adds two neighbours capped
  at 200 each (:443, :457) - at 1x1 both neighbours are the pixel itself,
  so the stored M saturates at 208 + 400 = 608.
"""
import numpy as np
import pytest

import oracle as O
from artemis_amd import look_at, scenes
from tests.test_oracle_estimators import KD, LE, direct_closed_form, two_light_scene


def split_light_scene():
    """The two-light scene with one emissive material per light triangle."""
    data, lights = two_light_scene()
    mats = np.concatenate([data.materials, data.materials[1:2]])
    tri_mat = np.array([0, 0, 1, 2], np.uint32)
    return scenes.SceneData(data.positions, data.normals, data.indices, tri_mat, mats), lights


@pytest.fixture(scope="module")
def setup():
    shared, lights = two_light_scene()
    split, _ = split_light_scene()
    cam = look_at((0.0, 3.0, 0.0), (0.0, 0.0, 0.0), up=(0.0, 0.0, -1.0), fov=1e-3)
    return O.OracleScene(shared), O.OracleScene(split), cam, direct_closed_form(lights)


def run(osc, cam, frames, temporal, spatial, w=1, h=1):
    m = O.OracleMamba(osc, w, h)
    for f in range(frames):
        col = m.frame(cam, temporal=temporal, spatial=spatial, nthreads=1)
    return m, col


def test_light_grouping(setup):
    shared, split, _, _ = setup
    assert O.OracleMamba(shared, 1, 1).light_count() == 1
    assert O.OracleMamba(split, 1, 1).light_count() == 2


@pytest.mark.parametrize("temporal,spatial", [(False, False), (True, False), (True, True)])
def test_converges_to_closed_form(setup, temporal, spatial):
    _, split, cam, expect = setup
    m, col = run(split, cam, 3000, temporal, spatial)
    got = col[0, 0, :3]
    assert np.allclose(got, got[0])  # gray
    assert col[0, 0, 3] == 1.0
    assert got[0] == pytest.approx(expect, rel=0.015)


def test_light_pick_omits_triangle_choice(setup):
    shared, _, cam, expect = setup
    _, col = run(shared, cam, 3000, False, False)
    assert col[0, 0, 0] == pytest.approx(0.5 * expect, rel=0.015)


@pytest.mark.parametrize("temporal,spatial,m_r,m_p", [(False, False, 8, 0), (True, False, 208, 208),
                                                      (True, True, 608, 608)])
def test_reservoir_counts(setup, temporal, spatial, m_r, m_p):
    _, split, cam, _ = setup
    m, _ = run(split, cam, 40, temporal, spatial)
    r, p = m.reservoirs(0)[0, 0], m.reservoirs(1)[0, 0]
    assert r[12] == m_r
    assert p[12] == m_p
    # the selected sample is a point of one of the two lights, facing down
    assert r[1] in (np.float32(1.0), np.float32(1.2))
    np.testing.assert_array_equal(r[3:6], [0, -1, 0])
    np.testing.assert_array_equal(r[6:9], [LE, LE, LE])
    assert r[11] > 0


def test_sky_pixels_are_empty(setup):
    _, split, _, _ = setup
    cam = look_at((0.0, 3.0, 0.0), (0.0, 10.0, 0.0), up=(0.0, 0.0, -1.0), fov=30.0)
    m, col = run(split, cam, 3, True, True, w=4, h=4)
    assert not np.any(col[..., :3])
    assert not np.any(m.reservoirs(0)[..., 12])
    assert not np.any(m.reservoirs(1)[..., 12])


def test_camera_move_restarts_accumulation(setup):
    """dirty (mamba.cu:1056-1068): a moved camera restarts the running mean,
    so the colour after the move is one frame's estimate."""
    _, split, cam, expect = setup
    m = O.OracleMamba(split, 1, 1)
    for f in range(50):
        m.frame(cam, temporal=False, spatial=False, nthreads=1)
    cam2 = look_at((0.0, 3.0, 0.0), (0.0, 0.0, 1e-4), up=(0.0, 0.0, -1.0), fov=1e-3)
    moved = m.frame(cam2, temporal=False, spatial=False, nthreads=1).copy()
    # an independent single frame at the same frame index and camera
    fresh = O.OracleMamba(split, 1, 1)
    for f in range(50):
        fresh.frame(cam2, temporal=False, spatial=False, accumulate=False, nthreads=1)
    single = fresh.frame(cam2, temporal=False, spatial=False, accumulate=False, nthreads=1)
    np.testing.assert_array_equal(moved, single)


def test_kd_scales_linearly(setup):
    """Halving the floor albedo halves every frame's estimate exactly (the
    target's luminance scales, the selection does not change)."""
    split, lights = split_light_scene()
    half = scenes.SceneData(split.positions, split.normals, split.indices, split.tri_material,
                            split.materials.copy())
    half.materials[0] = np.asarray(scenes.material((KD / 2,) * 3, 0.05), np.float32)
    cam = look_at((0.0, 3.0, 0.0), (0.0, 0.0, 0.0), up=(0.0, 0.0, -1.0), fov=1e-3)
    a = O.OracleMamba(O.OracleScene(split), 1, 1)
    b = O.OracleMamba(O.OracleScene(half), 1, 1)
    for f in range(5):
        ca = a.frame(cam, temporal=False, spatial=False, nthreads=1)[0, 0, 0]
        cb = b.frame(cam, temporal=False, spatial=False, nthreads=1)[0, 0, 0]
        assert cb == pytest.approx(ca / 2, rel=1e-5)
