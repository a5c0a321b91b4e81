"""Path-tracer NEE (amadeus_common.cuh:255-286) and ReSTIR DI RIS
(amadeus_restir.cu) pinned by the closed-form direct illumination of two small
triangle lights over a diffuse floor.

One camera pixel looks straight down at the floor point x = 0.  With
max_depth 0 the pixel value is Le(floor) + Ld; E[Ld] is
  sum_lights  Kd/pi * Le * int_A cos(th_x) cos(th_l) / r^2 dA,
evaluated here by quadrature over each triangle.  The floor's GGX lobe
(roughness 0.05, specular 0) contributes below 1e-6 of that at these angles.
"""
import numpy as np
import pytest

import oracle as O
from artemis_amd import look_at, scenes

KD = 0.5
LE = 10.0


def two_light_scene():
    floor_p = np.array([[-5, 0, -5], [5, 0, -5], [5, 0, 5], [-5, 0, 5]], np.float64)
    floor_i = np.array([[0, 2, 1], [0, 3, 2]])
    lights = [np.array([[0.7, 1.0, -0.1], [0.9, 1.0, -0.1], [0.8, 1.0, 0.1]]),
              np.array([[-0.9, 1.2, 0.0], [-0.7, 1.2, 0.0], [-0.8, 1.2, 0.25]])]
    pos = [floor_p] + lights
    nrm = [np.tile([0, 1, 0], (4, 1))] + [np.tile([0, -1, 0], (3, 1))] * 2
    idx = [floor_i, np.array([[4, 5, 6]]), np.array([[7, 8, 9]])]
    mats = [scenes.material((KD,) * 3, 0.05), scenes.material((0.0,) * 3, 1.0, emission=(LE,) * 3)]
    data = scenes.SceneData(np.concatenate(pos).astype(np.float32), np.concatenate(nrm).astype(np.float32),
                            np.concatenate(idx).astype(np.uint32), np.array([0, 0, 1, 1], np.uint32),
                            np.asarray(mats, np.float32))
    return data, lights


def direct_closed_form(lights, x=np.zeros(3), n=np.array([0, 1.0, 0])):
    total = 0.0
    m = 400
    for tri in lights:
        a, b, c = tri
        nl = np.cross(b - a, c - a)
        area = 0.5 * np.linalg.norm(nl)
        nl /= np.linalg.norm(nl)
        # midpoint rule on the m^2 sub-triangles' centroids (barycentric grid)
        acc = 0.0
        cnt = 0
        for i in range(m):
            for j in range(m - i):
                for (u, v) in (((i + 1 / 3) / m, (j + 1 / 3) / m),) + ((((i + 2 / 3) / m, (j + 2 / 3) / m),)
                                                                      if i + j < m - 1 else ()):
                    p = a + u * (b - a) + v * (c - a)
                    d = p - x
                    r2 = d @ d
                    w = d / np.sqrt(r2)
                    acc += abs(n @ w) * abs(nl @ w) / r2
                    cnt += 1
        total += KD / np.pi * LE * area * acc / cnt
    return total


@pytest.fixture(scope="module")
def setup():
    data, lights = two_light_scene()
    osc = O.OracleScene(data)
    cam = look_at((0.0, 3.0, 0.0), (0.0, 0.0, 0.0), up=(0.0, 0.0, -1.0), fov=1e-3)
    return osc, cam, direct_closed_form(lights)


def test_lights_from_emissive_triangles(setup):
    osc, _, _ = setup
    L = osc.lights()
    assert L.shape == (2, 12)
    np.testing.assert_allclose(L[:, 9:12], LE)


def test_pt_nee_matches_closed_form(setup):
    osc, cam, expect = setup
    n = 8192
    acc = osc.accumulate_pt_f64(cam, 1, 1, 0, n, max_depth=0, nthreads=1)
    got = acc[0, 0] / n
    assert np.allclose(got, got[0])  # gray
    assert got[0] == pytest.approx(expect, rel=0.01)


def test_restir_ris_matches_closed_form(setup):
    osc, cam, expect = setup
    rs = O.OracleReSTIR(1, 1)
    frames = 256
    for f in range(frames):
        rs.frame(osc, cam, max_depth=0, candidates=32, m_cap=20, spatial_passes=1, neighbours=0, nthreads=1)
    col = rs.buffers[0][0, 0, :3]
    assert col[0] == pytest.approx(expect, rel=0.01)
    r = rs.reservoirs(0)[0, 0]
    assert r[12] <= 20 + 32  # M after the temporal cap: prev capped at 20, plus 32 new candidates
