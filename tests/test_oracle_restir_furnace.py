"""The oracle's ReSTIR DI spatial reuse pinned by closed forms.

ReSTIR DI's reuse (amadeus_restir.cu:182-434: RIS over 32 candidates, the
temporal merge with an M cap, the spatial merge of 5 neighbours with the 1/Z
rule, W = w / (p_hat Z)) is unbiased, so the colour a pixel accumulates over
frames converges to its direct lighting whatever its neighbours see.  Two
scenes make that a closed form.

1. Sphere furnace (test_oracle_furnace.py): inside a tessellated unit sphere
   with a cap of black emitters and diffuse walls at roughness 1e4 (f = rho /
   pi), cos(x) cos(y) / R^2 = 1 / (4 r^2) for any two points, so every wall
   point's direct lighting is rho Le g (g = the cap's share of the area).
   Every target is the same constant there: this pins the normalisation of
   the merges (W, M, Z) with neighbours that all hold valid samples.
2. Two small triangle lights over a diffuse floor (test_oracle_estimators.py),
   seen from above by an 8 x 8 image: each pixel's direct lighting is the
   quadrature Kd/pi Le int cos cos / r^2 dA at its own floor point, and it
   varies across the image by a factor of several.  A neighbour's sample is
   reweighted by the pixel's own target; a merge that kept the neighbour's
   target, or normalised by M where Z is due, would pull each pixel toward
   its neighbours' values and fail the per-pixel comparison.
"""
import numpy as np
import pytest

import oracle as O
from artemis_amd import look_at

from .test_oracle_estimators import KD, LE, two_light_scene
from .test_oracle_furnace import furnace_scene


@pytest.fixture(scope="module")
def furnace():
    rho, le = 0.7, 4.0
    data, g = furnace_scene(rho, le)
    return O.OracleScene(data), rho * le * g


@pytest.mark.parametrize("neighbours,m_cap,passes", [(5, 20, 1), (5, 200, 2), (0, 20, 1)])
def test_furnace_reuse_converges_to_closed_form(furnace, neighbours, m_cap, passes):
    osc, expect = furnace
    W = H = 16
    cam = look_at((0.0, 0.0, 0.0), (1.0, -0.05, 0.2), fov=40.0)
    rs = O.OracleReSTIR(W, H, user_seed=3)
    for _ in range(64):
        rs.frame(osc, cam, max_depth=0, candidates=32, m_cap=m_cap, spatial_passes=passes, neighbours=neighbours,
                 radius=6, nthreads=8)
    col = rs.buffers[0][..., :3].astype(np.float64)
    assert np.all(np.isfinite(col))
    mean = col.mean(axis=(0, 1))
    np.testing.assert_allclose(mean, mean[0], rtol=1e-5)  # gray
    print(f"furnace neighbours={neighbours} m_cap={m_cap} passes={passes}: {mean[0]:.5f} vs {expect:.5f}")
    assert mean[0] == pytest.approx(expect, rel=0.01)
    if neighbours:
        # the spatial merge sums the neighbours' M: above what one pixel's
        # temporal chain (capped at m_cap, plus 32 new candidates) holds
        assert rs.reservoirs(0)[..., 12].max() > m_cap + 32


def direct_at(points, lights, m=160):
    """Kd/pi Le sum_lights int_A cos(th_x) cos(th_l) / r^2 dA at each floor
    point (normal +y): midpoint rule over the m^2 sub-triangles of a
    barycentric grid, vectorised over points."""
    i, j = np.meshgrid(np.arange(m), np.arange(m), indexing="ij")
    up = (i + 1 / 3) / m, (j + 1 / 3) / m
    keep = i + j < m
    dn = (i + 2 / 3) / m, (j + 2 / 3) / m
    keep2 = i + j < m - 1
    uv = np.concatenate([np.stack([up[0][keep], up[1][keep]], 1), np.stack([dn[0][keep2], dn[1][keep2]], 1)])
    out = np.zeros(len(points))
    for a, b, c in lights:
        nl = np.cross(b - a, c - a)
        area = 0.5 * np.linalg.norm(nl)
        nl = nl / np.linalg.norm(nl)
        p = a + uv[:, :1] * (b - a) + uv[:, 1:] * (c - a)
        d = p[None, :, :] - points[:, None, :]
        r2 = (d * d).sum(-1)
        w = d / np.sqrt(r2)[..., None]
        out += KD / np.pi * LE * area * (np.abs(w[..., 1]) * np.abs(w @ nl) / r2).mean(axis=1)
    return out


@pytest.mark.parametrize("neighbours,passes", [(5, 1), (5, 2)])
def test_floor_reuse_matches_each_pixels_closed_form(neighbours, passes):
    data, lights = two_light_scene()
    osc = O.OracleScene(data)
    W = H = 8
    cam = look_at((0.0, 3.0, 0.0), (0.0, 0.0, 0.0), up=(0.0, 0.0, -1.0), fov=36.0)
    rs = O.OracleReSTIR(W, H, user_seed=11)
    frames = 384
    for _ in range(frames):
        rs.frame(osc, cam, max_depth=0, candidates=32, m_cap=20, spatial_passes=passes, neighbours=neighbours,
                 radius=3, nthreads=8)
    col = rs.buffers[0][..., 0].astype(np.float64).ravel()
    x = rs.buffers[3][..., :3].astype(np.float64).reshape(-1, 3)
    # every pixel sees the floor (the position AOV holds the point offset by
    # 1e-3 along the normal, amadeus_common.cuh:22)
    assert np.allclose(x[:, 1], 1e-3, atol=1e-6)
    x[:, 1] = 0.0
    want = direct_at(x, lights)
    rel = col / want - 1
    print(f"floor neighbours={neighbours} passes={passes}: closed forms {want.min():.4f}..{want.max():.4f}, "
          f"per-pixel error max {np.abs(rel).max() * 100:.2f} %, mean {rel.mean() * 100:+.2f} %")
    assert want.max() > 2.0 * want.min()  # the pixels' targets really differ
    assert abs(rel.mean()) < 0.01
    assert np.abs(rel).max() < 0.05
