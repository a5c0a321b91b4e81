"""BSDF and light sampling of the oracle (include/cuda/brdf.cuh,
include/optix/lighting.cuh) pinned by quadrature, moments and symmetry."""
import ctypes as C

import numpy as np
import pytest

import oracle as O


def arr(*v):
    return np.array(v, np.float32)


def fptr(a):
    return a.ctypes.data_as(O.fp)


def pdf(n, wi, wo, kd, ks, r):
    return O.lib().orc_ggx_pdf(fptr(n), fptr(wi), fptr(wo), fptr(kd), fptr(ks), r)


def brdf(n, wi, wo, kd, ks, r):
    out = np.zeros(3, np.float32)
    O.lib().orc_ggx_brdf(fptr(n), fptr(wi), fptr(wo), fptr(kd), fptr(ks), r, fptr(out))
    return out


@pytest.mark.parametrize("rough", [0.3, 0.6, 1.0])
def test_pdf_mass_closed_form_at_normal_incidence(rough):
    # brdf.cuh:357-382: mixture (1-t) cos/pi + t D(h) (n.h) / (4 wi.h) with
    # t = max(Ks/(Kd+Ks), 0.25) = 0.25 for Ks = 0.  At wo = n a half vector at
    # angle th_h reflects to 2 th_h, so the GGX lobe keeps the mass of
    # th_h < 45 deg: P(tan^2 th_h < 1) = 1 / (1 + alpha^2) (GGX CDF).
    n = arr(0, 0, 1)
    wo = arr(0, 0, 1)
    kd, ks = arr(0.5, 0.5, 0.5), arr(0, 0, 0)
    nt, nph = 400, 64
    th = (np.arange(nt) + 0.5) / nt * (np.pi / 2)
    total = 0.0
    for t in th:
        # pdf is azimuthally symmetric at normal incidence
        wi = arr(np.sin(t), 0.0, np.cos(t))
        total += pdf(n, wi, wo, kd, ks, rough) * np.sin(t) * (np.pi / 2 / nt) * 2 * np.pi
    assert total == pytest.approx(0.75 + 0.25 / (1.0 + rough * rough), abs=2e-3)


def test_pdf_zero_below_surface():
    n, wo = arr(0, 0, 1), arr(0, 0.6, 0.8)
    assert pdf(n, arr(0, 0.6, -0.8), wo, arr(0.5, 0.5, 0.5), arr(0, 0, 0), 0.5) == 0.0


def test_brdf_reciprocity():
    rng = np.random.default_rng(1)
    n = arr(0, 0, 1)
    for _ in range(200):
        a, b = rng.normal(size=3), rng.normal(size=3)
        a[2], b[2] = abs(a[2]) + 0.1, abs(b[2]) + 0.1
        a, b = (a / np.linalg.norm(a)).astype(np.float32), (b / np.linalg.norm(b)).astype(np.float32)
        kd, ks = rng.uniform(0.2, 0.8, 3).astype(np.float32), rng.uniform(0, 0.5, 3).astype(np.float32)
        r = float(rng.uniform(0.05, 1))
        np.testing.assert_allclose(brdf(n, a, b, kd, ks, r), brdf(n, b, a, kd, ks, r), rtol=2e-5)


def test_lambert_part_and_zero_specular_lobe():
    # f = Kd/pi + GGX; with specular 0 Schlick still leaves (1 - wi.h)^5
    # (brdf.cuh:205-206): at wi = wo that term vanishes and f = Kd/pi + 0.
    n, w = arr(0, 0, 1), arr(0.0, 0.6, 0.8)
    kd = arr(0.3, 0.5, 0.7)
    np.testing.assert_allclose(brdf(n, w, w, kd, arr(0, 0, 0), 0.5), kd / np.float32(np.pi), rtol=1e-6)


def test_sampling_is_cosine_when_specular_zero():
    # brdf.cuh:397-398 picks the GGX lobe with t = Ks/(Kd+Ks) = 0 here, so every
    # sample is cosine-distributed: E[cos] = 2/3, E[cos^2] = 1/2.
    n, wo = arr(0, 0, 1), arr(0.3, 0.0, 0.954)
    kd, ks = arr(0.5, 0.5, 0.5), arr(0, 0, 0)
    cos = []
    seed = np.zeros(3, np.float32)
    wi = np.zeros(3, np.float32)
    for i in range(20000):
        seed[:] = np.array([i, 7, 11], np.uint32).view(np.float32)
        O.lib().orc_ggx_sample(fptr(n), fptr(wo), fptr(kd), fptr(ks), 0.5, fptr(seed), fptr(wi))
        cos.append(wi[2])
    cos = np.array(cos)
    assert cos.min() >= 0.0
    assert cos.mean() == pytest.approx(2 / 3, abs=0.01)
    assert (cos ** 2).mean() == pytest.approx(0.5, abs=0.01)


def test_light_samples_uniform_on_triangle():
    # lighting.cuh:133-146: fold u+v>1 -> uniform over the triangle
    light = arr(0, 0, 0, 2, 0, 0, 0, 1, 0, 5, 5, 5)
    seed = np.zeros(3, np.float32)
    p = np.zeros(3, np.float32)
    pts = []
    for i in range(20000):
        seed[:] = np.array([i, 3, 5], np.uint32).view(np.float32)
        O.lib().orc_light_sample(fptr(light), fptr(seed), fptr(p))
        pts.append(p.copy())
    pts = np.array(pts)
    assert np.all(pts[:, 0] >= 0) and np.all(pts[:, 1] >= 0) and np.all(pts[:, 0] / 2 + pts[:, 1] <= 1 + 1e-6)
    np.testing.assert_allclose(pts.mean(0), [2 / 3, 1 / 3, 0], atol=0.01)
    # half the mass lies below the line through the midpoints? check u-marginal
    u = pts[:, 0] / 2
    assert np.mean(u < 0.5) == pytest.approx(0.75, abs=0.01)
