"""The oracle's path-tracer estimator pinned by closed forms at depth > 0.

The estimator is the reference's (source/optix/amadeus_path_tracer.cu:133-201):
at every vertex value = Le + Ld + T * value(next), T = f |wi.n| / pdf, with
the next vertex's emission counted again although Ld already sampled it, and
a ray past max_depth returning 0 (LOAD_RAYPACKET,
source/optix/amadeus_common.cuh:480-488).  Two scenes make that expectation a
closed form:

1. Sphere furnace (the recursion).  The inside of a finely tessellated unit
   sphere; a polar cap of black emitters (emission Le), the rest diffuse
   albedo rho, specular 0.  Every material has roughness 1e4, where the GGX
   terms of brdf.cuh:337-382 vanish (D ~ 1/alpha^2, G ~ 1/alpha^2), so
   f = rho/pi and the evaluated pdf is the mixture 0.75 cos/pi (t = max(Ks /
   (Kd + Ks), 0.25) = 0.25, brdf.cuh:372-377) while the sample is cosine
   distributed (t = Ks / (Kd + Ks) = 0, brdf.cuh:397-398): T = 4 rho / 3 for
   every sample.  On a sphere cos(x) cos(y) / R^2 = 1 / (4 r^2) for any two
   points, so a cosine-distributed ray hits a region of area A with
   probability A / (4 pi r^2) and NEE's expectation is rho Le g, g = A_light /
   (4 pi r^2).  With D = max_depth, the expected value at a wall vertex of
   depth d is
       W_D = rho Le g
       W_d = rho Le g + (4 rho / 3) (g Le + (1 - g) W_{d+1})     (d < D)
   and a camera ray that hits a wall has expectation W_0.

2. GGX dome (one bounce with a glossy lobe).  A floor disk (Kd, Ks > 0,
   roughness alpha) under a black emitting hemisphere centred on the shading
   point; max_depth 1.  The pixel's expectation is
       Le [ int f(wi, wo) cos dwi + int p_s(wi) f cos / p_e(wi) dwi ]
   over the upper hemisphere: NEE plus the BSDF-sampled ray's T times the
   dome's emission.  f, the sampling density p_s and the evaluated pdf p_e are
   computed here in FP64 from the textbook GGX (Walter et al. 2007: D =
   alpha^2 / (pi cos^4 (alpha^2 + tan^2)^2), Smith G1 = 2 / (1 + sqrt(1 +
   alpha^2 tan^2)), Schlick F, half-vector sampling density D cos_h / (4 wi.h))
   with the reference's choices (the 0.999 clamp of G1, the +1e-6 of the
   denominator, t without the 0.25 floor for sampling, with it for the pdf).

Each test also evaluates the closed form with a plausible mistake (T dropped,
the next vertex's emission not counted again, NEE missing, the pdf floor
missing, one bounce more or less) and checks that every such variant is far
outside the tolerance, so the pin can fail.
"""
import numpy as np
import pytest

import oracle as O
from artemis_amd import look_at, scenes

BIG_ROUGH = 1.0e4


def uv_sphere(slices, stacks, r=1.0, theta_max=np.pi):
    """Vertices on rings of polar angle theta_k = k theta_max / stacks (k = 0
    the +y pole); radial normals; triangles (2 per quad, 1 at the pole)."""
    verts = [[0.0, r, 0.0]]
    for k in range(1, stacks + 1):
        th = theta_max * k / stacks
        for s in range(slices):
            ph = 2 * np.pi * s / slices
            verts.append([r * np.sin(th) * np.cos(ph), r * np.cos(th), r * np.sin(th) * np.sin(ph)])
    if theta_max >= np.pi - 1e-12:
        verts = verts[:-slices] + [[0.0, -r, 0.0]]
    V = np.array(verts)

    def ring(k, s):
        return 1 + (k - 1) * slices + (s % slices)

    tris = []
    for s in range(slices):
        tris.append([0, ring(1, s), ring(1, s + 1)])
    last = stacks if theta_max < np.pi - 1e-12 else stacks - 1
    for k in range(1, last):
        for s in range(slices):
            a, b, c, d = ring(k, s), ring(k, s + 1), ring(k + 1, s), ring(k + 1, s + 1)
            tris += [[a, c, d], [a, d, b]]
    if theta_max >= np.pi - 1e-12:
        south = len(V) - 1
        for s in range(slices):
            tris.append([ring(stacks - 1, s), south, ring(stacks - 1, s + 1)])
    I = np.array(tris)
    N = V / np.linalg.norm(V, axis=1, keepdims=True)
    return V, N, I


def tri_areas(V, I):
    return 0.5 * np.linalg.norm(np.cross(V[I[:, 1]] - V[I[:, 0]], V[I[:, 2]] - V[I[:, 0]]), axis=1)


# ---------------------------------------------------------------------------
# 1. Sphere furnace
# ---------------------------------------------------------------------------
def furnace_scene(rho, le, cap_cos=0.8):
    V, N, I = uv_sphere(128, 96)
    cen = V[I].mean(axis=1)
    light = cen[:, 1] / np.linalg.norm(cen, axis=1) > cap_cos
    mats = [scenes.material((rho,) * 3, BIG_ROUGH), scenes.material((0.0,) * 3, BIG_ROUGH, emission=(le,) * 3)]
    data = scenes.SceneData(V.astype(np.float32), N.astype(np.float32), I.astype(np.uint32),
                            light.astype(np.uint32), np.asarray(mats, np.float32))
    A = tri_areas(V, I)
    g = A[light].sum() / A.sum()
    return data, g


def furnace_closed_form(rho, le, g, depth, t_factor=None, double_count=True, nee=True):
    a = 4.0 * rho / 3.0 if t_factor is None else t_factor
    e = rho * le * g if nee else 0.0
    w = e
    for d in range(depth - 1, -1, -1):
        w = e + a * ((g * le if double_count else 0.0) + (1.0 - g) * w)
    return w


@pytest.fixture(scope="module", params=[(0.6, 4.0, 5), (0.9, 2.0, 2), (0.5, 3.0, 0)], ids=["rho.6-d5", "rho.9-d2",
                                                                                             "rho.5-d0"])
def furnace(request):
    rho, le, depth = request.param
    data, g = furnace_scene(rho, le)
    osc = O.OracleScene(data)
    cam = look_at((0.0, 0.0, 0.0), (1.0, -0.05, 0.2), fov=40.0)
    W = H = 32
    frames = 128
    acc = osc.accumulate_pt_f64(cam, W, H, 0, frames, max_depth=depth, nthreads=8)
    mean = acc.reshape(-1, 3).mean(axis=0) / frames
    return rho, le, depth, g, mean


def test_furnace_sphere_matches_closed_form(furnace):
    rho, le, depth, g, mean = furnace
    assert 0.05 < g < 0.15
    np.testing.assert_allclose(mean, mean[0], rtol=1e-6)  # gray
    expect = furnace_closed_form(rho, le, g, depth)
    print(f"furnace rho={rho} depth={depth}: oracle {mean[0]:.5f} closed form {expect:.5f} "
          f"({(mean[0] / expect - 1) * 100:+.2f} %)")
    assert mean[0] == pytest.approx(expect, rel=0.01)


def test_furnace_planted_mistakes_fail(furnace):
    rho, le, depth, g, mean = furnace
    wrong = {
        "T dropped": furnace_closed_form(rho, le, g, depth, t_factor=1.0),
        "no pdf floor (T = rho)": furnace_closed_form(rho, le, g, depth, t_factor=rho),
        "next emission not counted": furnace_closed_form(rho, le, g, depth, double_count=False),
        "no NEE": furnace_closed_form(rho, le, g, depth, nee=False),
        "one bounce more": furnace_closed_form(rho, le, g, depth + 1),
    }
    if depth > 0:
        wrong["one bounce less"] = furnace_closed_form(rho, le, g, depth - 1)
    if depth == 0:  # the recursion is absent at depth 0; only NEE-level mistakes apply
        wrong = {k: v for k, v in wrong.items() if k in ("no NEE", "one bounce more")}
    for name, w in wrong.items():
        assert abs(w - mean[0]) > 0.04 * mean[0], name


# ---------------------------------------------------------------------------
# 2. GGX dome
# ---------------------------------------------------------------------------
def dome_scene(kd, ks, alpha, le):
    slices, stacks = 192, 64
    V, N, I = uv_sphere(slices, stacks, theta_max=np.pi / 2)
    # floor: a fan over the dome's bottom ring (shared vertices: watertight)
    base = 1 + (stacks - 1) * slices
    centre = len(V)
    V = np.vstack([V, [[0.0, 0.0, 0.0]]])
    Nf = np.vstack([N, [[0.0, 1.0, 0.0]]])
    # the floor uses its own copies of the ring vertices (same positions) with +y normals
    ring = V[base:base + slices]
    fv0 = len(V)
    V = np.vstack([V, ring])
    Nf = np.vstack([Nf, np.tile([0.0, 1.0, 0.0], (slices, 1))])
    floor = np.array([[centre, fv0 + (s + 1) % slices, fv0 + s] for s in range(slices)])
    I = np.vstack([I, floor])
    tm = np.concatenate([np.ones(len(I) - slices, np.uint32), np.zeros(slices, np.uint32)])
    mats = [scenes.material((kd,) * 3, alpha, specular=(ks,) * 3),
            scenes.material((0.0,) * 3, BIG_ROUGH, emission=(le,) * 3)]
    return scenes.SceneData(V.astype(np.float32), Nf.astype(np.float32), I.astype(np.uint32), tm,
                            np.asarray(mats, np.float32))


def ggx_terms(kd, ks, alpha, theta_o, n_th=600, n_ph=720, pdf_floor=True, spec_jacobian=True):
    """FP64 quadrature over the upper hemisphere (n = +z, wo in the xz plane):
    returns (int f cos, int p_s f cos / p_e)."""
    th = (np.arange(n_th) + 0.5) / n_th * (np.pi / 2)
    ph = (np.arange(n_ph) + 0.5) / n_ph * (2 * np.pi)
    TH, PH = np.meshgrid(th, ph, indexing="ij")
    dw = np.sin(TH) * (np.pi / 2 / n_th) * (2 * np.pi / n_ph)
    wi = np.stack([np.sin(TH) * np.cos(PH), np.sin(TH) * np.sin(PH), np.cos(TH)], -1)
    wo = np.array([np.sin(theta_o), 0.0, np.cos(theta_o)])
    h = wi + wo
    h /= np.linalg.norm(h, axis=-1, keepdims=True)
    ci, co, ch = wi[..., 2], wo[2], h[..., 2]
    ih = np.sum(wi * h, -1)

    def G1(c):
        c = np.minimum(c, 0.999)
        tan2 = (1 - c * c) / (c * c)
        return 2.0 / (1.0 + np.sqrt(1.0 + alpha * alpha * tan2))

    tan2h = (1 - ch * ch) / (ch * ch)
    D = alpha ** 2 / (np.pi * ch ** 4 * (alpha ** 2 + tan2h) ** 2)
    F = ks + (1 - ks) * (1 - ih) ** 5
    f = kd / np.pi + F * G1(ci) * G1(co) * D / (4 * ci * co + 1e-6)
    t_s = ks / (kd + ks)
    t_e = max(t_s, 0.25) if pdf_floor else t_s
    spec = D * ch / (4 * ih)
    p_s = (1 - t_s) * ci / np.pi + t_s * spec
    p_e = (1 - t_e) * ci / np.pi + t_e * (spec if spec_jacobian else D * ch)
    nee = np.sum(f * ci * dw)
    bsdf = np.sum(p_s * f * ci / p_e * dw)
    return nee, bsdf


DOME_CASES = [(0.4, 0.5, 0.3, 30.0), (0.2, 0.8, 0.6, 60.0), (0.7, 0.0, 0.5, 45.0)]


@pytest.mark.parametrize("kd,ks,alpha,theta_deg", DOME_CASES)
def test_ggx_dome_single_bounce(kd, ks, alpha, theta_deg):
    le = 2.0
    data = dome_scene(kd, ks, alpha, le)
    osc = O.OracleScene(data)
    th = np.radians(theta_deg)
    eye = (0.5 * np.sin(th), 0.5 * np.cos(th), 0.0)
    cam = look_at(eye, (0.0013, 0.0, 0.0007), up=(0.0, 0.0, 1.0), fov=1e-3)  # off the fan's centre vertex
    W = H = 48
    frames = 96
    acc = osc.accumulate_pt_f64(cam, W, H, 0, frames, max_depth=1, nthreads=8)
    got = acc.reshape(-1, 3).mean(axis=0)[0] / frames
    nee, bsdf = ggx_terms(kd, ks, alpha, th)
    expect = le * (nee + bsdf)
    print(f"dome kd={kd} ks={ks} a={alpha} th={theta_deg}: oracle {got:.5f} closed form {expect:.5f} "
          f"(NEE {le * nee:.4f}, BSDF {le * bsdf:.4f}, {(got / expect - 1) * 100:+.2f} %)")
    assert got == pytest.approx(expect, rel=0.01)
    wrong = {
        "no NEE": le * bsdf,
        "no BSDF bounce": le * nee,
        "pdf without the 0.25 floor": le * (nee + ggx_terms(kd, ks, alpha, th, pdf_floor=False)[1]),
        "evaluated pdf without 1/(4 wi.h)": le * (nee + ggx_terms(kd, ks, alpha, th, spec_jacobian=False)[1]),
    }
    for name, w in wrong.items():
        if name == "pdf without the 0.25 floor" and ks / (kd + ks) >= 0.25:
            continue  # the floor is inactive
        assert abs(w - got) > 0.03 * got, name


# ---------------------------------------------------------------------------
# 3. Sampling density of the glossy lobe (Ks > 0) and Glass directions
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("kd,ks,alpha,theta_deg", [(0.3, 0.6, 0.4, 35.0), (0.5, 0.5, 0.15, 60.0)])
def test_glossy_sampling_histogram(kd, ks, alpha, theta_deg):
    """Samples of brdf.cuh:386-433 with Ks > 0 follow p_s = (1 - t) cos/pi +
    t D(h) cos_h / (4 wi.h), t = Ks / (Kd + Ks): chi-square over (cos, phi)
    bins against the textbook density integrated by quadrature."""
    th = np.radians(theta_deg)
    n = np.array([0, 0, 1], np.float32)
    wo = np.array([np.sin(th), 0, np.cos(th)], np.float32)
    kdv, ksv = np.full(3, kd, np.float32), np.full(3, ks, np.float32)
    count = 400_000
    wi = np.empty((count, 3), np.float32)
    O.lib().orc_ggx_sample_batch(O._f(n), O._f(wo), O._f(kdv), O._f(ksv), alpha, count, 77, O._f(wi))
    up = wi[:, 2] > 0
    nb_c, nb_p = 24, 24
    c_edges = np.linspace(0, 1, nb_c + 1)
    p_edges = np.linspace(-np.pi, np.pi, nb_p + 1)
    hist, _, _ = np.histogram2d(wi[up, 2], np.arctan2(wi[up, 1], wi[up, 0]), bins=[c_edges, p_edges])
    # expected mass per bin: quadrature of p_s on a fine (cos, phi) grid
    sub = 16
    cc = (np.arange(nb_c * sub) + 0.5) / (nb_c * sub)
    pp = -np.pi + (np.arange(nb_p * sub) + 0.5) / (nb_p * sub) * 2 * np.pi
    C_, P_ = np.meshgrid(cc, pp, indexing="ij")
    S_ = np.sqrt(1 - C_ * C_)
    w = np.stack([S_ * np.cos(P_), S_ * np.sin(P_), C_], -1)
    h = w + wo.astype(np.float64)
    h /= np.linalg.norm(h, axis=-1, keepdims=True)
    ch = h[..., 2]
    tan2h = (1 - ch * ch) / (ch * ch)
    D = alpha ** 2 / (np.pi * ch ** 4 * (alpha ** 2 + tan2h) ** 2)
    t = ks / (kd + ks)
    p = (1 - t) * C_ / np.pi + t * D * ch / (4 * np.sum(w * h, -1))
    cell = (1.0 / (nb_c * sub)) * (2 * np.pi / (nb_p * sub))  # d(cos) d(phi) = dw
    mass = (p * cell).reshape(nb_c, sub, nb_p, sub).sum(axis=(1, 3))
    expect = mass * count
    ok = expect > 20
    chi2 = np.sum((hist[ok] - expect[ok]) ** 2 / expect[ok])
    dof = ok.sum()
    # the below-horizon mass of the lobe is what the density leaves out
    assert (~up).sum() / count == pytest.approx(1.0 - mass.sum(), abs=3e-3)
    assert chi2 < dof + 5 * np.sqrt(2 * dof), (chi2, dof)
    # the same histogram against a density with the wrong lobe weight fails
    p_bad = (1 - 0.5 * t) * C_ / np.pi + 0.5 * t * D * ch / (4 * np.sum(w * h, -1))
    bad = (p_bad * cell).reshape(nb_c, sub, nb_p, sub).sum(axis=(1, 3)) * count
    chi2_bad = np.sum((hist[ok] - bad[ok]) ** 2 / np.maximum(bad[ok], 1))
    assert chi2_bad > 3 * (dof + 5 * np.sqrt(2 * dof))


def glass(nrm, wi, eta, entering):
    out = np.empty(3, np.float32)
    O.lib().orc_glass_sample(O._f(np.asarray(nrm, np.float32)), O._f(np.asarray(wi, np.float32)), eta,
                             int(entering), O._f(out))
    return out.astype(np.float64)


@pytest.mark.parametrize("eta", [1.5, 1.33])
def test_glass_directions_snell_and_tir(eta):
    """Glass::sample (brdf.cuh:144-175): refraction obeys Snell's law with
    n1/n2 = 1/eta entering and eta leaving; the reflection test is
    eta sin(theta) > 1 on both sides (quirk kept: an entering ray past
    asin(1/eta) reflects although physics would refract it)."""
    n = np.array([0.0, 0.0, 1.0])
    for deg in np.linspace(0.5, 89.5, 60):
        th = np.radians(deg)
        wi = np.array([np.sin(th), 0.0, -np.cos(th)])  # travelling into the surface
        for entering in (True, False):
            d = glass(n, wi, eta, entering)
            if eta * np.sin(th) > 1.0 + 1e-6:
                np.testing.assert_allclose(d, [np.sin(th), 0.0, np.cos(th)], atol=2e-6)  # mirror
                continue
            if eta * np.sin(th) > 1.0 - 1e-6:
                continue
            ratio = 1.0 / eta if entering else eta
            assert np.linalg.norm(d) == pytest.approx(1.0, abs=2e-6)
            assert d[0] == pytest.approx(ratio * np.sin(th), abs=2e-6)  # sin(theta_t) = ratio sin(theta_i)
            assert d[1] == pytest.approx(0.0, abs=1e-7)
            assert d[2] < 0.0  # transmitted through the surface
