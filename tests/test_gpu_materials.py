"""Every BSDF branch on the GPU against the oracle, bit for bit: glossy
surfaces (Ks > 0: the GGX lobe of brdf.cuh:386-433 is sampled with
probability Ks / (Kd + Ks)), a near-mirror box and a glass sphere
(eTransmission: Glass::sample, brdf.cuh:125-176, and the inward offset of
amadeus_path_tracer.cu:143).  The scene is scenes.cornell_glossy()."""
import numpy as np
import pytest

import oracle as O
from artemis_amd import ArmadaRTX, PathTracer, ReSTIR, ReSTIRGI, look_at, scenes

pytestmark = pytest.mark.gpu


def exact(a, b):
    return np.all(a.view(np.uint32) == b.view(np.uint32), axis=-1)


@pytest.fixture(scope="module")
def glossy_pair(krt_ctx):
    s = scenes.cornell_glossy()
    sc = krt_ctx.upload(s)
    osc = O.OracleScene(s, sc.serialized())
    yield s, sc, osc
    sc.free()


def test_scene_uses_every_branch(glossy_pair):
    s, _, _ = glossy_pair
    m = s.materials
    assert (m[:, 3:6].mean(axis=1) > 0).sum() >= 7  # glossy lobes
    assert (m[:, 11] == 2.0).sum() == 1  # one glass material


def test_pt_glossy_glass_frames_bit_exact(krt_ctx, glossy_pair):
    s, sc, osc = glossy_pair
    W = H = 64
    frames = 4
    cam = look_at(**s.camera, aspect=W / H)
    arm = ArmadaRTX(krt_ctx, W, H)
    arm.attach("path_tracer", PathTracer(krt_ctx))
    arm.set_depth(5)
    print(f"[materials] path tracer {W}x{H} frames={frames}", flush=True)
    for _ in range(frames):
        arm.render(sc, cam, accumulate=True)
    ref, _ = osc.render_pt(cam, W, H, frames=frames, max_depth=5)
    got = [arm.color_buffer(), arm.normal_buffer(), arm.albedo_buffer(), arm.position_buffer()]
    for name, g, r in zip(["color", "normal", "albedo", "position"], got, ref):
        ok = exact(g, r)
        print(f"  {name}: {ok.mean():.4f} of pixels bit-exact", flush=True)
        assert ok.all(), f"{name}: first differing pixels {np.argwhere(~ok)[:4].tolist()}"


def test_restir_glossy_glass_reservoirs_bit_exact(krt_ctx, glossy_pair):
    s, sc, osc = glossy_pair
    W = H = 32
    cam = look_at(**s.camera, aspect=W / H)
    arm = ArmadaRTX(krt_ctx, W, H)
    arm.attach("restir", ReSTIR(krt_ctx))
    arm.set_option("spatial_passes", 2)
    arm.set_option("m_cap", 20)
    ors = O.OracleReSTIR(W, H)
    print(f"[materials] ReSTIR DI {W}x{H} frames=3", flush=True)
    for _ in range(3):
        arm.render(sc, cam)
        ors.frame(osc, cam, max_depth=2, spatial_passes=2, m_cap=20)
    for which in (0, 1):
        ok = exact(arm.reservoirs(which), ors.reservoirs(which))
        print(f"  reservoirs[{which}]: {ok.mean():.4f} bit-exact", flush=True)
        assert ok.all()
    ok = exact(arm.color_buffer(), ors.buffers[0])
    print(f"  color: {ok.mean():.4f} bit-exact", flush=True)
    assert ok.all()


def test_gi_glossy_glass_bit_exact(krt_ctx, glossy_pair):
    s, sc, osc = glossy_pair
    W, H = 40, 32
    cam = look_at(**s.camera, aspect=W / H)
    arm = ArmadaRTX(krt_ctx, W, H)
    arm.attach("gi", ReSTIRGI(krt_ctx))
    arm.set_option("spatial_passes", 2)
    arm.set_option("radius", 8)
    arm.set_depth(3)
    ogi = O.OracleGI(W, H)
    print(f"[materials] ReSTIR GI {W}x{H} frames=2", flush=True)
    for _ in range(2):
        arm.render(sc, cam)
        ogi.frame(osc, cam, max_depth=3, m_cap=30, spatial_passes=2, neighbours=5, radius=8)
    for which in (0, 1):
        ok = exact(arm.reservoirs(which, gi=True), ogi.reservoirs(which))
        print(f"  reservoirs[{which}]: {ok.mean():.4f} bit-exact", flush=True)
        assert ok.all()
    ok = exact(arm.color_buffer(), ogi.color)
    print(f"  color: {ok.mean():.4f} bit-exact", flush=True)
    assert ok.all()
