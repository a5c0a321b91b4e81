"""The armada's environment map on the GPU (ArmadaRTX::set_envmap,
armada.cu:86-96) against the oracle, bit for bit: the path tracer's misses
(read at (u, 1 - v)) and next-event estimation with the environment as one
more light (Ld_Environment), ReSTIR DI's misses (at (u, v)) and environment
candidates (sample_direct type 2) through the temporal and spatial passes;
ReSTIR GI refuses a frame while the map is on."""
import numpy as np
import pytest

import oracle as O
from artemis_amd import ArmadaRTX, PathTracer, ReSTIR, ReSTIRGI, look_at, scenes
from artemis_amd._lib import KRT_ERR_STATE, KrtError

pytestmark = pytest.mark.gpu


def exact(a, b):
    return np.all(np.asarray(a, np.float32).view(np.uint32) == np.asarray(b, np.float32).view(np.uint32), axis=-1)


@pytest.fixture(scope="module")
def sky(krt_ctx):
    s = scenes.sky_room()
    sc = krt_ctx.upload(s)
    osc = O.OracleScene(s, sc.serialized())
    env = np.random.default_rng(11).uniform(0.0, 3.0, (32, 64, 4)).astype(np.float32)
    osc.set_envmap(env)
    yield s, sc, osc, env
    sc.free()


def test_path_tracer_with_envmap(krt_ctx, sky):
    s, sc, osc, env = sky
    W, H = 64, 48
    cam = look_at(**s.camera, aspect=W / H)
    arm = ArmadaRTX(krt_ctx, W, H)
    arm.set_envmap(env)
    arm.attach("pt", PathTracer(krt_ctx))
    arm.set_depth(5)
    print(f"[envmap] path tracer {W}x{H} depth 5, 3 frames", flush=True)
    for _ in range(3):
        arm.render(sc, cam)
    ref, _ = osc.render_pt(cam, W, H, frames=3, max_depth=5)
    for k in range(4):
        ok = exact(arm._read(k), ref[k])
        assert ok.all(), f"buffer {k}: {ok.mean():.4f} bit-exact"
    # the map changed the image: without it the same frames differ
    plain = O.OracleScene(s, sc.serialized())
    ref0, _ = plain.render_pt(cam, W, H, frames=3, max_depth=5)
    assert not np.array_equal(ref0[0], ref[0])


def test_restir_with_envmap_moving_camera(krt_ctx, sky):
    s, sc, osc, env = sky
    W, H = 48, 32
    arm = ArmadaRTX(krt_ctx, W, H)
    arm.set_envmap(env)
    arm.attach("restir", ReSTIR(krt_ctx))
    ors = O.OracleReSTIR(W, H)
    print(f"[envmap] ReSTIR DI {W}x{H}, 4 frames of a moving camera", flush=True)
    for f in range(4):
        eye = (0.25 * f, 3.0, 4.0 - 0.1 * f)
        cam = look_at(eye, (0.0, 0.5, 0.0), fov=50.0, aspect=W / H)
        arm.render(sc, cam, accumulate=False)
        ors.frame(osc, cam, max_depth=2, reset=True)
        for which in (0, 1):
            ok = exact(arm.reservoirs(which), ors.reservoirs(which))
            assert ok.all(), f"frame {f} reservoirs[{which}]: {ok.mean():.4f} bit-exact"
        ok = exact(arm.color_buffer(), ors.buffers[0])
        assert ok.all(), f"frame {f} colour: {ok.mean():.4f} bit-exact"
    types = arm.reservoirs(0)[..., 9]
    assert np.any(types == 2.0), "no environment sample was selected"


def test_gi_refuses_envmap(krt_ctx, sky):
    s, sc, _, env = sky
    arm = ArmadaRTX(krt_ctx, 16, 16)
    arm.set_envmap(env)
    arm.attach("gi", ReSTIRGI(krt_ctx))
    cam = look_at(**s.camera)
    with pytest.raises(KrtError) as e:
        arm.render(sc, cam)
    assert e.value.code == KRT_ERR_STATE
    arm.set_envmap_enabled(False)
    arm.render(sc, cam)
    assert np.all(np.isfinite(arm.color_buffer()))


def test_cached_path_tracer_with_envmap(krt_ctx, sky):
    s, sc, _, env = sky
    W, H = 32, 32
    cam = look_at(**s.camera, aspect=W / H)
    arm = ArmadaRTX(krt_ctx, W, H)
    arm.set_envmap(env)
    arm.attach("pt", PathTracer(krt_ctx))
    arm.set_option("radiance_cache", True)
    arm.set_option("train_samples", 256)
    for _ in range(4):
        arm.render(sc, cam)
    assert np.all(np.isfinite(arm.color_buffer()))
