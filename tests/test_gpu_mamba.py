"""The editor's G-buffer ReSTIR DI ("mamba" attachment) on the GPU against the
oracle's restatement (editor/optix/mamba_shader.cu:35-505, mamba.cu:1040-1100),
bit for bit: the colour buffer and both reservoir buffers after every frame,
for every combination of the temporal and spatial passes, a camera that
moves (reprojection through the previous projection * view), a scene with
several area lights and sky pixels, and glossy/glass materials."""
import numpy as np
import pytest

import oracle as O
from artemis_amd import ArmadaRTX, Mamba, look_at, scenes
from tests.test_gpu_motion import moving_cameras
from tests.test_oracle_mamba import split_light_scene

pytestmark = pytest.mark.gpu

FIELDS = ["pos", "pos", "pos", "n", "n", "n", "Le", "Le", "Le", "area", "w", "W", "M", "seed", "seed", "seed"]


def compare(arm, om, tag):
    col = arm.color_buffer()
    ok = np.all(col.view(np.uint32) == om.color.view(np.uint32), axis=-1)
    print(f"  {tag} color exact {np.mean(ok):.4f}", flush=True)
    if not ok.all():
        y, x = np.argwhere(~ok)[0]
        print(f"    ({y},{x}) gpu {col[y, x]} oracle {om.color[y, x]}", flush=True)
    bad = {}
    for which in (0, 1):
        g, r = arm.reservoirs(which), om.reservoirs(which)
        diff = np.any(g.view(np.uint32) != r.view(np.uint32), axis=(0, 1))
        names = [FIELDS[k] + str(k) for k in range(16) if diff[k]]
        if names:
            px = np.argwhere(np.any(g.view(np.uint32) != r.view(np.uint32), axis=-1))[0]
            print(f"    buffer {which} fields {names} first {px.tolist()}\n      gpu {g[tuple(px)].tolist()}"
                  f"\n      ora {r[tuple(px)].tolist()}", flush=True)
            bad[which] = names
    assert ok.all(), f"{tag}: colour differs in {np.mean(~ok):.4f} of pixels"
    assert not bad, f"{tag}: reservoir fields differ {bad}"


@pytest.fixture(scope="module")
def cornell_pair(krt_ctx):
    s = scenes.cornell()
    sc = krt_ctx.upload(s)
    osc = O.OracleScene(s, sc.serialized())
    yield s, sc, osc
    sc.free()


def mamba_armada(ctx, W, H, temporal, spatial):
    arm = ArmadaRTX(ctx, W, H)
    arm.attach("mamba", Mamba(ctx))
    arm.set_option("temporal", temporal)
    arm.set_option("spatial", spatial)
    assert arm.get_option("temporal") == temporal and arm.get_option("spatial") == spatial
    return arm


@pytest.mark.parametrize("temporal,spatial", [(False, False), (True, False), (False, True), (True, True)])
def test_static_camera(krt_ctx, cornell_pair, temporal, spatial):
    s, sc, osc = cornell_pair
    W, H = 48, 40
    cam = look_at(**s.camera, aspect=W / H)
    arm = mamba_armada(krt_ctx, W, H, temporal, spatial)
    om = O.OracleMamba(osc, W, H)
    print(f"[mamba static] {W}x{H} temporal={temporal} spatial={spatial}", flush=True)
    for f in range(3):
        arm.render(sc, cam)
        om.frame(cam, temporal=temporal, spatial=spatial)
        compare(arm, om, f"frame {f}")


def test_moving_camera(krt_ctx, cornell_pair):
    s, sc, osc = cornell_pair
    W, H = 48, 48
    arm = mamba_armada(krt_ctx, W, H, True, True)
    om = O.OracleMamba(osc, W, H)
    print(f"[mamba moving] {W}x{H}", flush=True)
    for f, cam in enumerate(moving_cameras(s.camera, W, H, 5)):
        arm.render(sc, cam)
        om.frame(cam, temporal=True, spatial=True)
        compare(arm, om, f"frame {f}")


def test_no_accumulation(krt_ctx, cornell_pair):
    s, sc, osc = cornell_pair
    W, H = 32, 32
    cam = look_at(**s.camera, aspect=1.0)
    arm = mamba_armada(krt_ctx, W, H, True, True)
    om = O.OracleMamba(osc, W, H)
    print("[mamba accumulate=False]", flush=True)
    for f in range(3):
        arm.render(sc, cam, accumulate=False)
        om.frame(cam, temporal=True, spatial=True, accumulate=False)
        compare(arm, om, f"frame {f}")


def test_several_lights_and_sky(krt_ctx):
    data, _ = split_light_scene()
    sc = krt_ctx.upload(data)
    osc = O.OracleScene(data, sc.serialized())
    W, H = 24, 16
    cam = look_at((0.0, 2.0, 3.0), (0.0, 0.5, 0.0), fov=70.0, aspect=W / H)
    arm = mamba_armada(krt_ctx, W, H, True, True)
    om = O.OracleMamba(osc, W, H)
    assert om.light_count() == 2
    print("[mamba split lights]", flush=True)
    for f in range(4):
        arm.render(sc, cam)
        om.frame(cam, temporal=True, spatial=True)
        compare(arm, om, f"frame {f}")
    assert np.any(om.reservoirs(0)[..., 12] == 0)  # sky pixels
    assert np.any(om.reservoirs(0)[..., 12] > 0)
    sc.free()


def test_glossy_and_glass(krt_ctx):
    s = scenes.cornell_glossy()
    sc = krt_ctx.upload(s)
    osc = O.OracleScene(s, sc.serialized())
    W, H = 40, 40
    cam = look_at(**s.camera, aspect=1.0)
    arm = mamba_armada(krt_ctx, W, H, True, True)
    om = O.OracleMamba(osc, W, H)
    print("[mamba glossy]", flush=True)
    for f in range(3):
        arm.render(sc, cam)
        om.frame(cam, temporal=True, spatial=True)
        compare(arm, om, f"frame {f}")
    sc.free()
