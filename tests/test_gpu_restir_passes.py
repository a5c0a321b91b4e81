"""ReSTIR DI pass-by-pass parity: reservoirs and AOVs after each frame."""
import numpy as np
import pytest

import oracle as O
from artemis_amd import ArmadaRTX, ReSTIR, look_at, scenes

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("passes,mcap,frames", [(1, 20, 1), (2, 200, 1), (2, 200, 2), (2, 20, 2), (3, 30, 2)])
def test_reservoirs_per_frame(krt_ctx, passes, mcap, frames):
    s = scenes.cornell()
    sc = krt_ctx.upload(s)
    osc = O.OracleScene(s, sc.serialized())
    W = H = 32
    cam = look_at(**s.camera, aspect=1.0)
    arm = ArmadaRTX(krt_ctx, W, H)
    arm.attach("restir", ReSTIR(krt_ctx))
    arm.set_option("spatial_passes", passes)
    arm.set_option("m_cap", mcap)
    ors = O.OracleReSTIR(W, H)
    print(f"[restir passes] passes={passes} mcap={mcap} frames={frames}", flush=True)
    for f in range(frames):
        arm.render(sc, cam)
        ors.frame(osc, cam, max_depth=2, spatial_passes=passes, m_cap=mcap)
    names = ["Le", "Le", "Le", "n", "n", "n", "p", "p", "p", "type", "w", "W", "M", "seed", "seed", "seed"]
    for which in (0, 1):
        g, r = arm.reservoirs(which), ors.reservoirs(which)
        diff = np.any(g.view(np.uint32) != r.view(np.uint32), axis=(0, 1))
        bad = [names[k] + str(k) for k in range(16) if diff[k]]
        frac = np.mean(np.any(g.view(np.uint32) != r.view(np.uint32), axis=2))
        print(f"  buffer {which}: differing fields {bad}, pixels {frac:.4f}")
        assert not bad, f"buffer {which} fields {bad} differ in {frac:.4f} of pixels"
    col = arm.color_buffer()
    ex = np.mean(np.all(col.view(np.uint32) == ors.buffers[0].view(np.uint32), axis=-1))
    print(f"  color exact {ex:.4f}")
    assert ex == 1.0
    sc.free()


def test_gbuffer_and_reservoirs_frame_by_frame(krt_ctx):
    """Locates the first frame and stage at which GPU and oracle part."""
    s = scenes.cornell()
    sc = krt_ctx.upload(s)
    osc = O.OracleScene(s, sc.serialized())
    W = H = 32
    cam = look_at(**s.camera, aspect=1.0)
    rays = O.camera_rays(cam, W, H, jitter=False)
    ref_hits, _ = osc.trace(rays, mode=0)
    gpu_hits = sc.trace_closest(rays)
    nbad = int(np.sum(np.any(ref_hits.view(np.uint32) != gpu_hits.view(np.uint32), axis=1)))
    print(f"  primary rays: {nbad} differing hits", flush=True)
    arm = ArmadaRTX(krt_ctx, W, H)
    arm.attach("restir", ReSTIR(krt_ctx))
    arm.set_option("spatial_passes", 2)
    arm.set_option("m_cap", 200)
    ors = O.OracleReSTIR(W, H)
    first_bad = None
    for f in range(3):
        arm.render(sc, cam)
        ors.frame(osc, cam, max_depth=2, spatial_passes=2, m_cap=200)
        for name, g, r in zip(["normal", "albedo", "position", "color"],
                              [arm.normal_buffer(), arm.albedo_buffer(), arm.position_buffer(), arm.color_buffer()],
                              [ors.buffers[1], ors.buffers[2], ors.buffers[3], ors.buffers[0]]):
            bad = np.argwhere(np.any(g.view(np.uint32) != r.view(np.uint32), axis=-1))
            print(f"  frame {f} {name}: {len(bad)} px differ {bad[:4].tolist()}", flush=True)
            if len(bad) and name != "color":
                y, x = bad[0]
                print(f"    gpu {g[y, x]} oracle {r[y, x]}", flush=True)
        for which in (0, 1):
            g, r = arm.reservoirs(which), ors.reservoirs(which)
            bad = np.argwhere(np.any(g.view(np.uint32) != r.view(np.uint32), axis=-1))
            print(f"  frame {f} reservoirs[{which}]: {len(bad)} px differ {bad[:4].tolist()}", flush=True)
            if len(bad):
                y, x = bad[0]
                print(f"    gpu {g[y, x].tolist()}\n    ora {r[y, x].tolist()}", flush=True)
                first_bad = first_bad or (f, which)
    assert nbad == 0
    assert first_bad is None, first_bad
    sc.free()
