"""ReSTIR GI on the GPU vs the oracle restatement (same seeds, same order of
operations): reservoirs and images per frame, bit for bit."""
import numpy as np
import pytest

import oracle as O
from artemis_amd import ArmadaRTX, ReSTIRGI, look_at, scenes

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("scene,passes,frames", [("sky_room", 2, 3), ("cornell", 2, 2), ("cornell", 0, 2)])
def test_gi_frames_match_oracle(krt_ctx, scene, passes, frames):
    s = scenes.sky_room() if scene == "sky_room" else scenes.cornell()
    sc = krt_ctx.upload(s)
    osc = O.OracleScene(s, sc.serialized())
    W, H = 40, 32
    cam = look_at(**s.camera, aspect=W / H)
    arm = ArmadaRTX(krt_ctx, W, H)
    arm.attach("gi", ReSTIRGI(krt_ctx))
    arm.set_option("spatial_passes", passes)
    arm.set_option("radius", 8)
    arm.set_depth(3)
    ogi = O.OracleGI(W, H)
    print(f"[gi] {scene} {W}x{H} passes={passes} frames={frames}", flush=True)
    for f in range(frames):
        arm.render(sc, cam)
        ogi.frame(osc, cam, max_depth=3, m_cap=30, spatial_passes=passes, neighbours=5, radius=8)
    g, r = arm.color_buffer(), ogi.color
    mad = float(np.abs(g.astype(np.float64) - r).mean())
    ex = float(np.mean(np.all(g.view(np.uint32) == r.view(np.uint32), axis=-1)))
    print(f"  color MAD={mad:.3e} exact={ex:.4f}")
    for which in (0, 1):
        gr, rr = arm.reservoirs(which, gi=True), ogi.reservoirs(which)
        frac = float(np.mean(np.any(gr.view(np.uint32) != rr.view(np.uint32), axis=-1)))
        print(f"  reservoirs[{which}] differing pixels {frac:.4f}")
        assert frac == 0.0, f"reservoirs[{which}]"
    assert mad <= 1e-3
    assert ex == 1.0
    sc.free()
