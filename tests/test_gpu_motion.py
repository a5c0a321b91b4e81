"""Temporal reuse under a moving camera.  With accumulate = False the
sample counter stays 0, so every frame's temporal pass reprojects the
surface point through the previous frame's projection * view
(amadeus_restir.cu:182-210; ReSTIR GI uses the same rule) instead of reading
the same pixel.  The GPU reservoirs and images are compared with the
oracle's bit for bit over several frames of a camera that pans and moves."""
import numpy as np
import pytest

import oracle as O
from artemis_amd import ArmadaRTX, ReSTIR, ReSTIRGI, look_at, scenes

pytestmark = pytest.mark.gpu


def exact(a, b):
    return np.all(a.view(np.uint32) == b.view(np.uint32), axis=-1)


def moving_cameras(camera, W, H, frames, px_per_frame=1.5, dolly=0.02):
    """Pans right by px_per_frame pixels and moves the eye forward by
    `dolly` of the look distance per frame."""
    eye = np.asarray(camera["eye"], np.float64)
    at = np.asarray(camera["lookat"], np.float64)
    up = np.asarray(camera.get("up", (0.0, 1.0, 0.0)), np.float64)
    fov = camera.get("fov", 45.0)
    w = at - eye
    right = np.cross(w, up)
    right /= np.linalg.norm(right)
    pitch = 2.0 * np.linalg.norm(w) * np.tan(np.radians(fov) / 2.0) / H
    out = []
    for f in range(frames):
        e = eye + dolly * f * w
        out.append(look_at(e, at + right * pitch * px_per_frame * f, up=tuple(up), fov=fov, aspect=W / H))
    return out


@pytest.fixture(scope="module")
def cornell_pair(krt_ctx):
    s = scenes.cornell()
    sc = krt_ctx.upload(s)
    osc = O.OracleScene(s, sc.serialized())
    yield s, sc, osc
    sc.free()


@pytest.mark.parametrize("mcap", [20, 200])
def test_restir_di_moving_camera(krt_ctx, cornell_pair, mcap):
    s, sc, osc = cornell_pair
    W = H = 48
    frames = 5
    arm = ArmadaRTX(krt_ctx, W, H)
    arm.attach("restir", ReSTIR(krt_ctx))
    arm.set_option("spatial_passes", 2)
    arm.set_option("m_cap", mcap)
    ors = O.OracleReSTIR(W, H)
    print(f"[motion] ReSTIR DI {W}x{H} frames={frames} mcap={mcap}", flush=True)
    for f, cam in enumerate(moving_cameras(s.camera, W, H, frames)):
        arm.render(sc, cam, accumulate=False)
        ors.frame(osc, cam, max_depth=2, spatial_passes=2, m_cap=mcap, reset=True)
        for which in (0, 1):
            ok = exact(arm.reservoirs(which), ors.reservoirs(which))
            assert ok.all(), f"frame {f} reservoirs[{which}]: {ok.mean():.4f} bit-exact"
        ok = exact(arm.color_buffer(), ors.buffers[0])
        print(f"  frame {f}: reservoirs bit-exact, color {ok.mean():.4f}", flush=True)
        assert ok.all()
    # the moving camera changed which reservoir a pixel merged with: the merged
    # sample counts differ from a static camera's
    st = ArmadaRTX(krt_ctx, W, H)
    st.attach("restir", ReSTIR(krt_ctx))
    st.set_option("spatial_passes", 2)
    st.set_option("m_cap", mcap)
    cam0 = moving_cameras(s.camera, W, H, 1)[0]
    for _ in range(frames):
        st.render(sc, cam0, accumulate=False)
    assert not np.array_equal(st.reservoirs(0)[..., 12], arm.reservoirs(0)[..., 12])


@pytest.mark.parametrize("scene", ["cornell", "sky_room"])
def test_restir_gi_moving_camera(krt_ctx, scene):
    s = scenes.cornell() if scene == "cornell" else scenes.sky_room()
    sc = krt_ctx.upload(s)
    osc = O.OracleScene(s, sc.serialized())
    W, H = 40, 32
    frames = 4
    arm = ArmadaRTX(krt_ctx, W, H)
    arm.attach("gi", ReSTIRGI(krt_ctx))
    arm.set_option("spatial_passes", 2)
    arm.set_option("radius", 8)
    arm.set_depth(3)
    ogi = O.OracleGI(W, H)
    print(f"[motion] ReSTIR GI {scene} {W}x{H} frames={frames}", flush=True)
    for f, cam in enumerate(moving_cameras(s.camera, W, H, frames)):
        arm.render(sc, cam, accumulate=False)
        ogi.frame(osc, cam, max_depth=3, m_cap=30, spatial_passes=2, neighbours=5, radius=8, accumulate=False)
        for which in (0, 1):
            ok = exact(arm.reservoirs(which, gi=True), ogi.reservoirs(which))
            assert ok.all(), f"frame {f} reservoirs[{which}]: {ok.mean():.4f} bit-exact"
        ok = exact(arm.color_buffer(), ogi.color)
        print(f"  frame {f}: reservoirs bit-exact, color {ok.mean():.4f}", flush=True)
        assert ok.all()
    sc.free()


def test_restir_gi_config_1080p_2m_panning(krt_ctx):
    """ReSTIR GI at 1920x1080 on the 2,097,152-triangle scene (256 x
    sphere(64, 65)), depth 3, M-cap 30, two spatial passes, the camera
    panning 1 px per frame (accumulate = False: reprojection every frame).
    The oracle cannot render such frames inside a test: two fresh runs give
    the same bits, and reservoirs and colours stay finite."""
    W, H = 1920, 1080
    data = scenes.spheres(instances=256, slices=64, stacks=65, emissive=4096)
    assert data.triangles == 2_097_152
    print(f"[motion] uploading {data.triangles} triangles", flush=True)
    sc = krt_ctx.upload(data)
    cams = moving_cameras(data.camera, W, H, 8, px_per_frame=1.0, dolly=0.0)

    def run():
        arm = ArmadaRTX(krt_ctx, W, H)
        arm.attach("restir_gi", ReSTIRGI(krt_ctx))
        arm.set_option("m_cap", 30)
        arm.set_option("spatial_passes", 2)
        arm.set_option("neighbours", 5)
        arm.set_option("radius", 30)
        arm.set_depth(3)
        for cam in cams:
            arm.render(sc, cam, accumulate=False)
        return arm

    a, b = run(), run()
    ca, cb = a.color_buffer(), b.color_buffer()
    assert np.array_equal(ca.view(np.uint32), cb.view(np.uint32))
    for which in (0, 1):
        ra, rb = a.reservoirs(which, gi=True), b.reservoirs(which, gi=True)
        assert np.array_equal(ra.view(np.uint32), rb.view(np.uint32))
        assert np.isfinite(ra).all()
    assert np.isfinite(ca).all()
    M = a.reservoirs(0, gi=True)[..., 14]
    print(f"  mean colour {ca[..., :3].mean():.4f}, M in [{M.min():.0f}, {M.max():.0f}]", flush=True)
    assert M.max() > 1  # temporal reuse through the reprojection carried samples over
    sc.free()
