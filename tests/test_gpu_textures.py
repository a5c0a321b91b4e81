"""Material textures and normal maps on the GPU (krt_scene_upload_textured;
calculate_material / calculate_normal, amadeus_common.cuh:74-187) against
the oracle, bit for bit, on a scene with every texture kind (diffuse
checker tiled with wrap addressing, noise at negative coordinates, normal
maps on walls and a sphere, specular, roughness and emission textures):

* the path tracer's colour and AOVs over several frames at depth 5;
* ReSTIR DI's reservoirs and colour under a moving camera (the G-buffer keeps
  each pixel's texture coordinates for the temporal and spatial passes);
* served requests, in the batched launch and on a lane;
* the triangle lights (textured emission: the material's constant);
* ReSTIR GI (its G-buffer keeps the primary hit's (u, v)) and the editor's
  ReSTIR (the diffuse texture only, no normal map, as its G-buffer and
  convert_material do), reservoirs and colour;
* invalid texture arrays are refused at upload."""
import numpy as np
import pytest

import oracle as O
from artemis_amd import ArmadaRTX, Mamba, PathTracer, ReSTIR, ReSTIRGI, Server, look_at, scenes
from artemis_amd._lib import KRT_ERR_INVALID, KrtError

pytestmark = pytest.mark.gpu


def exact(a, b):
    return np.all(np.asarray(a, np.float32).view(np.uint32) == np.asarray(b, np.float32).view(np.uint32), axis=-1)


@pytest.fixture(scope="module")
def room(krt_ctx):
    s = scenes.textured_room()
    sc = krt_ctx.upload(s)
    osc = O.OracleScene(s, sc.serialized())
    yield s, sc, osc
    sc.free()


def test_lights_include_emission_textures(room):
    s, sc, osc = room
    got, want = sc.lights(), osc.lights()
    assert len(got) == len(want) == 4
    np.testing.assert_array_equal(got, want)


def test_path_tracer_frames(krt_ctx, room):
    s, sc, osc = room
    W, H = 64, 48
    cam = look_at(**s.camera, aspect=W / H)
    arm = ArmadaRTX(krt_ctx, W, H)
    arm.attach("pt", PathTracer(krt_ctx))
    arm.set_depth(5)
    print(f"[textures] path tracer {W}x{H} depth 5, 3 frames", flush=True)
    for _ in range(3):
        arm.render(sc, cam)
    ref, _ = osc.render_pt(cam, W, H, frames=3, max_depth=5)
    for k in range(4):
        ok = exact(arm._read(k), ref[k])
        assert ok.all(), f"buffer {k}: {ok.mean():.4f} bit-exact"
    # the textures change the image: the same scene without them differs
    plain = scenes.SceneData(s.positions, s.normals, s.indices, s.tri_material, s.materials, camera=s.camera)
    ref0, _ = O.OracleScene(plain, sc.serialized()).render_pt(cam, W, H, frames=3, max_depth=5)
    assert not exact(ref0[2], ref[2]).all() and not exact(ref0[1], ref[1]).all()


def test_restir_moving_camera(krt_ctx, room):
    s, sc, osc = room
    W, H = 48, 32
    arm = ArmadaRTX(krt_ctx, W, H)
    arm.attach("restir", ReSTIR(krt_ctx))
    ors = O.OracleReSTIR(W, H)
    print(f"[textures] ReSTIR DI {W}x{H}, 4 frames of a moving camera", flush=True)
    for f in range(4):
        eye = (0.15 * f - 0.2, 1.0 + 0.05 * f, 3.5 - 0.1 * f)
        cam = look_at(eye, (0.0, 0.9, 0.0), fov=45.0, aspect=W / H)
        arm.render(sc, cam, accumulate=False)
        ors.frame(osc, cam, max_depth=2, reset=True)
        for which in (0, 1):
            ok = exact(arm.reservoirs(which), ors.reservoirs(which))
            assert ok.all(), f"frame {f} reservoirs[{which}]: {ok.mean():.4f} bit-exact"
        ok = exact(arm.color_buffer(), ors.buffers[0])
        assert ok.all(), f"frame {f} colour: {ok.mean():.4f} bit-exact"


@pytest.mark.parametrize("wave_pixels", [0, 1 << 40])
def test_served_request(krt_ctx, room, wave_pixels):
    s, sc, osc = room
    w, h, spp, depth, seed = 40, 30, 3, 4, 5
    cam = look_at(**s.camera, aspect=w / h)
    srv = Server(krt_ctx, wave_pixels=wave_pixels)
    img, _ = srv.wait(srv.submit(sc, cam, w, h, spp=spp, max_depth=depth, seed=seed))
    srv.close()
    ref, _ = osc.render_pt(cam, w, h, frames=spp, max_depth=depth, user_seed=seed)
    ok = exact(img, ref[0])
    assert ok.all(), f"{ok.mean():.4f} bit-exact"


def test_cached_path_tracer_runs(krt_ctx, room):
    s, sc, _ = room
    W, H = 32, 32
    arm = ArmadaRTX(krt_ctx, W, H)
    arm.attach("pt", PathTracer(krt_ctx))
    arm.set_option("radiance_cache", True)
    arm.set_option("train_samples", 256)
    for _ in range(3):
        arm.render(sc, look_at(**s.camera, aspect=1.0))
    assert np.all(np.isfinite(arm.color_buffer()))


def test_restir_gi_frames(krt_ctx, room):
    """ReSTIR GI: the textured primary surface (its G-buffer keeps (u, v)),
    textured and normal-mapped secondary vertices; two frames, reservoirs
    and colour bit for bit."""
    s, sc, osc = room
    W, H = 40, 32
    cam = look_at(**s.camera, aspect=W / H)
    arm = ArmadaRTX(krt_ctx, W, H)
    arm.attach("gi", ReSTIRGI(krt_ctx))
    arm.set_option("spatial_passes", 2)
    arm.set_option("radius", 8)
    arm.set_depth(3)
    ogi = O.OracleGI(W, H)
    print(f"[textures] ReSTIR GI {W}x{H}, 2 frames", flush=True)
    for f in range(2):
        arm.render(sc, cam)
        ogi.frame(osc, cam, max_depth=3, m_cap=30, spatial_passes=2, neighbours=5, radius=8)
        for which in (0, 1):
            ok = exact(arm.reservoirs(which, gi=True), ogi.reservoirs(which))
            assert ok.all(), f"frame {f} reservoirs[{which}]: {ok.mean():.4f} bit-exact"
        ok = exact(arm.color_buffer(), ogi.color)
        assert ok.all(), f"frame {f} colour: {ok.mean():.4f} bit-exact"


def test_editor_restir(krt_ctx, room):
    """The editor's ReSTIR (mamba): its ray-cast G-buffer keeps (u, 1 - v)
    and no normal map, convert_material reads the diffuse texture only
    (editor/optix/common.cuh:34-47); three frames of a moving camera."""
    s, sc, osc = room
    W, H = 40, 32
    arm = ArmadaRTX(krt_ctx, W, H)
    arm.attach("mamba", Mamba(krt_ctx))
    om = O.OracleMamba(osc, W, H)
    print(f"[textures] editor ReSTIR {W}x{H}, 3 frames", flush=True)
    for f in range(3):
        cam = look_at((0.2 * f - 0.2, 1.0, 3.5), (0.0, 0.9, 0.0), aspect=W / H)
        arm.render(sc, cam)
        om.frame(cam, temporal=True, spatial=True)
        for which in (0, 1):
            ok = exact(arm.reservoirs(which), om.reservoirs(which))
            assert ok.all(), f"frame {f} reservoirs[{which}]: {ok.mean():.4f} bit-exact"
        ok = exact(arm.color_buffer(), om.color)
        assert ok.all(), f"frame {f} colour: {ok.mean():.4f} bit-exact"


def test_invalid_texture_arrays_are_refused(krt_ctx):
    base = scenes.textured_room()
    bad_index = scenes.SceneData(base.positions, base.normals, base.indices, base.tri_material, base.materials)
    bad_index.uvs, bad_index.tangents, bad_index.textures = base.uvs, base.tangents, base.textures
    mt = base.material_textures.copy()
    mt[0, 0] = len(base.textures)
    bad_index.material_textures = mt
    no_tangents = scenes.SceneData(base.positions, base.normals, base.indices, base.tri_material, base.materials)
    no_tangents.uvs, no_tangents.tangents, no_tangents.textures = base.uvs, None, base.textures
    no_tangents.material_textures = base.material_textures
    for d in (bad_index, no_tangents):
        with pytest.raises(KrtError) as e:
            krt_ctx.upload(d, build=False)
        assert e.value.code == KRT_ERR_INVALID
