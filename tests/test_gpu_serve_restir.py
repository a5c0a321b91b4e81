"""The server with ReSTIR DI requests and against the oracle.

* a ReSTIR DI request's answer is the restir attachment's colour buffer after
  spp frames, bit for bit, and the oracle's (OracleReSTIR) too; its
  reservoirs stay resident on the device across the steps of its life;
* path-tracer requests equal the oracle's render_pt image bit for bit;
* a mix at the serving configuration's scale -- a 1080p request on the
  1,048,576-triangle scene beside small path-tracer and ReSTIR requests on
  other scenes -- gives every request the result it gets alone;
* while a request holds a scene, the scene can be neither freed nor rebuilt
  (KRT_ERR_STATE)."""
import numpy as np
import pytest

import oracle as O
from artemis_amd import ArmadaRTX, PathTracer, ReSTIR, Server, look_at, scenes
from artemis_amd._lib import KRT_ERR_STATE, KrtError

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


@pytest.fixture(scope="module")
def cornell(krt_ctx):
    s = scenes.cornell()
    sc = krt_ctx.upload(s)
    yield s, sc, O.OracleScene(s, sc.serialized())
    sc.free()


RESTIR = dict(candidates=32, m_cap=20, spatial_passes=1, neighbours=5, radius=30)


def restir_attachment(ctx, sc, data, w, h, spp, depth, seed, opts=RESTIR):
    arm = ArmadaRTX(ctx, w, h)
    arm.attach("restir", ReSTIR(ctx))
    for k, v in opts.items():
        arm.set_option(k, v)
    arm.set_option("seed", seed)
    arm.set_depth(depth)
    cam = look_at(**data.camera, aspect=w / h)
    for _ in range(spp):
        arm.render(sc, cam, accumulate=True)
    return arm.color_buffer()


def test_restir_request_equals_attachment_and_oracle(krt_ctx, cornell):
    data, sc, osc = cornell
    w, h, spp, depth, seed = 40, 32, 4, 2, 0
    cam = look_at(**data.camera, aspect=w / h)
    srv = Server(krt_ctx, step_pixels=1)  # one frame per step: the reservoirs live across steps
    print(f"[serve] restir request {w}x{h} spp={spp}", flush=True)
    rid = srv.submit(sc, cam, w, h, spp=spp, max_depth=depth, seed=seed, kind="restir", **RESTIR)
    img, _ = srv.wait(rid)
    st = srv.stats()
    srv.close()
    assert st["steps"] >= spp
    assert np.array_equal(bits(img), bits(restir_attachment(krt_ctx, sc, data, w, h, spp, depth, seed)))
    ors = O.OracleReSTIR(w, h, user_seed=seed)
    for _ in range(spp):
        ors.frame(osc, cam, max_depth=depth, **RESTIR)
    same = np.all(bits(img) == bits(ors.buffers[0]), axis=-1)
    print(f"  {same.mean():.4f} of pixels bit-exact vs the oracle", flush=True)
    assert same.all()


def test_path_tracer_request_equals_oracle(krt_ctx, cornell):
    data, sc, osc = cornell
    w, h, spp, depth, seed = 48, 36, 3, 5, 9
    cam = look_at(**data.camera, aspect=w / h)
    srv = Server(krt_ctx)
    rid = srv.submit(sc, cam, w, h, spp=spp, max_depth=depth, seed=seed)
    img, _ = srv.wait(rid)
    srv.close()
    ref, _ = osc.render_pt(cam, w, h, frames=spp, max_depth=depth, user_seed=seed)
    assert np.array_equal(bits(img), bits(ref[0]))


def test_config_scale_mix_matches_alone(krt_ctx, cornell):
    """A 1920x1080 ReSTIR DI request and a 1920x1080 path-tracer request on
    the 1M-triangle scene among small requests on the Cornell box and a
    50k-triangle scene: every result equals the same request served alone;
    1080p path-tracer pixels are spot-checked against the oracle."""
    data, sc, osc = cornell
    big = scenes.spheres()
    print(f"[serve] uploading {big.triangles} triangles", flush=True)
    sbig = krt_ctx.upload(big)
    mid = scenes.spheres(instances=12, slices=64, stacks=33, emissive=196, seed=11)
    smid = krt_ctx.upload(mid)
    reqs = [
        (sbig, big, 1920, 1080, 2, 2, 1, "restir"),
        (sbig, big, 1920, 1080, 2, 5, 2, "path_tracer"),
        (sc, data, 64, 64, 8, 2, 3, "restir"),
        (sc, data, 96, 54, 16, 5, 4, "path_tracer"),
        (smid, mid, 256, 256, 4, 2, 5, "restir"),
        (smid, mid, 512, 512, 1, 5, 6, "path_tracer"),
        (sc, data, 40, 40, 3, 2, 7, "restir"),
    ]

    def submit(srv, r):
        s_, d_, w, h, spp, depth, seed, kind = r
        kw = RESTIR if kind == "restir" else {}
        return srv.submit(s_, look_at(**d_.camera, aspect=w / h), w, h, spp=spp, max_depth=depth, seed=seed,
                          kind=kind, **kw)

    srv = Server(krt_ctx, max_active=16, hbm_budget=64 << 30, step_pixels=1 << 21)
    ids = [submit(srv, r) for r in reqs]
    mixed = [srv.wait(i)[0] for i in ids]
    st = srv.stats()
    srv.close()
    print(f"  mix: {st['completed']} done, max concurrent {st['max_concurrent']}, steps {st['steps']}", flush=True)
    assert st["completed"] == len(reqs) and st["max_concurrent"] > 1
    for r, img in zip(reqs, mixed):
        srv1 = Server(krt_ctx, max_active=1, hbm_budget=64 << 30, step_pixels=1 << 21)
        alone, _ = srv1.wait(submit(srv1, r))
        srv1.close()
        assert np.array_equal(bits(img), bits(alone)), r[2:]
        assert np.isfinite(img).all()
    # 1080p path tracer, 2 frames: spot pixels of the accumulated image vs the oracle
    w, h = 1920, 1080
    cam = look_at(**big.camera, aspect=w / h)
    obig = O.OracleScene(big, sbig.serialized())
    rng = np.random.default_rng(3)
    pix = np.unique(rng.integers(0, w * h, 1500)).astype(np.int32)
    f0 = obig.render_pt_pixels(cam, w, h, pix, sample=0, max_depth=5, user_seed=2)
    f1 = obig.render_pt_pixels(cam, w, h, pix, sample=1, max_depth=5, user_seed=2)
    acc = (f0 * np.float32(0.0) + f0) / np.float32(1.0)
    acc = (acc * np.float32(1.0) + f1) / np.float32(2.0)
    got = mixed[1].reshape(-1, 4)[pix]
    assert np.array_equal(bits(got), bits(acc))
    sbig.free()
    smid.free()


def test_scene_in_use_cannot_be_freed_or_rebuilt(krt_ctx):
    data = scenes.cornell()
    sc = krt_ctx.upload(data)
    srv = Server(krt_ctx, step_pixels=1)
    rid = srv.submit(sc, look_at(**data.camera), 64, 64, spp=64, max_depth=5)
    for call in (sc.free, sc.build):
        with pytest.raises(KrtError) as e:
            call()
        assert e.value.code == KRT_ERR_STATE
    img, _ = srv.wait(rid)
    srv.close()
    assert img is not None
    sc.build()
    sc.free()
