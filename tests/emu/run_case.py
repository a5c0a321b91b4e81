"""Runs one case of the host sanitizer harness: the Python API over
libkrt_emu.so (the library built for the CPU under AddressSanitizer / UBSan,
tests/emu/build.py) instead of libkrt.so, checked against the oracle.

Usage (the harness test starts it with clang's ASan runtime preloaded):
    python tests/emu/run_case.py <case>
Each case prints what it checked and exits 0; a sanitizer report or a failed
check exits non-zero.
"""
from __future__ import annotations

import ctypes as C
import os
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = Path(os.environ.get("KRT_EMU_ROOT", HERE.parents[1]))  # the checkout whose package and oracle run
EMU_LIB = Path(os.environ.get("KRT_EMU_BUILD", HERE / "_build")) / "libkrt_emu.so"
sys.path.insert(0, str(ROOT))

import artemis_amd._lib as L  # noqa: E402

_emu = C.CDLL(str(EMU_LIB))
for _name, (_res, _args) in L._SIGS.items():
    _f = getattr(_emu, _name)
    _f.restype = _res
    _f.argtypes = _args
L._lib = _emu  # every wrapper below now calls the emulated library

import oracle as O  # noqa: E402
from artemis_amd import ArmadaRTX, Context, PathTracer, ReSTIR, ReSTIRGI, look_at, scenes  # noqa: E402
from artemis_amd._lib import KRT_ERR_STATE, KrtError  # noqa: E402


def say(msg):
    print(msg, flush=True)


def bits_equal(a, b):
    return np.array_equal(np.asarray(a, np.float32).view(np.uint32), np.asarray(b, np.float32).view(np.uint32))


def case_flat_pt(ctx):
    """A path-traced Cornell frame: the emulated kernels give the oracle's bits."""
    s = scenes.cornell()
    sc = ctx.upload(s)
    nodes = sc.serialized()
    assert bits_equal(nodes, O.bvh_build(O.tri_boxes(s.positions, s.indices))), "BVH"
    osc = O.OracleScene(s, nodes)
    W = H = 16
    cam = look_at(**s.camera)
    arm = ArmadaRTX(ctx, W, H)
    arm.attach("pt", PathTracer(ctx))
    arm.set_depth(3)
    arm.render(sc, cam)
    ref, _ = osc.render_pt(cam, W, H, frames=1, max_depth=3)
    assert bits_equal(arm.color_buffer(), ref[0]), "path tracer colour"
    say("flat_pt: BVH and 16x16 depth-3 frame bit-exact vs oracle")
    sc.free()


def case_instancing(ctx, full=False):
    """The sequence of tests/test_gpu_instancing.py on the two-level path
    (full: at that file's sizes, 20,000 rays, 48x40 and 40x32 frames)."""
    from tests.test_gpu_instancing import rays_at_instances

    data = scenes.spheres(instances=24, slices=16, stacks=9, emissive=64, box=8.0, base_meshes=4, instanced=True)
    sc = ctx.upload(data)
    osc = O.OracleScene(data)
    for m, (pos, _, idx) in enumerate(data.meshes):
        assert bits_equal(sc.mesh_serialized(m), O.bvh_build(O.tri_boxes(pos, idx))), f"mesh tree {m}"
    assert bits_equal(sc.instance_tree(), osc.tlas()), "instance tree"
    say("instancing: mesh and instance trees bit-exact")
    rays = rays_at_instances(data, 20000 if full else 4000, 5)
    want, _ = osc.trace(rays, mode=0)
    got = sc.trace_closest(rays)
    assert bits_equal(got, want), "closest hits"
    occ = sc.trace_occluded(rays)
    assert np.array_equal(occ, osc.occluded(rays)), "occlusion"
    say(f"instancing: {len(rays):,} closest hits and occlusion queries bit-exact")
    got_l, want_l = sc.lights(), osc.lights()
    assert bits_equal(got_l, want_l), "lights"
    for layout in ("sah", "reference"):
        sc.build(layout=layout)
        W, H = (48, 40) if full else (24, 20)
        frames = 3 if full else 2
        cam = look_at(**data.camera, aspect=W / H)
        arm = ArmadaRTX(ctx, W, H)
        arm.attach("pt", PathTracer(ctx))
        arm.set_depth(5)
        for _ in range(frames):
            arm.render(sc, cam)
        ref, _ = osc.render_pt(cam, W, H, frames=frames, max_depth=5)
        assert bits_equal(arm.color_buffer(), ref[0]), f"path tracer [{layout}]"
        for k in (1, 2, 3):
            assert bits_equal(arm._read(k), ref[k]), f"AOV {k} [{layout}]"
        say(f"instancing: path tracer frames bit-exact [{layout}]")
        del arm
    sc.build(layout="sah")
    W, H = (40, 32) if full else (20, 16)
    cam = look_at(**data.camera, aspect=W / H)
    arm = ArmadaRTX(ctx, W, H)
    arm.attach("restir", ReSTIR(ctx))
    arm.set_option("m_cap", 20)
    arm.set_option("spatial_passes", 1)
    ors = O.OracleReSTIR(W, H)
    for _ in range(2):
        arm.render(sc, cam)
        ors.frame(osc, cam, max_depth=2, m_cap=20, spatial_passes=1)
    for which in (0, 1):
        assert bits_equal(arm.reservoirs(which), ors.reservoirs(which)), f"reservoirs {which}"
    say("instancing: ReSTIR DI reservoirs bit-exact")
    del arm
    # the flat-only attachments refuse the scene with KRT_ERR_STATE
    cam = look_at(**data.camera)
    arm = ArmadaRTX(ctx, 16, 16)
    arm.attach("gi", ReSTIRGI(ctx))
    try:
        arm.render(sc, cam)
        raise AssertionError("restir_gi rendered an instanced scene")
    except KrtError as e:
        assert e.code == KRT_ERR_STATE, e
    arm2 = ArmadaRTX(ctx, 16, 16)
    arm2.attach("pt", PathTracer(ctx))
    arm2.set_option("radiance_cache", True)
    try:
        arm2.render(sc, cam)
        raise AssertionError("the radiance cache rendered an instanced scene")
    except KrtError as e:
        assert e.code == KRT_ERR_STATE, e
    say("instancing: restir_gi and the radiance cache refuse the scene with KRT_ERR_STATE")
    del arm, arm2
    sc.free()


def case_attachments(ctx):
    """Every renderer on a small flat scene, bit for bit vs the oracle: path
    tracer, ReSTIR DI and ReSTIR GI under a moving camera, the editor's
    G-buffer ReSTIR, served requests (batched and on lanes); the cached path tracer (MLP inference
    and training on the side stream) runs for coverage."""
    from artemis_amd import Mamba, Server

    s = scenes.sky_room()
    sc = ctx.upload(s)
    osc = O.OracleScene(s, sc.serialized())
    W, H = 24, 16
    cams = [look_at((0.3 * f - 0.3, 1.0, 4.0), (0.0, 0.8, 0.0), aspect=W / H) for f in range(3)]
    arm = ArmadaRTX(ctx, W, H)
    arm.attach("pt", PathTracer(ctx))
    arm.set_depth(4)
    for _ in range(2):
        arm.render(sc, cams[0])
    ref, _ = osc.render_pt(cams[0], W, H, frames=2, max_depth=4)
    assert bits_equal(arm.color_buffer(), ref[0]), "path tracer"
    say("attachments: path tracer bit-exact")
    del arm
    arm = ArmadaRTX(ctx, W, H)
    arm.attach("restir", ReSTIR(ctx))
    ors = O.OracleReSTIR(W, H)
    for cam in cams:
        arm.render(sc, cam, accumulate=False)
        ors.frame(osc, cam, max_depth=2, reset=True)
    for which in (0, 1):
        assert bits_equal(arm.reservoirs(which), ors.reservoirs(which)), f"restir reservoirs {which}"
    assert bits_equal(arm.color_buffer(), ors.buffers[0]), "restir colour"
    say("attachments: ReSTIR DI under a moving camera bit-exact")
    del arm
    arm = ArmadaRTX(ctx, W, H)
    arm.attach("gi", ReSTIRGI(ctx))
    arm.set_option("radius", 8)
    arm.set_depth(3)
    ogi = O.OracleGI(W, H)
    for cam in cams:
        arm.render(sc, cam, accumulate=False)
        ogi.frame(osc, cam, max_depth=3, m_cap=30, spatial_passes=2, neighbours=5, radius=8, accumulate=False)
    for which in (0, 1):
        assert bits_equal(arm.reservoirs(which, gi=True), ogi.reservoirs(which)), f"gi reservoirs {which}"
    assert bits_equal(arm.color_buffer(), ogi.color), "gi colour"
    say("attachments: ReSTIR GI under a moving camera bit-exact")
    del arm
    arm = ArmadaRTX(ctx, W, H)
    arm.attach("mamba", Mamba(ctx))
    om = O.OracleMamba(osc, W, H)
    for cam in cams:
        arm.render(sc, cam)
        om.frame(cam)
    assert bits_equal(arm.color_buffer(), om.color), "mamba colour"
    say("attachments: editor G-buffer ReSTIR bit-exact")
    del arm
    srv = Server(ctx)
    rid = srv.submit(sc, cams[1], W, H, spp=2, max_depth=5, seed=3)
    img, _ = srv.wait(rid)
    srv.close()
    ref, _ = osc.render_pt(cams[1], W, H, frames=2, max_depth=5, user_seed=3)
    assert bits_equal(img, ref[0]), "served request"
    say("attachments: served path-tracer request bit-exact")
    # every request on a lane: the path tracer's wavefront passes with the
    # request's own buffers swapped in, beside a ReSTIR request
    srv = Server(ctx, wave_pixels=0)
    ropt = dict(candidates=8, m_cap=20, spatial_passes=1, neighbours=3, radius=10)
    rid = srv.submit(sc, cams[1], W, H, spp=2, max_depth=5, seed=3)
    rid2 = srv.submit(sc, cams[0], W, H, spp=2, max_depth=2, seed=4, kind="restir", **ropt)
    img, _ = srv.wait(rid)
    img2, _ = srv.wait(rid2)
    srv.close()
    assert bits_equal(img, ref[0]), "served request on a lane"
    ors = O.OracleReSTIR(W, H, user_seed=4)
    for _ in range(2):
        ors.frame(osc, cams[0], max_depth=2, **ropt)
    assert bits_equal(img2, ors.buffers[0]), "served ReSTIR request"
    say("attachments: served requests on lanes bit-exact")
    arm = ArmadaRTX(ctx, W, H)
    arm.attach("pt", PathTracer(ctx))
    arm.set_option("radiance_cache", True)
    arm.set_option("train_samples", 128)
    for _ in range(3):
        arm.render(sc, cams[0])
    assert np.all(np.isfinite(arm.color_buffer())), "cached path tracer"
    say("attachments: cached path tracer (inference + training) ran")
    del arm
    sc.free()


def case_envmap(ctx):
    """The armada's environment map: misses, next-event estimation and
    ReSTIR's environment candidates, bit for bit vs the oracle; ReSTIR GI
    refuses to render while it is on."""
    s = scenes.sky_room()
    sc = ctx.upload(s)
    osc = O.OracleScene(s, sc.serialized())
    rng = np.random.default_rng(5)
    env = rng.uniform(0.0, 3.0, (8, 16, 4)).astype(np.float32)
    osc.set_envmap(env)
    W, H = 24, 16
    cam = look_at(**s.camera, aspect=W / H)
    arm = ArmadaRTX(ctx, W, H)
    arm.set_envmap(env)
    arm.attach("pt", PathTracer(ctx))
    arm.set_depth(4)
    for _ in range(2):
        arm.render(sc, cam)
    ref, _ = osc.render_pt(cam, W, H, frames=2, max_depth=4)
    for k in range(4):
        assert bits_equal(arm._read(k), ref[k]), f"path tracer buffer {k}"
    say("envmap: path tracer bit-exact")
    del arm
    arm = ArmadaRTX(ctx, W, H)
    arm.set_envmap(env)
    arm.attach("restir", ReSTIR(ctx))
    ors = O.OracleReSTIR(W, H)
    for _ in range(2):
        arm.render(sc, cam)
        ors.frame(osc, cam, max_depth=2)
    for which in (0, 1):
        assert bits_equal(arm.reservoirs(which), ors.reservoirs(which)), f"restir reservoirs {which}"
    assert bits_equal(arm.color_buffer(), ors.buffers[0]), "restir colour"
    say("envmap: ReSTIR DI bit-exact")
    arm.attach("gi", ReSTIRGI(ctx))
    try:
        arm.render(sc, cam)
        raise AssertionError("restir_gi rendered with an environment map")
    except KrtError as e:
        assert e.code == KRT_ERR_STATE, e
    arm.set_envmap_enabled(False)
    arm.render(sc, cam)
    say("envmap: restir_gi refuses it (KRT_ERR_STATE) and renders once it is off")
    del arm
    sc.free()


def case_textures(ctx):
    """Material textures and normal maps: the path tracer, ReSTIR DI under a
    moving camera and a served request on a lane, bit for bit vs the
    oracle."""
    from artemis_amd import Server

    s = scenes.textured_room(coarse=True)
    sc = ctx.upload(s)
    osc = O.OracleScene(s, sc.serialized())
    assert bits_equal(sc.lights(), osc.lights()), "lights"
    W, H = 16, 12
    cam = look_at(**s.camera, aspect=W / H)
    arm = ArmadaRTX(ctx, W, H)
    arm.attach("pt", PathTracer(ctx))
    arm.set_depth(4)
    for _ in range(2):
        arm.render(sc, cam)
    ref, _ = osc.render_pt(cam, W, H, frames=2, max_depth=4)
    for k in range(4):
        assert bits_equal(arm._read(k), ref[k]), f"path tracer buffer {k}"
    say("textures: path tracer bit-exact")
    del arm
    arm = ArmadaRTX(ctx, W, H)
    arm.attach("restir", ReSTIR(ctx))
    ors = O.OracleReSTIR(W, H)
    for f in range(3):
        c = look_at((0.2 * f - 0.2, 1.0, 3.5), (0.0, 0.9, 0.0), aspect=W / H)
        arm.render(sc, c, accumulate=False)
        ors.frame(osc, c, max_depth=2, reset=True)
        for which in (0, 1):
            assert bits_equal(arm.reservoirs(which), ors.reservoirs(which)), f"restir reservoirs {which}"
        assert bits_equal(arm.color_buffer(), ors.buffers[0]), "restir colour"
    say("textures: ReSTIR DI (moving camera) bit-exact")
    del arm
    srv = Server(ctx, wave_pixels=0)
    img, _ = srv.wait(srv.submit(sc, cam, W, H, spp=2, max_depth=4, seed=0))
    srv.close()
    assert bits_equal(img, ref[0]), "served request"
    say("textures: served request on a lane bit-exact")
    arm = ArmadaRTX(ctx, W, H)
    arm.attach("gi", ReSTIRGI(ctx))
    arm.set_option("radius", 8)
    arm.set_depth(3)
    ogi = O.OracleGI(W, H)
    for _ in range(1):
        arm.render(sc, cam)
        ogi.frame(osc, cam, max_depth=3, m_cap=30, spatial_passes=2, neighbours=5, radius=8)
    for which in (0, 1):
        assert bits_equal(arm.reservoirs(which, gi=True), ogi.reservoirs(which)), f"gi reservoirs {which}"
    assert bits_equal(arm.color_buffer(), ogi.color), "gi colour"
    say("textures: ReSTIR GI bit-exact")
    del arm
    from artemis_amd import Mamba

    arm = ArmadaRTX(ctx, W, H)
    arm.attach("mamba", Mamba(ctx))
    om = O.OracleMamba(osc, W, H)
    for f in range(2):
        c = look_at((0.2 * f - 0.2, 1.0, 3.5), (0.0, 0.9, 0.0), aspect=W / H)
        arm.render(sc, c)
        om.frame(c, temporal=True, spatial=True)
    for which in (0, 1):
        assert bits_equal(arm.reservoirs(which), om.reservoirs(which)), f"mamba reservoirs {which}"
    assert bits_equal(arm.color_buffer(), om.color), "mamba colour"
    say("textures: editor ReSTIR bit-exact")
    del arm
    sc.free()
    # the PNG decoder: a written file reads back; a truncated one is refused
    import tempfile
    from artemis_amd import read_png, write_png

    with tempfile.TemporaryDirectory() as tmp:
        px = np.random.default_rng(1).integers(0, 2 ** 32, (9, 7), dtype=np.uint64).astype(np.uint32)
        write_png(f"{tmp}/a.png", px)
        assert np.array_equal(read_png(f"{tmp}/a.png").view(np.uint32)[..., 0], px), "png round trip"
        data = open(f"{tmp}/a.png", "rb").read()
        open(f"{tmp}/b.png", "wb").write(data[:len(data) - 20])
        try:
            read_png(f"{tmp}/b.png")
            raise AssertionError("truncated PNG decoded")
        except KrtError:
            pass
    say("textures: PNG decoder round trip and truncated file")


def case_refit(ctx):
    """Moved geometry: update_vertices makes the scene stale (renders refuse),
    refit gives the numpy restatement's tree bit for bit and frames equal the
    oracle's over the moved geometry and the refit tree; a light moves too."""
    from tests.test_oracle_refit import moved, refit_serialized

    s = scenes.sky_room()
    sc = ctx.upload(s)
    old = sc.serialized()
    new, first, count = moved(s, 2, (0.3, 0.1, -0.2))
    new, lfirst, lcount = moved(new, 3, (0.0, 0.5, 0.1))  # the ceiling light
    sc.update_vertices(first, new.positions[first:first + count])
    sc.update_vertices(lfirst, new.positions[lfirst:lfirst + lcount])
    W, H = 24, 16
    cam = look_at(**s.camera, aspect=W / H)
    arm = ArmadaRTX(ctx, W, H)
    arm.attach("pt", PathTracer(ctx))
    arm.set_depth(4)
    try:
        arm.render(sc, cam)
        raise AssertionError("a stale scene rendered")
    except KrtError as e:
        assert e.code == KRT_ERR_STATE, e
    sc.refit()
    b = O.tri_boxes(new.positions, new.indices)
    want = refit_serialized(old, b[:, 0:3], b[:, 3:6])
    assert bits_equal(sc.serialized(), want), "refit tree"
    osc = O.OracleScene(new, want)
    assert bits_equal(sc.lights(), osc.lights()), "moved light"
    for _ in range(2):
        arm.render(sc, cam)
    ref, _ = osc.render_pt(cam, W, H, frames=2, max_depth=4)
    assert bits_equal(arm.color_buffer(), ref[0]), "path tracer over the refit tree"
    say("refit: tree = numpy refit, lights moved, path tracer bit-exact")
    rays = np.zeros((2000, 8), np.float32)
    rng = np.random.default_rng(2)
    rays[:, :3] = rng.uniform(-3, 3, (2000, 3)) + np.array([0, 2, 0])
    d = rng.normal(size=(2000, 3))
    rays[:, 4:7] = d / np.linalg.norm(d, axis=1, keepdims=True)
    rays[:, 3] = 1e16
    want_h, _ = osc.trace(rays, mode=0)
    assert bits_equal(sc.trace_closest(rays), want_h), "closest hits"
    say("refit: 2,000 closest hits bit-exact")
    del arm
    sc.free()


def case_mlp(ctx):
    """The radiance-cache network's inference and gradient against the
    oracle's BF16 mode (the GPU tests' checks, on the emulated MFMA)."""
    from tests import test_gpu_mlp as T

    for n in (1, 65, 1000):
        T.test_infer_matches_oracle(ctx, n)
    for n in (1, 300):
        T.test_gradient_matches_oracle(ctx, n)
    say("mlp: inference and gradients within the BF16 bounds of the oracle")


CASES = {"flat_pt": case_flat_pt, "attachments": case_attachments, "envmap": case_envmap, "refit": case_refit,
         "textures": case_textures,
         "mlp": case_mlp, "instancing": case_instancing,
         "instancing_full": lambda ctx: case_instancing(ctx, full=True)}

if __name__ == "__main__":
    name = sys.argv[1]
    with Context(0) as ctx:
        CASES[name](ctx)
    say(f"case {name}: ok")
