"""GPU BVH build and traversal vs the oracle (bit-exact)."""
import numpy as np
import pytest

import oracle as O
from artemis_amd import scenes
from tests.rays import adversarial_cornell, random_rays

pytestmark = pytest.mark.gpu


def degenerate_scene(kind):
    s = scenes.random_soup(40, seed=11)
    if kind == "duplicates":
        s.positions[3:] = np.tile(s.positions[:3], (len(s.positions) // 3 - 1, 1))
    elif kind == "points":
        s.positions[:] = np.repeat(s.positions[::3], 3, axis=0)
    elif kind == "collinear":
        t = np.linspace(0, 1, len(s.positions), dtype=np.float32)
        s.positions[:] = np.stack([t, 2 * t, -t], axis=1)
    elif kind == "signed_zero":
        s.positions[::2, 1] = -0.0
        s.positions[1::2, 1] = 0.0
    elif kind == "signed_zero_large":  # several elements per lane, ties between +0 and -0 at every level
        s = scenes.random_soup(6000, seed=12)
        rng = np.random.default_rng(4)
        for a in range(3):
            sel = rng.random(len(s.positions)) < 0.3
            s.positions[sel, a] = np.where(rng.random(sel.sum()) < 0.5, -0.0, 0.0).astype(np.float32)
    elif kind == "ties_large":  # past the top levels' pre-pass threshold: equal extents and +-0 everywhere
        s = scenes.random_soup(70000, seed=13)
        rng = np.random.default_rng(5)
        s.positions[:] = np.round(s.positions * 4.0) / 4.0
        for a in range(3):
            sel = rng.random(len(s.positions)) < 0.2
            s.positions[sel, a] = np.where(rng.random(sel.sum()) < 0.5, -0.0, 0.0).astype(np.float32)
    return s


SCENES = {
    "soup1": lambda: scenes.random_soup(1, seed=1),
    "soup2": lambda: scenes.random_soup(2, seed=2),
    "soup3": lambda: scenes.random_soup(3, seed=3),
    "soup5": lambda: scenes.random_soup(5, seed=5),
    "soup4096": lambda: scenes.random_soup(4096, seed=6),
    "duplicates": lambda: degenerate_scene("duplicates"),
    "points": lambda: degenerate_scene("points"),
    "collinear": lambda: degenerate_scene("collinear"),
    "signed_zero": lambda: degenerate_scene("signed_zero"),
    "signed_zero_large": lambda: degenerate_scene("signed_zero_large"),
    "ties_large": lambda: degenerate_scene("ties_large"),
    "cornell": scenes.cornell,
}


@pytest.mark.parametrize("name", list(SCENES))
def test_build_bit_exact(krt_ctx, name):
    s = SCENES[name]()
    print(f"[bvh build] scene={name} tris={s.triangles}", flush=True)
    ref = O.bvh_build(O.tri_boxes(s.positions, s.indices))
    sc = krt_ctx.upload(s, layout="reference")
    got = sc.serialized()
    assert got.shape == ref.shape
    np.testing.assert_array_equal(got.view(np.uint32), ref.view(np.uint32))
    np.testing.assert_array_equal(sc.leaf_order(), O.bvh_leaf_order(ref))
    sc.free()


def test_build_bit_exact_large(krt_ctx):
    s = scenes.spheres(instances=16, slices=64, stacks=33, emissive=64)
    print(f"[bvh build] spheres tris={s.triangles}", flush=True)
    ref = O.bvh_build(O.tri_boxes(s.positions, s.indices))
    sc = krt_ctx.upload(s)
    np.testing.assert_array_equal(sc.serialized().view(np.uint32), ref.view(np.uint32))
    sc.free()


def compare_hits(got, ref):
    g, r = got.view(np.uint32), ref.view(np.uint32)
    bad = np.where((g != r).any(axis=1))[0]
    assert len(bad) == 0, f"{len(bad)} mismatching rays, first {bad[:5]}: got {got[bad[:3]]} ref {ref[bad[:3]]}"


@pytest.mark.parametrize("layout,leaf_max", [("reference", 1), ("reference", 4), ("sah", 1), ("sah", 4),
                                             ("sah", 8)])
def test_closest_hit_cornell(krt_ctx, layout, leaf_max):
    s = scenes.cornell()
    sc = krt_ctx.upload(s, leaf_max=leaf_max, layout=layout)
    assert sc.stats()["layout"] == (0 if layout == "sah" else 1)
    osc = O.OracleScene(s, sc.serialized())
    rays = np.concatenate([adversarial_cornell(0), random_rays(3000, [-0.95, 0.05, -0.95], [0.95, 1.95, 0.95], 1)])
    print(f"[trace] cornell layout={layout} leaf_max={leaf_max} rays={len(rays)}", flush=True)
    ref, _ = osc.trace(rays, mode=0)
    compare_hits(sc.trace_closest(rays), ref)
    occ_ref = osc.occluded(rays)
    np.testing.assert_array_equal(sc.trace_occluded(rays), occ_ref)
    sc.free()


@pytest.mark.parametrize("layout", ["reference", "sah"])
def test_closest_hit_soup(krt_ctx, layout):
    s = scenes.random_soup(20000, seed=21, extent=2.0, size=0.3)
    sc = krt_ctx.upload(s, layout=layout)
    osc = O.OracleScene(s, sc.serialized())
    rays = random_rays(2000, -2.5, 2.5, 3)
    print(f"[trace] soup20000 layout={layout} rays={len(rays)}", flush=True)
    ref, _ = osc.trace(rays, mode=0)
    compare_hits(sc.trace_closest(rays), ref)
    sc.free()


@pytest.mark.parametrize("name", ["soup1", "soup2", "soup5", "points", "duplicates"])
def test_small_and_degenerate_scenes_trace(krt_ctx, name):
    s = SCENES[name]()
    for layout in ("reference", "sah"):
        sc = krt_ctx.upload(s, layout=layout)
        osc = O.OracleScene(s, sc.serialized())
        rays = random_rays(500, -1.2, 1.2, 8)
        print(f"[trace] {name} layout={layout}", flush=True)
        ref, _ = osc.trace(rays, mode=0)
        compare_hits(sc.trace_closest(rays), ref)
        np.testing.assert_array_equal(sc.trace_occluded(rays), osc.occluded(rays))
        sc.free()


def test_trace_tmax_and_miss(krt_ctx):
    s = scenes.cornell()
    sc = krt_ctx.upload(s)
    osc = O.OracleScene(s, sc.serialized())
    rays = random_rays(1000, [-0.9, 0.1, -0.9], [0.9, 1.9, 0.9], 5)
    rays[:, 3] = np.linspace(1e-4, 0.5, len(rays))
    ref, _ = osc.trace(rays, mode=0)
    compare_hits(sc.trace_closest(rays), ref)
    out = rays.copy()
    out[:, 0:3] = [0.0, 1.0, 30.0]
    out[:, 4:7] = [0.0, 0.0, 1.0]  # pointing away from the room
    got = sc.trace_closest(out)
    assert np.all(got[:, 3].view(np.int32) == -1)
    sc.free()


@pytest.mark.parametrize("jitter", [False, True])
def test_camera_rays_bit_exact(krt_ctx, jitter):
    s = scenes.cornell()
    from artemis_amd import look_at

    cam = look_at(**s.camera, aspect=1.0)
    for W, H, sample in ((32, 32, 0), (48, 40, 3)):
        got = krt_ctx.camera_rays(cam, W, H, jitter=jitter, sample=sample)
        ref = O.camera_rays(cam, W, H, jitter=jitter, sample=sample)
        bad = np.where(np.any(got.view(np.uint32) != ref.view(np.uint32), axis=1))[0]
        print(f"[camera rays] jitter={jitter} {W}x{H}: {len(bad)} differ; first {bad[:4]}", flush=True)
        if len(bad):
            print(got[bad[:2]], ref[bad[:2]])
        assert len(bad) == 0
