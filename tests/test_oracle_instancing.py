"""Two-level scenes in the oracle (Accelerator::update / build_tlas,
include/amadeus/accelerator.cuh:145-277): instance tree over the instances'
world boxes, rays taken into each mesh's space by the inverse transform.

Pinned by: the affine inverse against its closed form for rigid+scale
transforms; the instance-tree walk against brute force over every instance
and triangle (the lexicographic (t, instance rank, triangle rank) minimum);
the same scene flattened to world space (the textbook single-level answer):
same triangle ids, t within rounding of the transforms; the light list of the
flattened scene bit for bit; containment of every instance's vertices in its
tree box."""
import numpy as np
import pytest

import oracle as O
from artemis_amd import scenes


@pytest.fixture(scope="module")
def pair():
    inst = scenes.spheres(instances=24, slices=16, stacks=9, emissive=64, box=8.0, base_meshes=4, instanced=True)
    return inst, O.OracleScene(inst), O.OracleScene(inst.flatten())


def rays_at_instances(inst, n, seed):
    rng = np.random.default_rng(seed)
    centers = inst.transforms[:, :, 3]
    o = rng.uniform(-6, 6, (n, 3))
    tgt = centers[rng.integers(len(centers), size=n)] + rng.normal(scale=0.4, size=(n, 3))
    d = tgt - o
    d[: n // 8] /= np.linalg.norm(d[: n // 8], axis=1, keepdims=True)  # some unit, some not
    # a few rays start inside an instance
    o[n // 8: n // 4] = tgt[n // 8: n // 4] + rng.normal(scale=0.05, size=(n // 4 - n // 8, 3))
    rays = np.zeros((n, 8), np.float32)
    rays[:, :3] = o
    rays[:, 3] = 1e16
    rays[:, 4:7] = d
    rays[: n // 16, 3] = 5.0  # finite tmax
    return rays


def test_affine_inverse_closed_form():
    rng = np.random.default_rng(1)
    for _ in range(20):
        q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
        s = rng.uniform(0.3, 2.0)
        t = rng.normal(size=3) * 5
        M = np.concatenate([s * q, t[:, None]], 1).astype(np.float32)
        inv = O.affine_inverse(M)
        R = M[:, :3].astype(np.float64)
        want_r = np.linalg.inv(R)
        want_t = -want_r @ M[:, 3].astype(np.float64)
        np.testing.assert_allclose(inv[:, :3], want_r, rtol=1e-6, atol=1e-6)
        np.testing.assert_allclose(inv[:, 3], want_t, rtol=1e-6, atol=1e-5)


def test_walk_equals_brute_force(pair):
    inst, osc, _ = pair
    rays = rays_at_instances(inst, 3000, 2)
    h0, _ = osc.trace(rays, 0)
    h2, ctr = osc.trace(rays, 2)
    assert np.array_equal(h0.view(np.uint32), h2.view(np.uint32))
    assert (h0[:, 3].view(np.int32) >= 0).mean() > 0.5
    assert np.array_equal(osc.occluded(rays), (h0[:, 3].view(np.int32) >= 0).astype(np.int32))


def test_matches_flattened_scene(pair):
    inst, osc, flat = pair
    rays = rays_at_instances(inst, 3000, 3)
    hi, _ = osc.trace(rays, 0)
    hf, _ = flat.trace(rays, 0)
    pi, pf = hi[:, 3].view(np.int32), hf[:, 3].view(np.int32)
    both = (pi >= 0) & (pf >= 0)
    assert np.mean(pi == pf) > 0.995  # differences only at edges, from rounding of the transforms
    np.testing.assert_allclose(hi[both & (pi == pf), 0], hf[both & (pi == pf), 0], rtol=1e-4, atol=2e-6)  # grazing hits amplify rounding


def test_lights_equal_flattened(pair):
    _, osc, flat = pair
    a, b = osc.lights(), flat.lights()
    assert len(a) == 64 and np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_instance_tree_contains_instances(pair):
    inst, osc, _ = pair
    nodes = osc.tlas()
    n = len(inst.inst_mesh)
    assert len(nodes) == 2 * n - 1
    leaf = nodes[:, 1].view(np.int32)
    seen = sorted(leaf[leaf >= 0].tolist())
    assert seen == list(range(n))
    for k in np.nonzero(leaf >= 0)[0]:
        i = leaf[k]
        pos = inst.meshes[inst.inst_mesh[i]][0]
        w = scenes.xform_points(inst.transforms[i], pos)
        assert (w >= nodes[k, 4:7]).all() and (w <= nodes[k, 8:11]).all()
