"""pcg3d / pcg3f (include/cuda/random.cuh:14-58, JCGT 2020 "Hash Functions for
GPU Rendering") and the quarter-turn sincos, pinned by an independent numpy
evaluation, committed known answers and distribution checks."""
import json
from pathlib import Path

import numpy as np

import oracle as O

KAT = json.loads((Path(__file__).parent / "golden" / "rng_kat.json").read_text())


def pcg3d_numpy(v):
    v = np.asarray(v, np.uint64) & 0xFFFFFFFF
    M = np.uint64(0xFFFFFFFF)
    v = (v * np.uint64(1664525) + np.uint64(1013904223)) & M
    x, y, z = v[:, 0].copy(), v[:, 1].copy(), v[:, 2].copy()
    x = (x + y * z) & M
    y = (y + z * x) & M
    z = (z + x * y) & M
    x ^= x >> np.uint64(16)
    y ^= y >> np.uint64(16)
    z ^= z >> np.uint64(16)
    x = (x + y * z) & M
    y = (y + z * x) & M
    z = (z + x * y) & M
    return np.stack([x, y, z], axis=1).astype(np.uint32)


def test_pcg3d_matches_independent_numpy():
    rng = np.random.default_rng(0)
    w = rng.integers(0, 2**32, (5000, 3), dtype=np.uint64).astype(np.uint32)
    np.testing.assert_array_equal(O.pcg3d(w), pcg3d_numpy(w))


def test_pcg3d_known_answers():
    np.testing.assert_array_equal(O.pcg3d(np.array(KAT["pcg3d_in"], np.uint32)), np.array(KAT["pcg3d_out"]))
    np.testing.assert_array_equal(pcg3d_numpy(np.array(KAT["pcg3d_in"])), np.array(KAT["pcg3d_out"]))


def test_pcg3f_bits_and_range():
    seeds = np.array(KAT["pcg3f_in"], np.uint32).view(np.float32)
    got = O.pcg3f(seeds)
    np.testing.assert_array_equal(got.view(np.uint32), np.array(KAT["pcg3f_out"], np.uint32))
    # pcg3f: mantissa of the hash under exponent 0 (value in [1,2)) minus 1
    h = pcg3d_numpy(seeds.view(np.uint32))
    expect = ((h & 0x007FFFFF) | 0x3F800000).view(np.float32) - np.float32(1.0)
    np.testing.assert_array_equal(got, expect)


def test_pcg3f_uniformity():
    n = 200000
    w = np.stack([np.arange(n) % 640, np.arange(n) // 640, np.full(n, 5)], axis=1).astype(np.uint32)
    u = O.pcg3f(w.view(np.float32))
    assert u.min() >= 0.0 and u.max() < 1.0
    hist, _ = np.histogram(u[:, 0], bins=32, range=(0, 1))
    chi2 = ((hist - n / 32) ** 2 / (n / 32)).sum()
    assert chi2 < 80  # 31 dof, p ~ 1e-6
    assert abs(np.corrcoef(u[:, 0], u[:, 1])[0, 1]) < 0.01


def test_sincos_turn():
    u = np.linspace(0, 1, 100001, endpoint=False).astype(np.float32)
    c, s = O.sincos_turn(u)
    ang = 2 * np.pi * u.astype(np.float64)
    assert np.max(np.abs(c - np.cos(ang))) < 4e-7
    assert np.max(np.abs(s - np.sin(ang))) < 4e-7
    k = np.array(KAT["sincos_u"], np.uint32).view(np.float32)
    c2, s2 = O.sincos_turn(k)
    np.testing.assert_array_equal(c2.view(np.uint32), KAT["sincos_c"])
    np.testing.assert_array_equal(s2.view(np.uint32), KAT["sincos_s"])
    # quarter turns are exact
    c3, s3 = O.sincos_turn(np.array([0.0, 0.25, 0.5, 0.75], np.float32))
    np.testing.assert_array_equal(c3, [1, 0, -1, 0])
    np.testing.assert_array_equal(s3, [0, 1, 0, -1])
