"""Radiance-cache MLP oracle (oracle/mlp.cpp) pinned against closed forms:
numpy's sin/cos for the frequency encoding, central finite differences of the
FP64 loss for every gradient, Adam's first two steps computed by hand, and
the parameter count of the 62 -> 5 x 64 -> 3 network."""
import numpy as np
import pytest

import oracle as O

P = 62 * 64 + 64 + 4 * (64 * 64 + 64) + 64 * 3 + 3


def random_params(rng, scale=1.0):
    # layer-wise xavier-uniform magnitudes (editor/mlp.cuh:13-16), wider so
    # that most ReLUs are active and gradients reach every layer
    dims = [(62, 64)] + [(64, 64)] * 4 + [(64, 3)]
    out = []
    for n_in, n_out in dims:
        a = scale * np.sqrt(6.0 / (n_in + n_out))
        out.append(rng.uniform(-a, a, n_in * n_out))
        out.append(rng.uniform(-0.1, 0.1, n_out))
    return np.concatenate(out)


def random_queries(rng, n):
    pos = rng.uniform(-3, 5, (n, 3))
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    nrm = rng.normal(size=(n, 3))
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    alb = rng.uniform(0, 1, (n, 3))
    rough = rng.uniform(0, 1, (n, 1))
    return np.concatenate([pos, d, nrm, alb, rough], axis=1).astype(np.float32)


PMIN = np.array([-3, -3, -3], np.float32)
PSCALE = np.array([1 / 8, 1 / 8, 1 / 8], np.float32)


def test_param_count():
    assert O.mlp_param_count() == P == 20867


def test_encoding_matches_numpy():
    rng = np.random.default_rng(1)
    raw = random_queries(rng, 256)
    f = O.mlp_encode(raw, PMIN, PSCALE).astype(np.float64)
    pn = np.concatenate([(raw[:, :3].astype(np.float64) - PMIN) * PSCALE, raw[:, 3:9]], axis=1)
    expect = np.zeros((256, 64))
    col = 0
    for vec, freqs in ((0, 4), (1, 2), (2, 2)):
        for k in range(freqs):
            for comp in range(3):
                x = pn[:, 3 * vec + comp]
                expect[:, col] = np.sin(2.0 ** k * np.pi * x)
                expect[:, col + 1] = np.cos(2.0 ** k * np.pi * x)
                col += 2
    assert col == 48
    expect[:, 48:51] = pn[:, :3]
    expect[:, 51:61] = raw[:, 3:13]
    expect[:, 61] = 1.0
    np.testing.assert_allclose(f, expect, atol=3e-6, rtol=0)
    assert np.all(f[:, 62:] == 0)


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(2)
    params = random_params(rng, 1.5)
    raw = random_queries(rng, 12)
    target = rng.uniform(0, 2, (12, 3)).astype(np.float32)
    y0 = O.mlp_forward(params, raw, PMIN, PSCALE)
    den = y0 ** 2 + 0.01
    loss, g = O.mlp_grad(params, raw, target, PMIN, PSCALE)
    assert loss == pytest.approx(O.mlp_loss(params, raw, target, PMIN, PSCALE, den), rel=1e-12)
    # every layer's weights and biases, including the first layer's last input
    idx = list(rng.choice(P, 60, replace=False))
    idx += [0, 61, 3968, 3968 + 63, 4032, 20672, 20864, 20866]
    h = 1e-6
    for p in idx:
        e = np.zeros(P)
        e[p] = h
        fd = (O.mlp_loss(params + e, raw, target, PMIN, PSCALE, den) -
              O.mlp_loss(params - e, raw, target, PMIN, PSCALE, den)) / (2 * h)
        assert g[p] == pytest.approx(fd, rel=1e-5, abs=1e-9), p
    # whole-vector check: directional derivative along a random direction
    d = rng.normal(size=P)
    h = 1e-7
    fd = (O.mlp_loss(params + h * d, raw, target, PMIN, PSCALE, den) -
          O.mlp_loss(params - h * d, raw, target, PMIN, PSCALE, den)) / (2 * h)
    assert float(g @ d) == pytest.approx(fd, rel=1e-5)


def test_bf16_mode_is_within_bf16_error_of_fp64():
    rng = np.random.default_rng(3)
    params = random_params(rng, 1.5)
    raw = random_queries(rng, 512)
    target = rng.uniform(0, 2, (512, 3)).astype(np.float32)
    y64 = O.mlp_forward(params, raw, PMIN, PSCALE, bf16=False)
    y16 = O.mlp_forward(params, raw, PMIN, PSCALE, bf16=True)
    err = np.linalg.norm(y16 - y64) / np.linalg.norm(y64)
    assert 1e-4 < err < 2e-2  # bf16 rounding is visible, and small
    _, g64 = O.mlp_grad(params, raw, target, PMIN, PSCALE, bf16=False)
    _, g16 = O.mlp_grad(params, raw, target, PMIN, PSCALE, bf16=True, loss_scale=1024.0)
    gerr = np.linalg.norm(g16 - g64) / np.linalg.norm(g64)
    assert gerr < 5e-2
    # the loss scale only moves the bf16 rounding points: scale-invariant to
    # within rounding
    _, g16b = O.mlp_grad(params, raw, target, PMIN, PSCALE, bf16=True, loss_scale=1.0)
    assert np.linalg.norm(g16b - g16) / np.linalg.norm(g16) < 5e-2


def test_adam_first_two_steps_closed_form():
    rng = np.random.default_rng(4)
    n = 1000
    p0 = rng.normal(size=n).astype(np.float32)
    g1 = rng.normal(size=n).astype(np.float32)
    g2 = rng.normal(size=n).astype(np.float32)
    lr, b1, b2, eps = 1e-2, 0.9, 0.99, 1e-15
    p = p0.copy()
    m = np.zeros(n, np.float32)
    v = np.zeros(n, np.float32)
    t1, t2 = O.mlp_adam(p, m, v, g1, lr, b1, b2, eps, 1.0, 1.0)
    # step 1: bias-corrected moments are g and g^2, so the step is lr * sign(g)
    np.testing.assert_allclose(p, p0 - lr * np.sign(g1), rtol=0, atol=2e-7)
    assert t1 == pytest.approx(b1) and t2 == pytest.approx(b2)
    p1 = p.copy()
    O.mlp_adam(p, m, v, g2, lr, b1, b2, eps, t1, t2)
    g1d, g2d = g1.astype(np.float64), g2.astype(np.float64)
    mh = (b1 * (1 - b1) * g1d + (1 - b1) * g2d) / (1 - b1 ** 2)
    vh = (b2 * (1 - b2) * g1d ** 2 + (1 - b2) * g2d ** 2) / (1 - b2 ** 2)
    np.testing.assert_allclose(p, p1 - lr * mh / np.sqrt(vh), rtol=0, atol=2e-6)
