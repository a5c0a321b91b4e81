"""Radiance-cache MLP kernels (BF16 MFMA, FP32 accumulation) against the
oracle through the C ABI: inference and gradients against the oracle's BF16
mode (same rounding points; the device sums in FP32, the oracle in FP64),
their distance to the FP64 model stated, Adam bit for bit, loss-scale
back-off on a non-finite gradient, and training that fits a smooth target."""
import numpy as np
import pytest

import oracle as O
from artemis_amd import RadianceField

from .test_oracle_mlp import PMIN, PSCALE, random_params, random_queries

pytestmark = pytest.mark.gpu


def make_net(ctx, params=None, **kw):
    net = RadianceField(ctx, seed=7, pos_min=PMIN, pos_scale=PSCALE, **kw)
    if params is not None:
        net.set_params(params.astype(np.float32))
    return net


def test_xavier_init(krt_ctx):
    # mlp.cuh:13-31: weights uniform in +-xavier(n, m) / 2
    p = make_net(krt_ctx).params().astype(np.float64)
    first = p[:62 * 64]
    a = np.sqrt(6.0 / (62 + 64))
    assert np.abs(first).max() <= a / 2
    assert np.abs(first).max() > 0.45 * a
    assert abs(first.mean()) < 0.02 * a


@pytest.mark.parametrize("n", [1, 15, 64, 65, 1000, 4099])
def test_infer_matches_oracle(krt_ctx, n):
    rng = np.random.default_rng(n)
    params = random_params(rng, 1.5).astype(np.float32)
    net = make_net(krt_ctx, params)
    raw = random_queries(rng, n)
    y = net.infer(raw).astype(np.float64)
    y16 = O.mlp_forward(params, raw, PMIN, PSCALE, bf16=True)
    y64 = O.mlp_forward(params, raw, PMIN, PSCALE, bf16=False)
    # same rounding points: only FP32-vs-FP64 summation differs (rare bf16
    # rounding flips of a hidden unit)
    scale = np.abs(y16).max() + 1e-3
    assert np.abs(y - y16).max() <= 4e-3 * scale
    assert np.median(np.abs(y - y16)) <= 1e-5 * scale
    # BF16 error of the network against the FP64 model
    assert np.linalg.norm(y - y64) / np.linalg.norm(y64) < 2e-2


def test_infer_empty(krt_ctx):
    net = make_net(krt_ctx)
    assert net.infer(np.zeros((0, 13), np.float32)).shape == (0, 3)


@pytest.mark.parametrize("n", [1, 128, 300, 2048])
def test_gradient_matches_oracle(krt_ctx, n):
    rng = np.random.default_rng(100 + n)
    params = random_params(rng, 1.5).astype(np.float32)
    net = make_net(krt_ctx, params)
    raw = random_queries(rng, n)
    target = rng.uniform(0, 2, (n, 3)).astype(np.float32)
    scale = net.state()["loss_scale"]
    loss, g = net.gradient(raw, target)
    l16, g16 = O.mlp_grad(params, raw, target, PMIN, PSCALE, bf16=True, loss_scale=scale)
    l64, g64 = O.mlp_grad(params, raw, target, PMIN, PSCALE, bf16=False)
    assert loss == pytest.approx(l16, rel=1e-4)
    g = g.astype(np.float64)
    # per layer, weights and biases separately
    bounds = [0, 3968, 4032]
    for _ in range(4):
        bounds += [bounds[-1] + 4096, bounds[-1] + 4096 + 64]
    bounds += [20864, 20867]
    for a, b in zip(bounds[:-1], bounds[1:]):
        ref = g16[a:b]
        assert np.linalg.norm(g[a:b] - ref) <= 2e-3 * np.linalg.norm(ref) + 1e-12, (a, b)
    assert np.linalg.norm(g - g64) / np.linalg.norm(g64) < 5e-2
    st = net.state()
    assert st["pending_nonfinite"] == 0 and st["step"] == 0


def test_adam_is_bit_exact(krt_ctx):
    rng = np.random.default_rng(5)
    net = make_net(krt_ctx)
    p = net.params().copy()
    m = np.zeros_like(p)
    v = np.zeros_like(p)
    cfg = net.config
    b1t = b2t = 1.0
    for step in range(3):
        g = (rng.normal(size=p.shape) * 10.0 ** rng.uniform(-6, 0, p.shape)).astype(np.float32)
        net.apply_gradient(g)
        b1t, b2t = O.mlp_adam(p, m, v, g, cfg.lr, cfg.beta1, cfg.beta2, cfg.eps, b1t, b2t)
        assert np.array_equal(net.params().view(np.uint32), p.view(np.uint32)), step
    st = net.state()
    assert st["step"] == 3 and st["skipped_steps"] == 0
    assert st["beta1_t"] == np.float32(b1t) and st["beta2_t"] == np.float32(b2t)


def test_train_step_equals_gradient_then_apply(krt_ctx):
    rng = np.random.default_rng(6)
    params = random_params(rng, 1.0).astype(np.float32)
    a, b = make_net(krt_ctx, params), make_net(krt_ctx, params)
    for _ in range(3):
        raw = random_queries(rng, 777)
        target = rng.uniform(0, 1, (777, 3)).astype(np.float32)
        la = a.train_step(raw, target)
        lb, g = b.gradient(raw, target)
        b.apply_gradient(g)
        assert la == lb
        assert np.array_equal(a.params().view(np.uint32), b.params().view(np.uint32))


def test_nonfinite_gradient_skips_step_and_halves_scale(krt_ctx):
    net = make_net(krt_ctx, loss_scale=4096.0)
    p0 = net.params()
    g = np.zeros(RadianceField.PARAMS, np.float32)
    g[1234] = np.inf
    net.apply_gradient(g)
    st = net.state()
    assert np.array_equal(net.params(), p0)
    assert st["skipped_steps"] == 1 and st["step"] == 0 and st["loss_scale"] == 2048.0
    assert st["pending_nonfinite"] == 0
    g[1234] = 0.5
    net.apply_gradient(g)
    st = net.state()
    assert st["step"] == 1 and st["skipped_steps"] == 1
    assert net.params()[1234] != p0[1234]


def test_loss_scale_grows_after_clean_steps(krt_ctx):
    net = make_net(krt_ctx, loss_scale=8.0, growth_interval=3)
    g = np.full(RadianceField.PARAMS, 1e-4, np.float32)
    for _ in range(7):
        net.apply_gradient(g)
    st = net.state()
    assert st["loss_scale"] == 32.0 and st["step"] == 7


def smooth_target(raw):
    pos, nrm, alb = raw[:, 0:3], raw[:, 6:9], raw[:, 9:12]
    shade = 0.6 + 0.4 * np.sin(1.3 * pos[:, :1]) * np.cos(0.7 * pos[:, 2:3]) + 0.2 * nrm[:, 1:2]
    return (alb * shade).astype(np.float32)


def test_training_fits_smooth_radiance(krt_ctx):
    rng = np.random.default_rng(9)
    net = make_net(krt_ctx)
    test_raw = random_queries(rng, 4096)
    test_t = smooth_target(test_raw)

    def rel_err():
        y = net.infer(test_raw)
        return float(np.mean((y - test_t) ** 2 / (test_t ** 2 + 1e-2)))

    e0 = rel_err()
    losses = []
    for _ in range(300):
        raw = random_queries(rng, 4096)
        losses.append(net.train_step(raw, smooth_target(raw)))
    e1 = rel_err()
    assert e1 < 0.05 * e0, (e0, e1)
    assert np.mean(losses[-20:]) < 0.05 * np.mean(losses[:5])
    assert net.state()["skipped_steps"] == 0
