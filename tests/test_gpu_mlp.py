"""The closure network's MFMA kernels against the CPU restatement
(oracle/mlp_oracle.py, which applies the same BF16 roundings and accumulates
in float64): inference, the gradient of the loss (every parameter), and a
training run that must fit a smooth function.  Tolerances: FP32 MFMA
accumulation against float64 flips an occasional BF16 rounding of a hidden
activation, so outputs agree to 1e-2 of their range at worst and the gradient
to 2e-2 in relative L2 norm; typical differences are ~1e-5."""
import numpy as np
import pytest

from peridem_amd import capi
from oracle import mlp_oracle as M

pytestmark = pytest.mark.gpu


def _norm(rng):
    mean = rng.normal(size=16).astype(np.float32) * 0.1
    scale = rng.uniform(0.5, 2.0, size=16).astype(np.float32)
    osc = np.array([1.0, 2.0, 0.5, 4.0], np.float32)
    ow = np.array([1.0, 0.5, 2.0, 1.0], np.float32)
    return mean, scale, osc, ow


def test_forward_matches_oracle():
    rng = np.random.default_rng(0)
    net = capi.Mlp(seed=3)
    mean, scale, osc, ow = _norm(rng)
    net.set_normalization(mean, scale, osc, ow)
    n0, s0, o0, w0 = net.normalization()
    np.testing.assert_array_equal(n0, mean)
    np.testing.assert_array_equal(o0, osc)
    x = rng.normal(size=(1000, 16)).astype(np.float32)  # not a multiple of 16 or 64
    y = net.forward(x)
    p = net.params().astype(np.float64)
    yr = M.predict(p, x, mean, scale, osc)
    err = np.abs(y - yr).max() / np.abs(yr).max()
    med = np.median(np.abs(y - yr)) / np.abs(yr).max()
    print(f"[mlp] forward max rel {err:.3e} median {med:.3e}", flush=True)
    assert err < 1e-2 and med < 1e-4


def test_gradient_matches_oracle():
    rng = np.random.default_rng(1)
    net = capi.Mlp(seed=5)
    mean, scale, osc, ow = _norm(rng)
    net.set_normalization(mean, scale, osc, ow)
    x = rng.normal(size=(200, 16)).astype(np.float32)
    t = rng.normal(size=(200, 4)).astype(np.float32)
    loss, g = net.gradient(x, t)
    p = net.params().astype(np.float64)
    lr_, gr = M.loss_and_grad(p, x, t, mean, scale, osc, ow)
    rel = np.linalg.norm(g - gr) / np.linalg.norm(gr)
    print(f"[mlp] loss {loss:.6e} oracle {lr_:.6e}; gradient rel L2 {rel:.3e}", flush=True)
    assert abs(loss - lr_) <= 1e-3 * lr_
    assert rel < 2e-2
    # per block (every layer's weights and biases carry gradient)
    o = 0
    for s in M.SHAPES:
        k = int(np.prod(s))
        a, b = g[o:o + k], gr[o:o + k]
        assert np.linalg.norm(b) > 0
        assert np.linalg.norm(a - b) < 5e-2 * np.linalg.norm(b), s
        o += k
    # the gradient call does not move the parameters
    np.testing.assert_array_equal(net.params(), p.astype(np.float32))


def test_params_round_trip():
    net = capi.Mlp(seed=7)
    p = np.random.default_rng(2).normal(size=net.nparam).astype(np.float32) * 0.05
    net.set_params(p)
    np.testing.assert_array_equal(net.params(), p)


def _target(x):
    return np.stack([np.sin(2 * x[:, 0]) + 0.5 * x[:, 1] ** 2, x[:, 2] * x[:, 3],
                     np.maximum(x[:, 4], 0.0) - 0.3 * x[:, 5], np.abs(x[:, 6]) * 0.5], 1)


def test_training_fits_a_smooth_function():
    rng = np.random.default_rng(3)
    x = rng.normal(size=(20000, 16)).astype(np.float32)
    t = _target(x).astype(np.float32)
    net = capi.Mlp(seed=11)
    net.set_normalization(np.zeros(16), np.ones(16), 1.0 / t.std(0), np.ones(4))
    y0 = net.forward(x[:2000])
    curve = net.train(x, t, iters=3000, batch=512, lr=2e-3, seed=1)
    xv = rng.normal(size=(2000, 16)).astype(np.float32)
    tv = _target(xv)
    yv = net.forward(xv)
    rel = np.sqrt(((yv - tv) ** 2).mean(0)) / tv.std(0)
    rel0 = np.sqrt(((y0 - t[:2000]) ** 2).mean(0)) / t.std(0)
    print(f"[mlp] loss {curve[0]:.4f} -> {curve[-50:].mean():.5f}; held-out rel RMS {rel} (init {rel0})",
          flush=True)
    assert np.all(np.isfinite(curve))
    assert curve[-50:].mean() < 0.05 * curve[:10].mean()
    assert np.all(rel < 0.2)
    # the trained network still matches its CPU restatement
    mean, scale, osc, _ = net.normalization()
    yr = M.predict(net.params().astype(np.float64), xv, mean, scale, osc)
    assert np.abs(yv - yr).max() < 1e-2 * np.abs(yr).max()
