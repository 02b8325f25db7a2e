"""The closure network's CPU restatement (oracle/mlp_oracle.py), pinned by
mathematics: BF16 rounding on known values, the float64 gradient against
central finite differences, and the BF16-emulating gradient close to it."""
import numpy as np

from oracle import mlp_oracle as M


def _net(seed=0, scale=1.0):
    rng = np.random.default_rng(seed)
    parts = []
    for s in M.SHAPES:
        fan = s[0] if len(s) == 2 else 1
        a = np.sqrt(6.0 / fan) if len(s) == 2 else 0.1
        parts.append(rng.uniform(-a, a, size=s) * scale)
    return M.pack(parts)


def test_bf16_rounding_known_values():
    one = 1.0
    e = 2.0 ** -8  # bf16 spacing at 1.0 is 2^-7
    vals = np.array([one, one + e, one + 3 * e, one + 2 ** -7, -2.5, 0.0, 3.0e38], np.float32)
    got = M.bf16(vals)
    # 1 + 2^-8 is a tie: to even (1.0); 1 + 3*2^-8 rounds up to 1 + 2^-6
    want = np.array([1.0, 1.0, 1.0 + 2 ** -6, 1.0 + 2 ** -7, -2.5, 0.0], np.float64)
    np.testing.assert_array_equal(got[:6], want)
    assert abs(got[6] / 3.0e38 - 1) < 2 ** -8


def test_float64_gradient_matches_finite_differences():
    rng = np.random.default_rng(1)
    p = _net(1)
    x = rng.normal(size=(12, M.NIN))
    t = rng.normal(size=(12, M.NOUT))
    mean, scale = np.zeros(M.NIN), np.ones(M.NIN)
    osc, ow = np.array([1.0, 2.0, 0.5, 1.0]), np.array([1.0, 0.5, 2.0, 1.0])
    loss, g = M.loss_and_grad(p, x, t, mean, scale, osc, ow, bf16_on=False)
    assert np.isfinite(loss) and loss > 0
    idx = np.concatenate([rng.choice(M.NPARAM, 160, replace=False),
                          np.arange(M.NPARAM - 4, M.NPARAM)])  # output biases too
    eps = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += eps
        pm[i] -= eps
        lp, _ = M.loss_and_grad(pp, x, t, mean, scale, osc, ow, bf16_on=False)
        lm, _ = M.loss_and_grad(pm, x, t, mean, scale, osc, ow, bf16_on=False)
        fd = (lp - lm) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(g[i])) + 1e-8, (i, fd, g[i])


def test_bf16_emulation_stays_close_to_float64():
    rng = np.random.default_rng(2)
    p = _net(2)
    x = rng.normal(size=(64, M.NIN))
    t = rng.normal(size=(64, M.NOUT))
    args = (np.zeros(M.NIN), np.ones(M.NIN), np.ones(M.NOUT), np.ones(M.NOUT))
    l64, g64 = M.loss_and_grad(p, x, t, *args, bf16_on=False)
    l16, g16 = M.loss_and_grad(p, x, t, *args, bf16_on=True)
    assert abs(l16 - l64) < 0.05 * l64
    assert np.linalg.norm(g16 - g64) < 0.05 * np.linalg.norm(g64)


def test_relu_network_is_piecewise_linear_in_inputs():
    # doubling all weights of the last layer and its bias doubles the output
    p = _net(3)
    x = np.random.default_rng(3).normal(size=(5, M.NIN))
    parts = M.unpack(p)
    parts[6] = parts[6] * 2
    parts[7] = parts[7] * 2
    a = M.predict(p, x, np.zeros(16), np.ones(16), np.ones(4), bf16_on=False)
    b = M.predict(M.pack(parts), x, np.zeros(16), np.ones(16), np.ones(4), bf16_on=False)
    np.testing.assert_allclose(b, 2 * a, rtol=1e-12, atol=1e-12)
