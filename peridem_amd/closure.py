"""Learned pair-contact closure: training data from a batched ensemble, GPU
training of the network with the engine's own MFMA kernels, and the
closure-vs-plain comparison on a model.

Pipeline (all on the GPU through the C-ABI):
  1. run an ensemble of independent 2D collisions with the plain contact
     method and log, every few steps, the pair features and the plain pair
     forces of every particle pair near contact (pdem_pair_log);
  2. normalise features and targets and train the 16-128-128-128-4 network
     (pdem_mlp_train: minibatch Adam, BF16 MFMA forward/backward);
  3. enable the closure on a model (pdem_closure_enable) and compare it with
     the plain method (observables, sampled exact checks, step time).
"""
from __future__ import annotations

import time

import numpy as np

from . import capi

# rows used for training: pairs whose penetration Rc - gap is below this many
# mesh sizes (twice the closure band, so the fit is smooth at its edge and the
# target scale is that of the band)
TRAIN_PEN_H = 0.1
OUT_WEIGHT = (1.0, 1.0, 1.0, 0.1)


def particle_codes(spec):
    """Feature 11/12 of a pair: the shape class of each particle's zone."""
    shape = spec.info.get("shape_of_zone")
    if shape is None:
        return np.zeros(len(spec.particles), np.float32)
    return np.array([shape[p["zone"]] for p in spec.particles], np.float32)


def collect(spec, steps, log_every=10, band_h=0.5, precision=capi.PREC_MIXED, cap=1 << 20, verbose=False):
    """Run `spec` with the plain method and log pair rows every `log_every`
    steps.  Returns features X and the network targets T: the plain pair
    force and moment minus their normal-contact baseline, and the local Z."""
    eng = capi.GpuEngine(spec, precision)
    eng.closure_setup(particle_codes(spec))
    eng.integrate_begin()
    X, T = [], []
    t0 = time.perf_counter()
    for k in range(steps // log_every):
        eng.integrate_cd_graph(log_every)
        x, t, b = eng.pair_log(band_h, cap, baseline=True)
        t[:, :3] -= b
        X.append(x)
        T.append(t)
        if verbose and (k + 1) % max(1, (steps // log_every) // 10) == 0:
            n = sum(len(a) for a in X)
            print(f"# collect: step {(k + 1) * log_every}, {n} rows, {time.perf_counter() - t0:.1f} s",
                  flush=True)
    eng.close()
    X = np.concatenate(X) if X else np.zeros((0, 16), np.float32)
    T = np.concatenate(T) if T else np.zeros((0, 4), np.float32)
    return X, T


# feature columns (pdem_closure.hpp, pair_features)
F_GAP, F_HMIN, F_DMIN = 0, 1, 14


def select(X, T, pen_h=TRAIN_PEN_H):
    """Rows the network would see: node pairs in contact (nearest patch pair
    closer than Rc, extent gap >= 0) with penetration Rc - dmin below pen_h
    mesh sizes of the finer particle."""
    hmin_rc = X[:, F_HMIN]
    keep = (X[:, F_GAP] >= 0.0) & (X[:, F_DMIN] < 1.0) & (X[:, F_DMIN] >= 1.0 - pen_h * hmin_rc)
    keep &= T[:, 3] < 1.0  # damaged pairs stay with the plain method
    return X[keep], T[keep]


def fit(X, T, iters=4000, batch=1024, lr=2e-3, seed=0, verbose=False):
    """Train a network on (X, T); returns (network, loss curve)."""
    if len(X) < 64:
        raise ValueError("too few training rows")
    mean = X.mean(0)
    std = X.std(0)
    scale = np.where(std > 1e-6, 1.0 / np.maximum(std, 1e-6), 1.0)
    tstd = T.std(0)
    oscale = np.where(tstd > 1e-9, 1.0 / np.maximum(tstd, 1e-9), 1.0)
    net = capi.Mlp(seed=seed)
    net.set_normalization(mean, scale, oscale, OUT_WEIGHT)
    t0 = time.perf_counter()
    curve = net.train(X, T, iters=iters, batch=batch, lr=lr, seed=seed + 1)
    if verbose:
        print(f"# fit: {len(X)} rows, {iters} iterations in {time.perf_counter() - t0:.2f} s, "
              f"loss {curve[:10].mean():.4f} -> {curve[-50:].mean():.5f}", flush=True)
    return net, curve


def fit_report(net, X, T):
    """Held-out quality: RMS error per target over the target's RMS."""
    Y = net.forward(X)
    rms = np.sqrt(((Y - T) ** 2).mean(0))
    ref = np.sqrt((T ** 2).mean(0))
    return rms / np.maximum(ref, 1e-30)
