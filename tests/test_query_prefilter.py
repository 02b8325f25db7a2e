"""The contact query's FP32 prefilter (query_node in pdem_kernels.hpp) must
never drop a true candidate: for every pair with FP64 distance below the
search radius r, the FP32 distance of the rounded positions (relative to the
grid origin) stays below Rf = r + 2^-19 (max|a| + r).  Emulated here with
numpy float32 in the kernel's operation order, on pairs placed just inside r
in random directions, with offsets from the grid origin up to 10^4 r, in 2D
and 3D.  The kernel's FP64 test alone decides the candidate set, so this
bound is all the prefilter has to satisfy."""
import numpy as np
import pytest

f32 = np.float32


def _rf(a_f, r):
    amax = np.max(np.abs(a_f), axis=1)
    rf = f32(r)
    return (rf + f32(2.0 ** -19) * (amax + rf)).astype(f32)


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("scale", [1.0, 30.0, 1e3, 1e4])
def test_prefilter_keeps_every_pair_within_r(dim, scale):
    rng = np.random.default_rng(17 + dim + int(scale))
    n = 200_000
    r = 3.7e-4
    glo = rng.uniform(-1.0, 1.0, 3) * 0.01
    a = glo + rng.uniform(-scale, scale, (n, 3)) * r
    u = rng.normal(size=(n, 3))
    u /= np.linalg.norm(u, axis=1)[:, None]
    dist = r * (1.0 - rng.uniform(0.0, 1e-6, n) ** 2)  # just inside r
    b = a + u * dist[:, None]
    if dim == 2:
        a[:, 2] = b[:, 2] = 0.0
        glo[2] = 0.0
    d2 = ((a - b) ** 2)[:, :dim].sum(1)
    inside = d2 < r * r
    assert inside.mean() > 0.5
    a_f = (a - glo).astype(f32)  # (float)(x - glo), FP64 difference rounded once
    b_f = (b - glo).astype(f32)
    t = (a_f - b_f).astype(f32)
    d2f = f32(0.0)
    for k in range(dim):
        d2f = (d2f + (t[:, k] * t[:, k]).astype(f32)).astype(f32)
    Rf = _rf(a_f[:, :dim], r)
    R2f = (Rf * Rf).astype(f32)
    kept = d2f < R2f
    assert np.all(kept[inside]), f"{(~kept[inside]).sum()} true candidates dropped"


def test_prefilter_rejects_far_entries():
    """The margin is small: entries at 1.001 r are rejected in FP32 already
    (so the filter saves the FP64 loads it is meant to save)."""
    rng = np.random.default_rng(3)
    n = 100_000
    r = 3.7e-4
    a = rng.uniform(-30, 30, (n, 3)) * r
    u = rng.normal(size=(n, 3))
    u /= np.linalg.norm(u, axis=1)[:, None]
    b = a + u * (1.001 * r)
    a_f, b_f = a.astype(f32), b.astype(f32)
    t = (a_f - b_f).astype(f32)
    d2f = ((t * t).astype(f32)).sum(1, dtype=f32)
    Rf = _rf(a_f, r)
    assert np.all(d2f >= (Rf * Rf).astype(f32))
