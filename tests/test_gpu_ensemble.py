"""Batched ensemble: many independent 2D simulations advanced as one model.

Instances sit far apart, so no contact, damping or candidate pair joins two
of them and the batch must reproduce each instance run on its own: in FP64
parity mode bit for bit (the contact sums run in ascending node id and the
damping sums in ascending particle id in both runs; instances keep their
relative order).  The oracle shows the same identity on the CPU
(test_oracle_invariants.test_ensemble_instances_independent).  The launch
shape autotuner must not change results beyond FP32 rounding of the mixed
kernels' block sums (1e-6 of the largest force; in practice identical)."""
import numpy as np
import pytest

from peridem_amd import capi, configs
import oracle

from parity_util import compare_state

pytestmark = pytest.mark.gpu


def _small_ensemble(n=4, seed=5):
    # equal radii on the tightest grid: gaps of 1.2 contact radii, closed
    # within a few hundred steps at 2-4 m/s
    return configs.ensemble_2d(n_instances=n, n_min=2, n_max=6, ratios=(6,), speed_lo=2.0,
                               speed_hi=4.0, pitch_factor=1.0, R_lo=1.0e-3, R_hi=1.0e-3, seed=seed)


def _slice(spec, k):
    first, last, _ = spec.info["instances"][k]
    return spec.particles[first]["glob_start"], spec.particles[last - 1]["glob_end"]


def test_ensemble_batched_equals_separate():
    spec = _small_ensemble()
    K = 400
    print(f"[ensemble] nodes={spec.n_nodes} particles={len(spec.particles)} steps={K}", flush=True)
    g = capi.GpuEngine(spec, capi.PREC_FP64)
    g.integrate_begin()
    g.integrate_cd_graph(K)
    r = oracle.OracleEngine(spec, 8)
    r.integrate_begin()
    r.integrate_cd(K)
    compare_state(g, r, tag="ensemble batched")
    F, U, V = g.field(capi.FIELD_F), g.field(capi.FIELD_U), g.field(capi.FIELD_V)
    Z = g.field(capi.FIELD_Z)
    touched = 0
    for k in range(len(spec.info["instances"])):
        s = configs.ensemble_instance(spec, k)
        e = capi.GpuEngine(s, capi.PREC_FP64)
        e.integrate_begin()
        e.integrate_cd_graph(K)
        n0, n1 = _slice(spec, k)
        np.testing.assert_array_equal(e.field(capi.FIELD_F), F[n0:n1])
        np.testing.assert_array_equal(e.field(capi.FIELD_U), U[n0:n1])
        np.testing.assert_array_equal(e.field(capi.FIELD_V), V[n0:n1])
        touched += int(Z[n0:n1].max() > 0.0)
        print(f"[ensemble] instance {k} equal", flush=True)
    assert touched >= 2, "too few instances deformed by contact in the test window"


def test_ensemble_mixed_against_oracle():
    spec = _small_ensemble(n=6, seed=11)
    K = 400
    print(f"[ensemble-mixed] nodes={spec.n_nodes} steps={K}", flush=True)
    g = capi.GpuEngine(spec, capi.PREC_MIXED)
    r = oracle.OracleEngine(spec, 8)
    g.integrate_begin()
    r.integrate_begin()
    g.integrate_cd_graph(K)
    r.integrate_cd(K)
    fg, fr = g.field(capi.FIELD_F), r.field(capi.FIELD_F)
    den = np.abs(fr).max()
    emax = float(np.abs(fg - fr).max() / den)
    erms = float(np.sqrt(((fg - fr) ** 2).sum(1).mean()) / den)
    print(f"[ensemble-mixed] max |f| {den:.3e} force max err {emax:.3e} rms {erms:.3e}", flush=True)
    assert den > 1e8  # contact forces, particles deform
    assert emax < 1e-4 and erms < 1e-5


@pytest.mark.parametrize("precision", [capi.PREC_FP64, capi.PREC_MIXED])
def test_autotune_keeps_results(precision):
    spec = _small_ensemble(n=4, seed=7)
    a = capi.GpuEngine(spec, precision)
    b = capi.GpuEngine(spec, precision)
    a.integrate_begin()
    b.integrate_begin()
    b.integrate_cd_graph(5)
    a.integrate_cd_graph(5)
    ms = b.autotune(reps=3)
    cfg = b.launch_config()
    print(f"[autotune] precision={precision} best {ms:.4f} ms config {cfg}", flush=True)
    assert ms > 0.0
    a.integrate_cd_graph(60)
    b.integrate_cd_graph(60)
    fa, fb = a.field(capi.FIELD_F), b.field(capi.FIELD_F)
    if precision == capi.PREC_FP64:
        np.testing.assert_array_equal(fa, fb)
    else:
        assert np.abs(fa - fb).max() <= 1e-6 * np.abs(fa).max()
    np.testing.assert_array_equal(a.field(capi.FIELD_BONDS), b.field(capi.FIELD_BONDS))


def test_light_contact_kernel_2d_matches_oracle():
    """The lane-per-node contact kernel in 2D (used by default for 2D models
    of 2^20 nodes and more) on a small ensemble with contacts, forced through
    pdem_set_contact_light_max: FP64 parity state equal to the oracle, and
    the same as the wavefront-per-node kernel bit for bit."""
    spec = _small_ensemble(n=4, seed=5)
    K = 400
    runs = {}
    for lm in (16, 0):
        g = capi.GpuEngine(spec, capi.PREC_FP64)
        g.set_contact_light_max(lm)
        g.integrate_begin()
        g.integrate_cd_graph(K)
        runs[lm] = g
    r = oracle.OracleEngine(spec, 8)
    r.integrate_begin()
    r.integrate_cd(K)
    out = compare_state(runs[16], r, tag="2D light contact kernel")
    print(f"[light-2d] nodes={spec.n_nodes} {out}", flush=True)
    assert runs[16].field(capi.FIELD_Z).max() > 0.0, "no contact deformation in the test window"
    for fid in (capi.FIELD_F, capi.FIELD_U, capi.FIELD_V):
        np.testing.assert_array_equal(runs[16].field(fid), runs[0].field(fid))
    with pytest.raises(RuntimeError):
        runs[0].set_contact_light_max(17)
