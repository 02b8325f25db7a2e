"""Mixed precision (FP32 bond arithmetic, FP64 sums and integration) against
the FP64 oracle on the state-based 3D packing.  Tolerances are those of FP32
bond arithmetic: the force error is reported as max |f - f_64| / max |f_64|
and RMS over nodes, and must stay below 1e-4 (max) and 1e-5 (RMS)."""
import numpy as np
import pytest

from peridem_amd import capi, configs
import oracle

pytestmark = pytest.mark.gpu


def _errors(fg, fr):
    den = np.abs(fr).max()
    d = np.abs(fg - fr)
    return float(d.max() / den), float(np.sqrt((d ** 2).sum(1).mean()) / den)


@pytest.mark.parametrize("steps", [0, 40])
def test_mixed_state_based_force_error(steps):
    spec = configs.packing_3d(n_spheres=6, nodes_per_sphere=300, L=5.0e-3, seed=2,
                              gap_factor=1.0, init_speed=2.0)
    print(f"[mixed] nodes={spec.n_nodes} steps={steps}", flush=True)
    g = capi.GpuEngine(spec, capi.PREC_MIXED)
    r = oracle.OracleEngine(spec, 8)
    g.integrate_begin()
    r.integrate_begin()
    if steps:
        g.integrate_cd(steps)
        r.integrate_cd(steps)
    emax, erms = _errors(g.field(capi.FIELD_F), r.field(capi.FIELD_F))
    th = g.field(capi.FIELD_THETA)
    thr = r.field(capi.FIELD_THETA)
    eth = float(np.abs(th - thr).max() / max(np.abs(thr).max(), 1e-300))
    print(f"[mixed] steps={steps} force max err {emax:.3e} rms {erms:.3e} theta err {eth:.3e}",
          flush=True)
    assert emax < 1e-4 and erms < 1e-5
    # neighbor sets are the same integer construction in both modes
    a, b = g.neighbors(capi.NEIGH_PD), r.neighbors(capi.NEIGH_PD)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


def test_mixed_graph_matches_eager():
    spec = configs.packing_3d(n_spheres=6, nodes_per_sphere=200, L=5.0e-3, seed=3,
                              gap_factor=1.0, init_speed=2.0)
    a = capi.GpuEngine(spec, capi.PREC_MIXED)
    b = capi.GpuEngine(spec, capi.PREC_MIXED)
    a.integrate_begin()
    b.integrate_begin()
    a.integrate_cd(30)
    b.integrate_cd_graph(30)
    np.testing.assert_array_equal(a.field(capi.FIELD_F), b.field(capi.FIELD_F))
