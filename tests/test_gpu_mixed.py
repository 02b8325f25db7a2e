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


def _smooth_displacement(xr, amp, lam):
    """A smooth non-rigid displacement field (peak strain ~ 2 pi amp / lam):
    every bond carries force."""
    u = np.zeros_like(xr)
    for d in range(xr.shape[1]):
        u[:, d] = amp * np.sin(2 * np.pi * xr[:, (d + 1) % xr.shape[1]] / lam + d)
    return u


@pytest.mark.parametrize("case", ["deformed", "steps40"])
def test_mixed_state_based_force_error(case):
    spec = configs.packing_3d(n_spheres=6, nodes_per_sphere=300, L=5.0e-3, seed=2,
                              gap_factor=1.0, init_speed=2.0)
    print(f"[mixed] nodes={spec.n_nodes} case={case}", flush=True)
    g = capi.GpuEngine(spec, capi.PREC_MIXED)
    r = oracle.OracleEngine(spec, 8)
    g.integrate_begin()
    r.integrate_begin()
    if case == "deformed":
        u = _smooth_displacement(spec.x_ref, 3.0e-7, 2.0e-3)
        for e in (g, r):
            e.set_field(capi.FIELD_U, u)
            e.set_field(capi.FIELD_X, spec.x_ref + u)
            e.compute_forces()
    else:
        g.integrate_cd(40)
        r.integrate_cd(40)
    fr = r.field(capi.FIELD_F)
    emax, erms = _errors(g.field(capi.FIELD_F), fr)
    th = g.field(capi.FIELD_THETA)
    thr = r.field(capi.FIELD_THETA)
    eth = float(np.abs(th - thr).max() / max(np.abs(thr).max(), 1e-300))
    print(f"[mixed] {case}: max |f| {np.abs(fr).max():.3e} force max err {emax:.3e} rms {erms:.3e} "
          f"theta err {eth:.3e}", flush=True)
    assert np.abs(thr).max() > 0.0
    assert 0.0 < emax < 1e-4 and erms < 1e-5
    assert eth < 1e-5
    # neighbor sets are the same integer construction in both modes
    a, b = g.neighbors(capi.NEIGH_PD), r.neighbors(capi.NEIGH_PD)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


def test_mixed_2d_large_window_against_oracle():
    """2D particles of more than 384 nodes (mesh size R/12): the mixed kernels'
    large-window instantiation in 2D, on a deformed state and after contact
    steps, against the oracle."""
    spec = configs.ensemble_2d(n_instances=2, n_min=3, n_max=4, ratios=(12,), speed_lo=2.0, speed_hi=4.0,
                               pitch_factor=1.0, R_lo=1.0e-3, R_hi=1.0e-3, seed=3)
    sizes = [p["glob_end"] - p["glob_start"] for p in spec.particles]
    print(f"[mixed-2d] nodes={spec.n_nodes} particle sizes {min(sizes)}..{max(sizes)}", flush=True)
    assert min(sizes) > 384
    g = capi.GpuEngine(spec, capi.PREC_MIXED)
    r = oracle.OracleEngine(spec, 8)
    for e in (g, r):
        e.integrate_begin()
    u = _smooth_displacement(spec.x_ref, 2.0e-7, 1.5e-3)
    for e in (g, r):
        e.set_field(capi.FIELD_U, u)
        e.set_field(capi.FIELD_X, spec.x_ref + u)
        e.compute_peridynamic_forces()
    emax, erms = _errors(g.field(capi.FIELD_F), r.field(capi.FIELD_F))
    print(f"[mixed-2d] deformed: force max err {emax:.3e} rms {erms:.3e}", flush=True)
    assert 0.0 < emax < 1e-4 and erms < 1e-5
    g.integrate_cd(200)
    r.integrate_cd(200)
    emax, erms = _errors(g.field(capi.FIELD_F), r.field(capi.FIELD_F))
    print(f"[mixed-2d] after 200 steps: force max err {emax:.3e} rms {erms:.3e}", flush=True)
    assert emax < 1e-4 and erms < 1e-5


@pytest.mark.parametrize("which", ["packing", "hbm_wall", "ensemble"])
def test_state_fusion_is_bit_exact(which):
    """The fused state-based kernel (whole particles in one workgroup:
    dilation, barrier, force) against the two-pass kernels: identical forces,
    dilations, bond states and damage after deforming steps, eager and
    captured."""
    if which == "packing":
        spec = configs.packing_3d(n_spheres=6, nodes_per_sphere=300, L=5.0e-3, seed=2, gap_factor=1.0,
                                  init_speed=2.0)
    elif which == "hbm_wall":
        spec = configs.packing_3d(n_spheres=2, nodes_per_sphere=1000, L=20.0e-3, seed=2, gap_factor=1.0,
                                  init_speed=2.0)
    else:
        spec = configs.ensemble_2d(n_instances=4, n_min=2, n_max=6, ratios=(6, 10), speed_lo=2.0,
                                   speed_hi=4.0, pitch_factor=1.0, R_lo=1.0e-3, R_hi=1.0e-3, seed=5)
    out = []
    for on in (True, False):
        g = capi.GpuEngine(spec, capi.PREC_MIXED)
        g.set_state_fusion(on)
        g.integrate_begin()
        u = _smooth_displacement(spec.x_ref, 3.0e-7, 2.0e-3)
        g.set_field(capi.FIELD_U, u)
        g.set_field(capi.FIELD_X, spec.x_ref + u)
        g.integrate_cd(20)
        g.integrate_cd_graph(40)
        out.append({f: g.field(f) for f in (capi.FIELD_F, capi.FIELD_THETA, capi.FIELD_BONDS, capi.FIELD_Z,
                                              capi.FIELD_U)})
        print(f"[state-fusion] {which} fused={on} tiles {g.tile_counts()} config {g.launch_config()}",
              flush=True)
        g.close()
    assert np.abs(out[0][capi.FIELD_F]).max() > 0.0
    for k in out[0]:
        np.testing.assert_array_equal(out[0][k], out[1][k])


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


def test_mixed_tiles_stage_in_lds():
    """Spheres of ~300 nodes fit the LDS window whatever the tile size: every
    tile of the mixed bond kernels stages its neighbor window in LDS, at the
    default shape, at each shape set explicitly and after the autotuner (a
    tile planned from stale neighbor ranges would fall back to HBM gathers).
    Results do not depend on the shape."""
    spec = configs.packing_3d(n_spheres=6, nodes_per_sphere=300, L=5.0e-3, seed=2,
                              gap_factor=1.0, init_speed=2.0)
    g = capi.GpuEngine(spec, capi.PREC_MIXED)
    g.integrate_begin()
    g.integrate_cd(5)
    n_state = sum(1 for p in spec.particles if p["compute_force"])
    lds, hbm = g.tile_counts()
    print(f"[tiles] default {g.launch_config()}: lds {lds} hbm {hbm}", flush=True)
    assert hbm == 0 and lds >= n_state
    g.compute_forces()
    ref = g.field(capi.FIELD_F)
    for tpb, tn in ((256, 256), (1024, 512), (1024, 1024)):
        g.set_launch_config(256, tpb, tn)
        lds, hbm = g.tile_counts()
        print(f"[tiles] mixed_tpb {tpb} tile {tn}: lds {lds} hbm {hbm}", flush=True)
        assert hbm == 0 and lds >= n_state
        g.compute_forces()
        np.testing.assert_array_equal(g.field(capi.FIELD_F), ref)
    g.autotune(reps=2)
    lds, hbm = g.tile_counts()
    print(f"[tiles] autotuned {g.launch_config()}: lds {lds} hbm {hbm}", flush=True)
    assert hbm == 0 and lds >= n_state



def test_mixed_hbm_tiles_against_oracle():
    """A model whose wall is numbered so that its neighbor windows do not fit
    LDS: the wall's tiles gather from HBM on the side stream, beside the LDS
    tiles of the spheres.  A smooth non-rigid displacement field (strains up
    to ~1e-3) is imposed on every body so that all bonds, the wall's
    included, carry force; then forces within the FP32 bond arithmetic bar of
    the oracle after eager steps, and the captured step equal to the eager one
    bit for bit (the side stream is forked and joined inside the graph)."""
    spec = configs.packing_3d(n_spheres=2, nodes_per_sphere=1000, L=20.0e-3, seed=2,
                              gap_factor=1.0, init_speed=2.0)
    steps = 5
    a = capi.GpuEngine(spec, capi.PREC_MIXED)
    b = capi.GpuEngine(spec, capi.PREC_MIXED)
    lds, hbm = a.tile_counts()
    print(f"[hbm-tiles] nodes={spec.n_nodes} lds {lds} hbm {hbm}", flush=True)
    assert hbm > 0 and lds > 0
    r = oracle.OracleEngine(spec, 8)
    xr = spec.x_ref
    lam = 2.0e-3
    u = np.zeros_like(xr)
    for d in range(3):
        u[:, d] = 3.0e-7 * np.sin(2 * np.pi * xr[:, (d + 1) % 3] / lam + d)
    for e in (a, b, r):
        e.integrate_begin()
        e.set_field(capi.FIELD_U, u)
        e.set_field(capi.FIELD_X, xr + u)
        e.compute_forces()
    wall = slice(spec.particles[-1]["glob_start"], spec.particles[-1]["glob_end"])
    f0 = r.field(capi.FIELD_F)
    emax, erms = _errors(a.field(capi.FIELD_F), f0)
    print(f"[hbm-tiles] deformed: max |f| wall {np.abs(f0[wall]).max():.3e}, "
          f"force max err {emax:.3e} rms {erms:.3e}", flush=True)
    assert np.abs(f0[wall]).max() > 1e-3 * np.abs(f0).max()
    assert 0 < emax < 1e-4 and erms < 1e-5
    a.integrate_cd(steps)
    b.integrate_cd_graph(steps)
    r.integrate_cd(steps)
    emax, erms = _errors(a.field(capi.FIELD_F), r.field(capi.FIELD_F))
    print(f"[hbm-tiles] after {steps} steps: force max err {emax:.3e} rms {erms:.3e}", flush=True)
    assert emax < 1e-4 and erms < 1e-5
    np.testing.assert_array_equal(a.field(capi.FIELD_F), b.field(capi.FIELD_F))
    np.testing.assert_array_equal(a.field(capi.FIELD_U), b.field(capi.FIELD_U))
