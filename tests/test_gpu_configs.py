"""The benchmark configurations at their own sizes, on the GPU, against the
oracle.

  C1  2D compressive test, 500 particles x ~400 nodes (198k nodes), FP64:
      from t = 0, state after each of 10 chunks; and from the compacted
      state (the top rows packed against the wall, contacts everywhere), the
      per-step reaction force on the moving wall (demModel.cpp:2202-2250),
      broken-bond count and largest damage (:770-773) and the load carried
      by the particles over 1000 steps at 1e-9 of each curve's largest
      value, plus the full state at the end.
  C2  3D packing, 2000 spheres x ~1000 nodes (2.08M nodes), mixed precision:
      max and RMS force error against the FP64 oracle over 20 steps and on a
      deformed state.
  C3  batched ensemble, 512 instances (4.19M nodes): the batch equals
      instances run alone bit for bit (FP64) for a sample of instances;
      mixed precision against the oracle on a subset of instances.
  C4  the closure's gating at full C1 size on the compacted state: without a
      network every pair keeps the plain kernels, bit for bit.
"""
import numpy as np
import pytest

from peridem_amd import capi, configs
import oracle

from parity_util import compare_state, same_sets

pytestmark = pytest.mark.gpu

NTHREADS = 16


def _force_errors(fg, fr):
    den = np.abs(fr).max()
    d = np.abs(fg - fr)
    return float(d.max() / den), float(np.sqrt((d ** 2).sum(1).mean()) / den)


@pytest.fixture(scope="module")
def c1_spec():
    return configs.compressive_2d(n_particles=500, nodes_per_particle=400, seed=0)


@pytest.fixture(scope="module")
def c1_compacted(c1_spec):
    print(f"[C1] preparing the compacted state: wall at {configs.COMPACT_PREP_SPEED} m/s for "
          f"{configs.COMPACT_PREP_STEPS} steps", flush=True)
    st = configs.compacted_state(c1_spec)
    print(f"[C1] compacted: restart step {st['step']}, wall penetration {st['penetration']:.4e}, "
          f"reaction {st['reaction']:.4e}, broken bonds {st['broken_bonds']}", flush=True)
    return st


def test_c1_full_size_from_rest(c1_spec):
    spec = c1_spec
    print(f"[C1] nodes={spec.n_nodes} particles={spec.n_particles}", flush=True)
    g = capi.GpuEngine(spec, capi.PREC_FP64)
    r = oracle.OracleEngine(spec, NTHREADS)
    a, b = g.neighbors(capi.NEIGH_PD), r.neighbors(capi.NEIGH_PD)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    g.integrate_begin()
    r.integrate_begin()
    for k in range(10):
        g.integrate_cd_graph(100)
        r.integrate_cd(100)
        errs = compare_state(g, r, tag=f"C1 chunk {k}")
        assert same_sets(*g.neighbors(capi.NEIGH_CONTACT), *r.neighbors(capi.NEIGH_CONTACT))
        sg, sr = g.contact_state(), r.contact_state()
        for key in ("interval", "counter", "n_rebuilds"):
            assert sg[key] == sr[key], (key, sg[key], sr[key])
        print(f"[C1] step {(k + 1) * 100}: {errs}", flush=True)
    pg = g.pp_compressive_test(spec.wall_id, spec.f_dir)
    pr = r.pp_compressive_test(spec.wall_id, spec.f_dir)
    for a_, b_ in zip(pg, pr):
        assert abs(a_ - b_) <= 1e-9 * max(abs(b_), 1e-300)


def test_c1_compacted_curves_1000_steps(c1_spec, c1_compacted):
    """Per-step curves over 1000 steps from the compacted state: the
    reference's wall reaction (ppCompressiveTest: PD and body forces on the
    wall's nodes only, since wall nodes take no contact force,
    demModel.cpp:1725-1726; constant while the wall moves rigidly), the
    broken-bond count, the largest damage Z, and the load the particles
    carry, sum over non-wall nodes of f_y V (contact, damping and bond forces
    of the compacted packing)."""
    spec, st = c1_spec, c1_compacted
    g = capi.GpuEngine(spec, capi.PREC_FP64)
    r = oracle.OracleEngine(spec, NTHREADS)
    for e in (g, r):
        configs.apply_state(e, st)
        e.integrate_begin()
    wall = np.array([p["type_index"] == 1 for p in spec.particles])[spec.pt_id]
    vol = spec.vol
    curves = {"reaction": ([], []), "broken": ([], []), "zmax": ([], []), "load": ([], [])}
    for n in range(1000):
        for k, e in enumerate((g, r)):
            e.integrate_cd(1)
            curves["reaction"][k].append(e.pp_compressive_test(spec.wall_id, spec.f_dir)[1])
            curves["broken"][k].append(e.broken_bonds()[0])
            curves["zmax"][k].append(float(e.field(capi.FIELD_Z).max()))
            f = e.field(capi.FIELD_F)
            curves["load"][k].append(float((f[~wall, 1] * vol[~wall]).sum()))
        if n % 250 == 249:
            print(f"[C1-compacted] step {n + 1}: load {curves['load'][1][-1]:.6e} zmax {curves['zmax'][1][-1]:.4f} "
                  f"broken {curves['broken'][1][-1]} candidates {r.neighbor_count(capi.NEIGH_CONTACT)}", flush=True)
    for name, (a, b) in curves.items():
        a, b = np.array(a, np.float64), np.array(b, np.float64)
        err = float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))
        print(f"[C1-compacted] {name}: range {b.min():.6e}..{b.max():.6e}, max rel err {err:.3e}", flush=True)
        assert err <= 1e-9, name
    assert curves["broken"][0] == curves["broken"][1]
    # the particles carry a load that changes over the window (contacts everywhere)
    load = np.array(curves["load"][1])
    assert load.max() - load.min() > 1e-3 * np.abs(load).max()
    zmax = np.array(curves["zmax"][1])
    assert zmax.max() > 0.0
    assert r.neighbor_count(capi.NEIGH_CONTACT) > 0
    assert same_sets(*g.neighbors(capi.NEIGH_CONTACT), *r.neighbors(capi.NEIGH_CONTACT))
    # The FP64 kernels evaluate the RNP law's exp with the device library,
    # which can differ from the host's in the last ulp; in the dense packing
    # such a difference grows over the steps (chaotic contact dynamics), so
    # after 1000 steps the end state agrees to about 1e-10, not bit for bit.
    out = compare_state(g, r, tol=1e-8, tag="C1 compacted, step 1000")
    print(f"[C1-compacted] end state after 1000 steps: {out}", flush=True)
    # Per-step forces: the oracle takes the GPU's state and both compute the
    # forces of that state: equal to 1e-12 of the largest force.
    for fid in (capi.FIELD_X, capi.FIELD_U, capi.FIELD_V, capi.FIELD_BONDS):
        r.set_field(fid, g.field(fid))
    g.compute_forces()
    r.compute_forces()
    fg, fr = g.field(capi.FIELD_F), r.field(capi.FIELD_F)
    ef = float(np.abs(fg - fr).max() / np.abs(fr).max())
    print(f"[C1-compacted] forces of one state: max |f| {np.abs(fr).max():.4e}, rel err {ef:.3e}", flush=True)
    assert ef <= 1e-12


def test_c2_full_size_mixed_force_error():
    spec = configs.packing_3d(n_spheres=2000, nodes_per_sphere=1000, seed=0)
    print(f"[C2] nodes={spec.n_nodes} spheres={spec.n_particles - 2}", flush=True)
    g = capi.GpuEngine(spec, capi.PREC_MIXED)
    r = oracle.OracleEngine(spec, NTHREADS)
    g.integrate_begin()
    r.integrate_begin()
    out = []
    for k in range(4):
        g.integrate_cd(5)
        r.integrate_cd(5)
        emax, erms = _force_errors(g.field(capi.FIELD_F), r.field(capi.FIELD_F))
        out.append((emax, erms))
        print(f"[C2] step {(k + 1) * 5}: force max err {emax:.3e} rms {erms:.3e}", flush=True)
        assert emax < 1e-4 and erms < 1e-5
    assert np.array_equal(g.field(capi.FIELD_BONDS), r.field(capi.FIELD_BONDS))
    # a deformed state: every bond carries force (peak strain ~1e-3)
    xr = spec.x_ref
    u = np.zeros_like(xr)
    for d in range(3):
        u[:, d] = 3.0e-7 * np.sin(2 * np.pi * xr[:, (d + 1) % 3] / 2.0e-3 + d)
    for e in (g, r):
        e.set_field(capi.FIELD_U, r.field(capi.FIELD_U) + u)
        e.set_field(capi.FIELD_X, r.field(capi.FIELD_X) + u)
        e.compute_peridynamic_forces()
    emax, erms = _force_errors(g.field(capi.FIELD_F), r.field(capi.FIELD_F))
    th, thr = g.field(capi.FIELD_THETA), r.field(capi.FIELD_THETA)
    eth = float(np.abs(th - thr).max() / np.abs(thr).max())
    print(f"[C2] deformed: max |f| {np.abs(r.field(capi.FIELD_F)).max():.3e} force max err {emax:.3e} "
          f"rms {erms:.3e} theta err {eth:.3e}", flush=True)
    assert 0.0 < emax < 1e-4 and erms < 1e-5 and eth < 1e-5


def test_c3_full_size_ensemble():
    spec = configs.ensemble_2d(n_instances=512, seed=0)
    K = 200
    print(f"[C3] nodes={spec.n_nodes} particles={spec.n_particles} instances=512 steps={K}", flush=True)
    g = capi.GpuEngine(spec, capi.PREC_FP64)
    g.integrate_begin()
    g.integrate_cd_graph(K)
    F, U, V = g.field(capi.FIELD_F), g.field(capi.FIELD_U), g.field(capi.FIELD_V)
    g.close()
    rng = np.random.default_rng(1)
    sample = sorted(set(rng.choice(512, 5, replace=False).tolist()) | {0, 511})
    for k in sample:
        s = configs.ensemble_instance(spec, k)
        e = capi.GpuEngine(s, capi.PREC_FP64)
        e.integrate_begin()
        e.integrate_cd_graph(K)
        first, last, _ = spec.info["instances"][k]
        n0, n1 = spec.particles[first]["glob_start"], spec.particles[last - 1]["glob_end"]
        np.testing.assert_array_equal(e.field(capi.FIELD_F), F[n0:n1])
        np.testing.assert_array_equal(e.field(capi.FIELD_U), U[n0:n1])
        np.testing.assert_array_equal(e.field(capi.FIELD_V), V[n0:n1])
        e.close()
        print(f"[C3] instance {k} ({last - first} particles) equals the batch", flush=True)
    # mixed precision on a subset of instances against the oracle: on a
    # deformed state (every bond loaded) and after 2000 steps of collisions
    for k in sample[:3]:
        s = configs.ensemble_instance(spec, k)
        m = capi.GpuEngine(s, capi.PREC_MIXED)
        r = oracle.OracleEngine(s, NTHREADS)
        xr = s.x_ref
        u = np.zeros_like(xr)
        for d in range(2):
            u[:, d] = 2.0e-7 * np.sin(2 * np.pi * xr[:, (d + 1) % 2] / 1.5e-3 + d)
        for e in (m, r):
            e.integrate_begin()
            e.set_field(capi.FIELD_U, u)
            e.set_field(capi.FIELD_X, xr + u)
            e.compute_peridynamic_forces()
        fr = r.field(capi.FIELD_F)
        emax, erms = _force_errors(m.field(capi.FIELD_F), fr)
        print(f"[C3] instance {k} mixed vs oracle, deformed: max |f| {np.abs(fr).max():.3e} "
              f"force max err {emax:.3e} rms {erms:.3e}", flush=True)
        assert 0.0 < emax < 1e-4 and erms < 1e-5
        m.close()
        r.close()
        m = capi.GpuEngine(s, capi.PREC_MIXED)
        r = oracle.OracleEngine(s, NTHREADS)
        for e in (m, r):
            e.integrate_begin()
        m.integrate_cd_graph(2000)
        r.integrate_cd(2000)
        fr = r.field(capi.FIELD_F)
        emax, erms = _force_errors(m.field(capi.FIELD_F), fr)
        print(f"[C3] instance {k} mixed vs oracle after 2000 steps: max |f| {np.abs(fr).max():.3e} "
              f"force max err {emax:.3e} rms {erms:.3e}", flush=True)
        if np.abs(fr).max() > 1e3:  # particles in contact: bond forces to compare
            assert emax < 1e-4 and erms < 1e-5


def test_c4_closure_gating_full_size(c1_spec, c1_compacted):
    spec, st = c1_spec, c1_compacted
    runs = []
    for gated in (False, True):
        e = capi.GpuEngine(spec, capi.PREC_FP64)
        if gated:
            e.closure_setup()
            e.closure_enable(None)
        configs.apply_state(e, st)
        e.integrate_begin()
        e.integrate_cd_graph(300)
        runs.append({f: e.field(f) for f in (capi.FIELD_F, capi.FIELD_U, capi.FIELD_V)})
        if gated:
            stats = e.closure_stats()
            print(f"[C4] gating without a network at full size: {stats}", flush=True)
            assert stats["closure"] == 0 and stats["plain"] > 0
        e.close()
    for f in runs[0]:
        np.testing.assert_array_equal(runs[0][f], runs[1][f])
