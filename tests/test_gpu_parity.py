"""GPU engine vs the FP64 oracle through the C-ABI, step by step.

Bar (FP64 parity mode): neighbor and contact-candidate sets identical, bond
states identical, forces / displacements / velocities / positions within
1e-12 of the field's largest entry, damage within its 1e-5 slack, and the
compressive-test observables within 1e-9.
"""
import numpy as np
import pytest

from peridem_amd import capi, configs
import oracle

from parity_util import compare_state, rel_err, same_sets

pytestmark = pytest.mark.gpu


def _pair(spec, nthreads=8):
    assert capi.device_count() > 0
    return capi.GpuEngine(spec), oracle.OracleEngine(spec, nthreads)


def _run_compare(spec, chunks, chunk_steps, tag):
    print(f"[{tag}] nodes={spec.n_nodes} particles={spec.n_particles}", flush=True)
    g, r = _pair(spec)
    rp_g, c_g = g.neighbors(capi.NEIGH_PD)
    rp_r, c_r = r.neighbors(capi.NEIGH_PD)
    assert np.array_equal(rp_g, rp_r) and np.array_equal(c_g, c_r), "PD neighbor lists differ"
    np.testing.assert_array_equal(g.field(capi.FIELD_MX), r.field(capi.FIELD_MX))
    g.integrate_begin()
    r.integrate_begin()
    compare_state(g, r, tag=f"{tag} n=0")
    for k in range(chunks):
        g.integrate_cd(chunk_steps)
        r.integrate_cd(chunk_steps)
        errs = compare_state(g, r, tag=f"{tag} chunk {k}")
        if spec.multi_particle:
            a = g.neighbors(capi.NEIGH_CONTACT)
            b = r.neighbors(capi.NEIGH_CONTACT)
            assert same_sets(*a, *b), f"{tag}: contact candidate sets differ"
            sg, sr = g.contact_state(), r.contact_state()
            for key in ("interval", "counter", "n_rebuilds"):
                assert sg[key] == sr[key], f"{tag}: controller {key} {sg[key]} != {sr[key]}"
        print(f"[{tag}] step {(k + 1) * chunk_steps}: {errs}", flush=True)
    assert g.time() == r.time()
    bg, br = g.broken_bonds(), r.broken_bonds()
    assert bg == br
    return g, r


def test_two_particle_rnp_fp64():
    # config-0 geometry at reduced size: the collision starts inside the run
    spec = configs.two_particle(R=0.5e-3, h=1.0e-4, gap=-1.0e-4, speed=1.0)
    g, r = _run_compare(spec, chunks=8, chunk_steps=50, tag="two_particle")
    assert np.abs(r.field(capi.FIELD_F)).max() > 0.0


def test_two_particle_bond_breaking():
    # fast impact with full contact stiffness: bonds break near the contact
    spec = configs.two_particle(R=0.5e-3, h=1.0e-4, gap=-2.0e-5, speed=100.0, dt=1e-9,
                                kn_factor=1.0)
    g, r = _run_compare(spec, chunks=4, chunk_steps=100, tag="breaking")
    b, t = r.broken_bonds()
    print("broken bonds", b, "of", t)
    assert b > 100


def test_compressive_small_fp64():
    spec = configs.compressive_2d(n_particles=12, nodes_per_particle=120, seed=3, interval=10,
                                  wall_speed=2.0, R_lo=1e-3, R_hi=1e-3, pitch_factor=1.03,
                                  init_speed=2.0, dt=2e-8, friction=True)
    g, r = _run_compare(spec, chunks=5, chunk_steps=40, tag="compressive12")
    assert r.neighbor_count(capi.NEIGH_CONTACT) > 0
    pg = g.pp_compressive_test(spec.wall_id, spec.f_dir)
    pr = r.pp_compressive_test(spec.wall_id, spec.f_dir)
    assert abs(pg[0] - pr[0]) <= 1e-9 * max(abs(pr[0]), 1e-30)
    assert abs(pg[1] - pr[1]) <= 1e-9 * max(abs(pr[1]), 1e-30)


def test_compressive_rebuild_every_step():
    # interval 1: the controller pins the radius to the contact radius and
    # rebuilds every step
    spec = configs.compressive_2d(n_particles=9, nodes_per_particle=100, seed=4, interval=1,
                                  wall_speed=1.0, R_lo=1e-3, R_hi=1e-3, pitch_factor=1.03,
                                  init_speed=1.0, dt=2e-8, friction=True)
    _run_compare(spec, chunks=3, chunk_steps=30, tag="interval1")


def test_packing_3d_state_fp64():
    spec = configs.packing_3d(n_spheres=6, nodes_per_sphere=150, L=5.0e-3, seed=1,
                              gap_factor=1.0, init_speed=5.0)
    g, r = _run_compare(spec, chunks=3, chunk_steps=20, tag="packing3d")
    assert r.neighbor_count(capi.NEIGH_CONTACT) > 0


def test_graph_replay_matches_eager():
    spec = configs.compressive_2d(n_particles=12, nodes_per_particle=120, seed=5, interval=10,
                                  wall_speed=2.0, friction=True)
    a = capi.GpuEngine(spec)
    b = capi.GpuEngine(spec)
    a.integrate_begin()
    b.integrate_begin()
    for _ in range(4):
        a.integrate_cd(25)
        b.integrate_cd_graph(25)
        for fid in (capi.FIELD_F, capi.FIELD_X, capi.FIELD_V):
            np.testing.assert_array_equal(a.field(fid), b.field(fid))
    assert a.time() == b.time()


@pytest.mark.parametrize("which", ["compressive", "packing", "packing_2m"])
def test_step_fusion_is_bit_exact(which):
    """Fused step kernels (node update with BC and centre data, one-workgroup
    controller, damping pass finalising the nodes) against the unfused
    launches, graph and eager: identical fields after contact, damping,
    gravity and the moving wall (compressive) or a 3D mixed-precision run.
    Above 2^20 nodes (the 2.08M-node packing) the fused node update and
    controller run with the finalisation deferred to the next node update."""
    if which == "compressive":
        spec = configs.compressive_2d(n_particles=12, nodes_per_particle=120, seed=5, interval=10,
                                      wall_speed=2.0, init_speed=1.0, friction=True)
        prec = capi.PREC_FP64
    elif which == "packing_2m":
        spec = configs.packing_3d(init_speed=2.0)
        prec = capi.PREC_MIXED
    else:
        spec = configs.packing_3d(n_spheres=6, nodes_per_sphere=200, L=5.0e-3, seed=3, gap_factor=1.0,
                                  init_speed=2.0)
        prec = capi.PREC_MIXED
    out = []
    for on in (True, False):
        g = capi.GpuEngine(spec, prec)
        g.set_step_fusion(on)
        g.integrate_begin()
        g.integrate_cd(60)
        g.integrate_cd_graph(240)
        out.append({f: g.field(f) for f in (capi.FIELD_F, capi.FIELD_U, capi.FIELD_V, capi.FIELD_VMAG, capi.FIELD_Z)}
                   | {"t": g.time(), "nc": g.neighbor_count(capi.NEIGH_CONTACT)})
        g.close()
    for k in out[0]:
        if k in ("t", "nc"):
            assert out[0][k] == out[1][k]
        else:
            np.testing.assert_array_equal(out[0][k], out[1][k])


def test_two_particle_test_observables():
    spec = configs.two_particle(gap=0.0, speed=0.5)
    g = capi.GpuEngine(spec)
    r = oracle.OracleEngine(spec, 8)
    g.integrate_begin()
    r.integrate_begin()
    for k in range(4):
        g.integrate_cd_graph(60)
        r.integrate_cd(60)
        a, b = g.pp_two_particle_test(), r.pp_two_particle_test()
        print(f"[pp2] step {(k + 1) * 60}: {a}", flush=True)
        for key in b:
            assert abs(a[key] - b[key]) <= 1e-12 * max(abs(b[key]), 1e-30), (key, a[key], b[key])
    assert a["pen_dist"] < 0 and a["contact_area_radius"] > 0


def test_check_stop_quantities():
    spec = configs.compressive_2d(n_particles=12, nodes_per_particle=120, seed=5, interval=10, wall_speed=2.0,
                                  init_speed=1.0, friction=True)
    g = capi.GpuEngine(spec)
    r = oracle.OracleEngine(spec, 8)
    g.integrate_begin()
    r.integrate_begin()
    g.integrate_cd(150)
    r.integrate_cd(150)
    for crit, param in ((capi.STOP_MAX_PARTICLE_DIST, 1e-3), (capi.STOP_MAX_NODE_DIST, 1e-2)):
        a, b = g.check_stop(crit, param), r.check_stop(crit, param)
        print(f"[stop] criterion {crit}: gpu {a} oracle {b}", flush=True)
        assert a == b


def test_deferred_finalisation_is_bit_exact():
    """Unfused steps leave damping and gravity to the next node update (forces
    pending); reads of f, checkpoints, partial force calls and Verlet steps
    finalise them first.  Against the fused step (damping pass finalising the
    nodes), bit for bit: fields after eager and replayed steps, a checkpoint
    taken with forces pending and resumed in another context, and Verlet
    steps after a deferred start."""
    spec = configs.packing_3d(n_spheres=6, nodes_per_sphere=200, L=5.0e-3, seed=3, gap_factor=1.0, init_speed=2.0)
    fields = (capi.FIELD_F, capi.FIELD_U, capi.FIELD_V)

    def run(mode, verlet=False):
        g = capi.GpuEngine(spec, capi.PREC_MIXED)
        g.set_step_fusion(mode)
        g.integrate_begin()
        out = []
        if verlet:
            g.integrate_verlet(40)
            out.append([g.field(f) for f in fields])
            g.close()
            return out, None
        g.integrate_cd(30)
        out.append([g.field(f) for f in fields])
        g.integrate_cd_graph(30)
        blob = g.checkpoint()
        g.integrate_cd_graph(30)
        out.append([g.field(f) for f in fields])
        pg = g.pp_compressive_test(0, 2)
        g.close()
        return out, (blob, pg)

    fused, (_, pg_f) = run(True)
    unfused, (blob, pg_u) = run(False)
    for a, b in zip(fused, unfused):
        for x, y in zip(a, b):
            np.testing.assert_array_equal(x, y)
    assert pg_f == pg_u
    r = capi.GpuEngine(spec, capi.PREC_MIXED)
    r.set_step_fusion(False)
    r.restore(blob)
    r.integrate_cd_graph(30)
    for x, y in zip([r.field(f) for f in fields], unfused[1]):
        np.testing.assert_array_equal(x, y)
    r.close()
    vf, _ = run(True, verlet=True)
    vu, _ = run(False, verlet=True)
    for x, y in zip(vf[0], vu[0]):
        np.testing.assert_array_equal(x, y)


@pytest.mark.parametrize("prec", [capi.PREC_FP64, capi.PREC_MIXED])
def test_peridynamic_forces_leave_other_nodes(prec):
    """computePeridynamicForces alone (demModel.cpp:602) writes f of the nodes
    of force-computing particles only: the walls keep the forces of the last
    step (contact, damping, gravity), also when the engine holds them pending
    after unfused steps."""
    spec = configs.packing_3d(n_spheres=6, nodes_per_sphere=150, L=5.0e-3, seed=1, gap_factor=1.0,
                              init_speed=5.0)
    g = capi.GpuEngine(spec, prec)
    g.set_step_fusion(False)  # forces pending after every step
    r = oracle.OracleEngine(spec, 8)
    for e in (g, r):
        e.integrate_begin()
        e.integrate_cd(30)
    walls = np.concatenate([np.arange(p["glob_start"], p["glob_end"]) for p in spec.particles
                            if not p["compute_force"]])
    assert walls.size > 0
    f_before = r.field(capi.FIELD_F)[walls]
    for e in (g, r):
        e.compute_peridynamic_forces()
    fg, fr = g.field(capi.FIELD_F), r.field(capi.FIELD_F)
    np.testing.assert_array_equal(fr[walls], f_before)
    print(f"[pd-only] prec {prec}: max |f| walls {np.abs(fr[walls]).max():.3e} all {np.abs(fr).max():.3e}",
          flush=True)
    assert np.abs(fr[walls]).max() > 0.0
    bar = 1e-12 if prec == capi.PREC_FP64 else 1e-4
    assert np.abs(fg - fr).max() <= bar * np.abs(fr).max()
    assert np.abs(fg[walls] - fr[walls]).max() <= 1e-12 * np.abs(fr).max()


def test_contact_skin_after_positions_set():
    """The contact loops skip nodes whose candidates cannot be in contact,
    from each node's nearest candidate at the last rebuild and a bound of the
    node moves since.  Positions set from the host between rebuilds void
    that bound: a particle moved onto a candidate particle mid-interval (no
    rebuild in between) must get the same contact forces as the oracle's
    from the same candidate lists."""
    spec = configs.compressive_2d(n_particles=12, nodes_per_particle=120, seed=3, interval=10,
                                  wall_speed=2.0, R_lo=1e-3, R_hi=1e-3, pitch_factor=1.03,
                                  init_speed=0.5, dt=2e-8, friction=True)
    g, r = _pair(spec)
    for e in (g, r):
        e.integrate_begin()
        e.integrate_cd(14)

    def quiet(st):
        # the controller's next call neither rebuilds nor revisits its radius
        # (demModel.cpp:1808-1955): counter off both the interval and its
        # fifth
        iv, c = st["interval"], st["counter"]
        upi = int(0.2 * iv) if iv > 5 else 1
        return iv > 5 and c % iv != 0 and c % upi != 0

    for _ in range(40):
        if quiet(g.contact_state()):
            break
        for e in (g, r):
            e.integrate_cd(1)
    assert quiet(g.contact_state())
    rp, col = r.neighbors(capi.NEIGH_CONTACT)
    x = r.field(capi.FIELD_X)
    rc = r.contact_state()["max_contact_r"]
    ii = np.repeat(np.arange(spec.n_nodes), np.diff(rp))
    d = np.linalg.norm(x[ii] - x[col], axis=1)
    pid = spec.pt_id
    far = (d > 1.5 * rc) & (pid[ii] != pid[col])
    far &= np.array([spec.particles[p]["type_index"] == 0 for p in pid[col]])
    assert far.any()
    k = int(np.argmin(np.where(far, d, np.inf)))
    i, j = int(ii[k]), int(col[k])
    # move j's particle straight towards node i until j sits 0.5 rc from it
    shift = (x[i] - x[j]) * (1.0 - 0.5 * rc / d[k])
    q = spec.particles[pid[j]]
    sl = slice(q["glob_start"], q["glob_end"])
    u = r.field(capi.FIELD_U)
    x2, u2 = x.copy(), u.copy()
    x2[sl] += shift
    u2[sl] += shift
    before = g.contact_state()["n_rebuilds"]
    for e in (g, r):
        e.set_field(capi.FIELD_X, x2)
        e.set_field(capi.FIELD_U, u2)
        e.compute_forces()
    assert g.contact_state()["n_rebuilds"] == before  # the lists of the last rebuild
    errs = compare_state(g, r, tag="skin")
    print(f"[skin] node {i} particle {pid[i]}, candidate {j} particle {pid[j]} moved by "
          f"{np.linalg.norm(shift):.3e} (contact radius {rc:.3e}): {errs}", flush=True)
    fi = r.field(capi.FIELD_F)[i]
    assert np.abs(fi).max() > 0.0
