"""Checkpoint and restore of the complete step state: a context restored from
a checkpoint taken at step n and run to step m equals, bit for bit, the
context that ran from 0 to m (fields, bond bits, damage, the cadence
controller), across contact rebuilds, in FP64 and mixed precision."""
import numpy as np
import pytest

from peridem_amd import capi, configs

pytestmark = pytest.mark.gpu

FIELDS = (capi.FIELD_X, capi.FIELD_U, capi.FIELD_V, capi.FIELD_F, capi.FIELD_Z, capi.FIELD_THETA, capi.FIELD_BONDS)


@pytest.mark.parametrize("which", ["compressive_fp64", "packing_mixed"])
def test_restore_continues_bit_for_bit(which):
    if which == "compressive_fp64":
        spec = configs.compressive_2d(n_particles=12, nodes_per_particle=150, seed=3, init_speed=1.0,
                                      pitch_factor=1.01, R_lo=1.2e-3, R_hi=1.2e-3, friction=True)
        prec = capi.PREC_FP64
    else:
        spec = configs.packing_3d(n_spheres=6, nodes_per_sphere=300, L=5.0e-3, seed=2, gap_factor=1.0,
                                  init_speed=2.0)
        prec = capi.PREC_MIXED
    a = capi.GpuEngine(spec, prec)
    a.integrate_begin()
    a.integrate_cd_graph(170)
    blob = a.checkpoint()
    a.integrate_cd_graph(230)
    b = capi.GpuEngine(spec, prec)
    b.integrate_begin()
    b.restore(blob)
    b.integrate_cd_graph(230)
    print(f"[ckpt] {which}: {len(blob)} bytes, state {a.contact_state()}", flush=True)
    for fid in FIELDS:
        np.testing.assert_array_equal(a.field(fid), b.field(fid))
    assert a.contact_state() == b.contact_state()
    assert a.time() == b.time()
    assert a.contact_state()["n_rebuilds"] > 3
    # a checkpoint of another model is rejected
    other = capi.GpuEngine(configs.two_particle(), capi.PREC_FP64)
    with pytest.raises(RuntimeError, match="checkpoint"):
        other.restore(blob)


def test_restart_from_vtu_matches_oracle(tmp_path):
    """DEMModel::restart from a VTU (positions, displacement, velocity; step
    counter, time and the contact-search controller's restart rule): the GPU
    and the oracle restarted from the same file at step 120 stay equal."""
    from peridem_amd import output
    import oracle
    spec = configs.compressive_2d(n_particles=12, nodes_per_particle=120, seed=5, interval=10, wall_speed=2.0,
                                  init_speed=1.0, friction=True)
    a = oracle.OracleEngine(spec, 8)
    a.integrate_begin()
    a.integrate_cd(120)
    path = str(tmp_path / "output_3.vtu")
    output.write_vtu(path, a, spec)
    g = capi.GpuEngine(spec, capi.PREC_FP64)
    r = oracle.OracleEngine(spec, 8)
    for e in (g, r):
        output.restart_from_vtu(e, spec, path, 120)
        e.integrate_begin()
        e.integrate_cd(200)
    st_g, st_r = g.contact_state(), r.contact_state()
    print(f"[restart] gpu {st_g} oracle {st_r}", flush=True)
    for k in ("interval", "counter", "n_rebuilds"):
        assert st_g[k] == st_r[k]
    for fid in (capi.FIELD_F, capi.FIELD_U, capi.FIELD_V, capi.FIELD_X):
        np.testing.assert_array_equal(g.field(fid), r.field(fid))
    assert g.time() == r.time() and g.time()[0] == 320
