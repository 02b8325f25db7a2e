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
                                      pitch_factor=1.01, R_lo=1.2e-3, R_hi=1.2e-3)
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
