"""VTU and compressive-test CSV output (peridem_amd/output.py), on the oracle's
state: the arrays read back equal the state bit for bit (17 significant
digits), with the reference's array names and component counts."""
import numpy as np
import pytest

from peridem_amd import capi, configs, output
import oracle


def test_vtu_round_trip(tmp_path):
    spec = configs.two_particle(gap=0.0, speed=0.1, dt=1.0e-7)
    o = oracle.OracleEngine(spec, 4)
    o.integrate_begin()
    o.integrate_cd(15)
    path = tmp_path / "out_0.vtu"
    output.write_vtu(str(path), o, spec)
    pts, data = output.read_vtu(str(path))
    np.testing.assert_array_equal(pts, o.field(capi.FIELD_X))
    np.testing.assert_array_equal(data["Displacement"], o.field(capi.FIELD_U))
    np.testing.assert_array_equal(data["Velocity"], o.field(capi.FIELD_V))
    np.testing.assert_array_equal(data["Force_Density"], o.field(capi.FIELD_F))
    np.testing.assert_array_equal(data["Force"], o.field(capi.FIELD_F) * spec.vol[:, None])
    np.testing.assert_array_equal(data["Damage_Z"], o.field(capi.FIELD_Z).astype(np.float64))
    np.testing.assert_array_equal(data["Particle_ID"], spec.pt_id)
    np.testing.assert_array_equal(data["Nodal_Volume"], spec.vol)
    assert set(data) == set(output.DEFAULT_TAGS)
    assert data["Displacement"].shape == (spec.n_nodes, 3) and data["Theta"].shape == (spec.n_nodes,)


def test_compressive_csv(tmp_path):
    rows = [(1e-7 * k, -1e-9 * k, 0.5 * k) for k in range(5)]
    path = tmp_path / "pp.csv"
    output.write_compressive_csv(str(path), rows)
    back = np.loadtxt(str(path), delimiter=",", skiprows=1)
    np.testing.assert_array_equal(back, np.array(rows))


def test_strain_vtu_round_trip(tmp_path):
    """Strain_Stress output: points and the 6-component arrays read back
    exactly (float32 values written with 9 significant digits)."""
    rng = np.random.default_rng(0)
    xq = rng.normal(size=(17, 3))
    eps = rng.normal(size=(17, 6)).astype(np.float32)
    sig = (rng.normal(size=(17, 6)) * 1e6).astype(np.float32)
    path = tmp_path / "output_strain_0.vtu"
    output.write_strain_vtu(str(path), xq, eps, sig)
    pts, data = output.read_vtu(str(path))
    np.testing.assert_array_equal(pts, xq)
    np.testing.assert_array_equal(data["Strain"].astype(np.float32), eps)
    np.testing.assert_array_equal(data["Stress"].astype(np.float32), sig)


@pytest.mark.parametrize("with_bc", [False, True])
def test_restart_from_vtu_continues(tmp_path, with_bc):
    """DEMModel::restart semantics on the oracle: a single elastic particle
    (no contact, no bond breaking) written to VTU at step 100 and restarted
    from it continues as the uninterrupted run: bit for bit without a
    time-dependent BC (the VTU numbers carry 17 significant digits); with a
    linear displacement BC to rounding only, because the restart's time is
    100 dt while the running time is a sum of 100 dt (as in the reference)."""
    from peridem_amd.geometry import center_node, disk_mesh
    from peridem_amd.model import BcDeck, ModelBuilder, ZoneDeck
    from peridem_amd.deck import _Region
    R = 1e-3
    x, vol = disk_mesh(R, 1e-4)
    b = ModelBuilder(2, 2e-8, [ZoneDeck(configs.pdelastic_material())], {}, multi_particle=False,
                     gravity=(0.0, -9.81, 0.0))
    b.add_particle(0, x, vol, center_node(x, (0, 0, 0), R), R)
    b.ic.append(([0], (0.3, 0.0, 0.0)))
    if with_bc:
        b.ubc.append(BcDeck([], [1], capi.UFN_LINEAR, [-2.0],
                            region=_Region("rectangle", [-R, -R, 0.0, -0.7 * R, R, 0.0], 2)))
    spec = b.build()
    a = oracle.OracleEngine(spec, 4)
    a.integrate_begin()
    a.integrate_cd(100)
    path = str(tmp_path / "output_1.vtu")
    output.write_vtu(path, a, spec)
    a.integrate_cd(100)
    r = oracle.OracleEngine(spec, 4)
    output.restart_from_vtu(r, spec, path, 100)
    n, t = r.time()
    assert n == 100 and t == 100 * spec.dt
    r.integrate_begin()
    r.integrate_cd(100)
    for fid in (capi.FIELD_X, capi.FIELD_U, capi.FIELD_V, capi.FIELD_F):
        ga, gr = a.field(fid), r.field(fid)
        if with_bc:
            np.testing.assert_allclose(gr, ga, rtol=0, atol=1e-9 * np.abs(ga).max())
        else:
            np.testing.assert_array_equal(gr, ga)
    assert np.abs(a.field(capi.FIELD_U)).max() > 0
