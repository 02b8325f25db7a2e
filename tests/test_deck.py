"""Reference input decks read unmodified (peridem_amd/deck.py).

Fixtures in tests/golden/decks are data files shipped with the reference:
apps/twoparticle_demo/example (YAML deck, two gmsh meshes, particle file) and
test/test_data/peridem/single_particle_circle/inp (YAML deck, gmsh mesh).
Checks: the FE nodal volumes against closed forms, every deck value that the
step uses against the deck text and the reference's formulas (dt, horizon,
contact radius factor times the smallest mesh size, Kn times Kn_Factor,
beta_n, friction switched off, gravity, initial velocity, the two_particle
force switch, zero-displacement fixity, region selections), the particle
transform, and a short oracle run of each deck."""
import math
import os

import numpy as np
import pytest
import yaml

from peridem_amd import capi, deck
import oracle

HERE = os.path.dirname(os.path.abspath(__file__))
DEMO = os.path.join(HERE, "golden", "decks", "twoparticle_demo", "input_0.yaml")
SINGLE = os.path.join(HERE, "golden", "decks", "single_particle_circle", "input_0.yaml")


def test_fe_volumes_closed_forms(tmp_path):
    # two triangles of a unit square: corner nodes get 1/6 or 1/3
    x = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0]], float)
    v = deck.fe_nodal_volumes(x, np.array([[0, 1, 2], [0, 2, 3]]), 2)
    np.testing.assert_allclose(v, [1 / 3, 1 / 6, 1 / 3, 1 / 6], rtol=1e-15)
    # a quadrangle (trapezoid): volumes sum to its area, symmetric pairs equal
    q = np.array([[0, 0, 0], [2, 0, 0], [1.5, 1, 0], [0.5, 1, 0]], float)
    vq = deck.fe_nodal_volumes(q, np.array([[0, 1, 2, 3]]), 2)
    assert abs(vq.sum() - 1.5) < 1e-14 and abs(vq[0] - vq[1]) < 1e-15 and abs(vq[2] - vq[3]) < 1e-15
    # a tetrahedron: a quarter of 1/6 per node
    t = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], float)
    np.testing.assert_allclose(deck.fe_nodal_volumes(t, np.array([[0, 1, 2, 3]]), 3), [1 / 24] * 4, rtol=1e-15)
    # .msh 2.2 with a point and a line element that must be skipped
    msh = tmp_path / "sq.msh"
    msh.write_text("$MeshFormat\n2.2 0 8\n$EndMeshFormat\n$Nodes\n4\n1 0 0 0\n2 1 0 0\n3 1 1 0\n4 0 1 0\n"
                   "$EndNodes\n$Elements\n4\n1 15 2 0 1 1\n2 1 2 0 1 1 2\n3 2 2 0 1 1 2 3\n4 2 2 0 1 1 3 4\n"
                   "$EndElements\n")
    xs, vs = deck.read_msh(str(msh), 2)
    np.testing.assert_allclose(vs, [1 / 3, 1 / 6, 1 / 3, 1 / 6], rtol=1e-15)


def test_demo_deck_values():
    spec = deck.load_deck(DEMO)
    cfg = yaml.safe_load(open(DEMO))
    nodes1 = deck.read_msh(os.path.join(os.path.dirname(DEMO), "mesh_cir_1_0.msh"), 2)[0]
    nodes2 = deck.read_msh(os.path.join(os.path.dirname(DEMO), "mesh_cir_2_0.msh"), 2)[0]
    assert spec.n_nodes == len(nodes1) + len(nodes2) == 1039
    assert spec.dim == 2 and spec.multi_particle
    assert spec.dt == pytest.approx(5e-3 / 50000, rel=1e-15)
    p0, p1 = spec.particles
    assert (p0["zone"], p1["zone"]) == (0, 1)
    assert p0["compute_force"] == 0 and p1["compute_force"] == 1   # two_particle test
    assert p0["horizon"] == p1["horizon"] == 6e-4
    assert p0["material_type"] == p1["material_type"] == capi.MAT_PDSTATE
    # particle 1: zone-2 mesh (R = 1 mm, reference radius 1 mm) rotated by o and moved
    site, th = np.array([0.003, 0.0076]), 1.570796
    rot = np.array([[math.cos(th), -math.sin(th)], [math.sin(th), math.cos(th)]])
    np.testing.assert_allclose(spec.x_ref[916:, :2], site + nodes2[:, :2] @ rot.T, rtol=0, atol=1e-15)
    np.testing.assert_allclose(spec.x_ref[p1["center_node"], :2], site, atol=1e-15)
    # contact table: factor x smallest mesh size, Kn x Kn_Factor, beta_n, friction off
    c01 = spec.contact[0][1]
    e = cfg["Contact"]["Zone_12"]
    assert c01["contact_r"] == pytest.approx(0.95 * spec.h_min, rel=1e-15)
    assert c01["Kn"] == e["Kn"] * e["Kn_Factor"]
    le = math.log(0.95)
    assert c01["betan"] == pytest.approx(100.0 * (-2 * le / math.sqrt(math.pi ** 2 + le * le)), rel=1e-14)
    assert c01["mu"] == 0.0
    assert spec.contact[0][0]["Kn"] == cfg["Contact"]["Zone_11"]["Kn"]
    assert spec.gravity == (0.0, -10.0, 0.0)
    np.testing.assert_array_equal(spec.v[916:], np.tile([0.0, -8.944272e-02, 0.0], (123, 1)))
    assert not spec.v[:916].any()
    assert (spec.fix[:916] == 3).all() and (spec.fix[916:] == 0).all()
    assert spec.ubc[0]["is_zero"] == 1 and len(spec.ubc[0]["nodes"]) == 916
    assert spec.search_factor == 10.0 and spec.neigh_update_interval == 100


def test_single_particle_deck_regions():
    spec = deck.load_deck(SINGLE)
    assert not spec.multi_particle and spec.n_particles == 1
    x = spec.x_ref
    lo = (x[:, 0] <= -0.0005 + 1e-12) & (x[:, 1] <= -0.0005 + 1e-12) & (x[:, 0] >= -0.001) & (x[:, 1] >= -0.001)
    hi = (x[:, 0] >= 0.0005 - 1e-12) & (x[:, 1] >= 0.0005 - 1e-12) & (x[:, 0] <= 0.001) & (x[:, 1] <= 0.001)
    z, lin = spec.ubc
    np.testing.assert_array_equal(np.sort(z["nodes"]), np.where(lo)[0])
    np.testing.assert_array_equal(np.sort(lin["nodes"]), np.where(hi)[0])
    assert z["is_zero"] == 1 and lin["time_fn"] == capi.UFN_LINEAR and lin["time_params"][0] == 0.005
    assert (spec.fix[lo | hi] == 3).all() and (spec.fix[~(lo | hi)] == 0).all()


@pytest.mark.parametrize("path", [DEMO, SINGLE])
def test_decks_run_on_the_oracle(path):
    spec = deck.load_deck(path)
    o = oracle.OracleEngine(spec, 4)
    o.integrate_begin()
    o.integrate_cd(20)
    u = o.field(capi.FIELD_U)
    assert np.isfinite(u).all() and np.abs(u).max() > 0


def test_bond_based_material_types(tmp_path):
    """The demo deck with its zones switched to PDElasticBond and PMBBond:
    material ids and constants (PdElastic c = 24 lambda / (pi delta^3); 2D
    PMB c = 24 E / (pi delta^3 (1 - nu)), s0 = sqrt(5 pi Gc / (9 E delta)),
    nu = 1/4, E from K: material.h:925-1000, 1260-1306), and an oracle run."""
    cfg = yaml.safe_load(open(DEMO))
    d = os.path.dirname(DEMO)
    for z in ("Zone_1", "Zone_2"):
        cfg["Mesh"][z]["File"] = os.path.join(d, os.path.basename(cfg["Mesh"][z]["File"].strip()))
    cfg["Particle_Generation"]["From_File"] = os.path.join(d, "particle_locations_0.csv")
    cfg["Material"]["Zone_1"]["Type"] = "PDElasticBond"
    cfg["Material"]["Zone_2"]["Type"] = "PMBBond"
    path = tmp_path / "bond.yaml"
    path.write_text(yaml.safe_dump(cfg))
    spec = deck.load_deck(str(path))
    p0, p1 = spec.particles
    assert p0["material_type"] == capi.MAT_PDELASTIC and p1["material_type"] == capi.MAT_PMB
    delta, nu = 6e-4, 0.25
    for p, m in ((p0, cfg["Material"]["Zone_1"]), (p1, cfg["Material"]["Zone_2"])):
        E = 3.0 * m["K"] * (1.0 - 2.0 * nu)
        lam = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu))
        if p is p0:
            assert p["mat"][0] == pytest.approx(24.0 * lam / (math.pi * delta ** 3), rel=1e-14)
        else:
            assert p["mat"][0] == pytest.approx(24.0 * E / (math.pi * delta ** 3 * (1.0 - nu)), rel=1e-14)
            assert p["mat"][1] == pytest.approx(math.sqrt(5.0 * math.pi * m["Gc"] / (9.0 * E * delta)), rel=1e-14)
    o = oracle.OracleEngine(spec, 4)
    o.integrate_begin()
    o.integrate_cd(20)
    u = o.field(capi.FIELD_U)
    assert np.isfinite(u).all() and np.abs(u).max() > 0
