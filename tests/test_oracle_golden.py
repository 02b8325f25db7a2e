"""The oracle and the host-side setup against vectors written by the
reference's own sources (tests/golden/make_golden.py): material parameters,
bond laws, strains, influence functions, relative compares and nanoflann
neighbor sets.  Bit-exact where the arithmetic is the same sequence of IEEE
operations; the exp()-based RNP force to 1 ulp."""
import math
import os

import numpy as np
import pytest

from peridem_amd import capi
from peridem_amd.material import (InfluenceFn, MaterialDeck, pdelastic_params, pdstate_params, pmb_params,
                                  rnp_params)
from peridem_amd.model import ContactPairDeck, ModelBuilder, ZoneDeck
import oracle

from oracle_api import bond_f, lib, rnp_bond_f

G = np.load(os.path.join(os.path.dirname(__file__), "golden", "ref_vectors.npz"))


def ulp_close(a, b, ulps=2):
    if a == b:
        return True
    return abs(a - b) <= ulps * np.spacing(max(abs(a), abs(b)))


def test_rnp_parameters_match_reference():
    for row in G["rnp_params"]:
        dim, inf, E, Gc, hor = int(row[0]), int(row[1]), row[2], row[3], row[4]
        C_, beta, rbar, invF, fsc, K, J5, M = row[5:]
        deck = MaterialDeck(type="RNPBond", E=E, Gc=Gc, density=1200.0, influence_fn_type=inf)
        f = InfluenceFn(inf, [], dim)
        mp = rnp_params(deck, dim, hor, f)
        # beta and C are recovered from the reference's public bond law: to a few ulps
        assert ulp_close(mp.mat[2], rbar, 0), (dim, inf)
        assert ulp_close(mp.mat[1], beta, 4)
        assert ulp_close(mp.mat[0], C_, 16)
        assert ulp_close(mp.mat[3], invF, 0)
        assert mp.mat[4] == fsc
        assert ulp_close(mp.K_for_contact, K, 4)
        assert ulp_close(f(0.5), J5, 0)
        assert ulp_close(f.moment(dim), M, 0)


def test_rnp_bond_law_matches_reference():
    L = lib()
    for row in G["rnp_params"]:
        dim, inf, E, Gc, hor = int(row[0]), int(row[1]), row[2], row[3], row[4]
        f = InfluenceFn(inf, [], dim)
        mp = rnp_params(MaterialDeck(type="RNPBond", E=E, Gc=Gc, density=1200.0,
                                     influence_fn_type=inf), dim, hor, f)
        sel = G["rnp_bond"]
        sel = sel[(sel[:, 0] == dim) & (sel[:, 1] == inf) & (sel[:, 2] == E) & (sel[:, 4] == hor)]
        assert len(sel) > 0
        for r in sel:
            rr, s, fs_in, fs_ref, f_ref = r[5], r[6], int(r[7]), int(r[8]), r[10]
            J = f(rr / hor)
            fs, fo = rnp_bond_f(L, J, rr, s, mp.mat, 1, fs_in)
            assert fs == fs_ref, (rr, s, fs_in)
            assert ulp_close(fo, f_ref, 64), (rr, s, fo, f_ref)


def test_pdstate_parameters_and_law_match_reference():
    L = lib()
    for row in G["pdstate_params"]:
        dim, E, K, G_, nu, hor = int(row[0]), row[1], row[2], row[3], row[4], row[5]
        Kr, Gr, s0r = row[6], row[7], row[8]
        deck = MaterialDeck(type="PDState", E=E, K=K, G=G_, nu=nu, Gc=50.0, density=1200.0,
                            influence_fn_type=1)
        mp = pdstate_params(deck, dim, hor)
        assert ulp_close(mp.mat[0], Kr, 0)
        assert ulp_close(mp.mat[1], Gr, 0)
        assert ulp_close(mp.mat[2], s0r, 0)
    for r in G["pdstate_bond"]:
        dim, hor, rr, s, mx, th, K, G_, f_ref = int(r[0]), r[5], r[6], r[7], r[8], r[9], r[10], r[11], r[12]
        # the reference law evaluates J with the constant default (d + 1) here
        J = float(dim + 1)
        fo = L.oracle_pdstate_bond_f(J, rr, s, K, G_, mx, th)
        assert fo == f_ref, (fo, f_ref)


def test_pmb_and_pdelastic_match_reference():
    """PmbMaterial / PdElastic: parameters from E and Gc, the bond law with
    fracture (PMB) or without (PdElastic), the actual stretch and Sc."""
    L = lib()
    for row in G["bond_params"]:
        law, dim, inf, E, Gc, hor = int(row[0]), int(row[1]), int(row[2]), row[3], row[4], row[5]
        c_ref, s0_ref, K_ref, J5 = row[6:10]
        deck = MaterialDeck(type="PMBBond" if law == 2 else "PDElasticBond", E=E, Gc=Gc, density=1200.0,
                            influence_fn_type=inf)
        mp = (pmb_params if law == 2 else pdelastic_params)(deck, dim, hor)
        f = InfluenceFn(inf, [], dim)
        assert mp.material_type == law
        assert ulp_close(mp.mat[0], c_ref, 4), (law, dim, inf)
        if law == 2:
            assert ulp_close(mp.mat[1], s0_ref, 0)
        assert ulp_close(mp.K_for_contact, K_ref, 4)
        assert ulp_close(f(0.5), J5, 0)
        sel = G["bond_law"]
        sel = sel[(sel[:, 0] == law) & (sel[:, 1] == dim) & (sel[:, 2] == inf) & (sel[:, 3] == E)]
        assert len(sel) > 0
        for r in sel:
            rr, s_, fs_in, fs_ref, f_ref = r[6], r[7], int(r[8]), int(r[9]), r[11]
            fs, fo = bond_f(L, law, f(rr / hor), s_, mp.mat[0], mp.mat[1], fs_in)
            assert fs == fs_ref, (law, rr, s_, fs_in)
            assert ulp_close(fo, f_ref, 2), (law, rr, s_, fo, f_ref)
            dx, du = r[12:15], r[15:18]
            dy = dx + du
            ly = math.sqrt(dy[0] * dy[0] + dy[1] * dy[1] + dy[2] * dy[2])
            lx = math.sqrt(dx[0] * dx[0] + dx[1] * dx[1] + dx[2] * dx[2])
            assert (ly - lx) / lx == r[18]
            assert r[19] == (mp.mat[1] if law == 2 else np.finfo(np.float64).max)


def test_strain_measures_match_reference():
    # the oracle's strains are written out in its force loops; evaluate the same
    # expressions on the reference's inputs
    for r in G["strain"]:
        dx, du = r[0:3], r[3:6]
        s_rnp = (dx[0] * du[0] + dx[1] * du[1] + dx[2] * du[2]) / (dx[0] * dx[0] + dx[1] * dx[1] + dx[2] * dx[2])
        y = dx + du
        ly = math.sqrt(y[0] * y[0] + y[1] * y[1] + y[2] * y[2])
        lx = math.sqrt(dx[0] * dx[0] + dx[1] * dx[1] + dx[2] * dx[2])
        assert s_rnp == r[6]
        assert (ly - lx) / lx == r[7]


def test_influence_functions_match_reference():
    L = lib()
    for r in G["influence"]:
        dim, inf, npar, a0, a1, x, J = int(r[0]), int(r[1]), int(r[2]), r[3], r[4], r[5], r[6]
        f = InfluenceFn(inf, [a0, a1][:npar], dim)
        assert ulp_close(f(x), J, 0)
        assert ulp_close(L.oracle_influence(inf, f.p[0], f.p[1], x), J, 0)
        for i in range(4):
            assert ulp_close(f.moment(i), r[7 + i], 2)


def test_relative_compares_match_reference():
    L = oracle.library()
    for a, b, g1, l1, g2, l2 in G["compare"]:
        assert L.oracle_is_greater(a, b) == g1
        assert L.oracle_is_less(a, b) == l1
        assert L.oracle_is_greater(b, a) == g2
        assert L.oracle_is_less(b, a) == l2


def _cloud_spec(pts, tags, r, dim):
    mat = MaterialDeck(type="RNPBond", E=2e9, K=2e9 / 1.5, Gc=500.0, density=1200.0, horizon=r)
    zones = [ZoneDeck(mat)]
    cp = ContactPairDeck(contact_radius_factor=None, contact_radius=r, Kn=1.0)
    b = ModelBuilder(dim, 1e-9, zones, {(0, 0): cp}, search_factor=1.0, neigh_update_interval=1)
    for t in range(int(tags.max()) + 1):
        x = pts[tags == t]
        b.add_particle(0, x, np.full(len(x), 1e-6), 0, 1e-3)
    return b.build(h_override=[r / 3.0] * (int(tags.max()) + 1))


@pytest.mark.parametrize("dim", [2, 3])
def test_neighbor_sets_match_nanoflann(dim):
    pts, tags, r = G[f"nbr{dim}_pts"], G[f"nbr{dim}_tags"], float(G[f"nbr{dim}_r"][0])
    spec = _cloud_spec(pts, tags, r, dim)
    o = oracle.OracleEngine(spec, 2)
    for kind, key in ((capi.NEIGH_PD, "inc"), (capi.NEIGH_CONTACT, "exc")):
        rp, cols = o.neighbors(kind)
        rr, rc = G[f"nbr{dim}_{key}_rp"], G[f"nbr{dim}_{key}_cols"]
        assert np.array_equal(rp, rr), key
        for i in range(len(rp) - 1):
            assert np.array_equal(np.sort(cols[rp[i]:rp[i + 1]]), np.sort(rc[rr[i]:rr[i + 1]])), (key, i)
    # the lattice part puts pairs exactly at the radius: strict < excludes them
    assert rp[-1] > 0
