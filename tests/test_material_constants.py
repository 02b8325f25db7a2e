"""pdem_material_constants (host code of the engine library, for C++ hosts)
against the Python set-up (peridem_amd/material.py), which
test_oracle_golden.py pins on the reference's own material classes: equal bit
for bit for every law, dimension, influence function and parameter source.
Host code only: runs without a GPU."""
import math

import pytest

from peridem_amd import capi
from peridem_amd.material import InfluenceFn, MaterialDeck, material_params

INFS = [(capi.INF_CONST, ()), (capi.INF_LINEAR, ()), (capi.INF_GAUSSIAN, ()),
        (capi.INF_CONST, (2.5,)), (capi.INF_LINEAR, (3.0, -2.0)), (capi.INF_GAUSSIAN, (1.3, 0.25))]
TYPES = {"RNPBond": capi.MAT_RNP, "PDState": capi.MAT_PDSTATE, "PMBBond": capi.MAT_PMB,
         "PDElasticBond": capi.MAT_PDELASTIC}


def _decks():
    yield MaterialDeck(type="RNPBond", E=2.0e9, Gc=500.0, check_sc_factor=1.0)
    yield MaterialDeck(type="RNPBond", K=1.3e9, KIc=2.0e6, check_sc_factor=10.0)
    yield MaterialDeck(type="RNPBond", compute_from_classical=False, bond_potential_params=[7.5e3, 4.1e6])
    yield MaterialDeck(type="PDState", K=2.16e7, G=1.296e7, Gc=50.0)
    yield MaterialDeck(type="PDState", E=3.0e7, nu=0.3, KIc=1.0e4)
    yield MaterialDeck(type="PDState", E=3.0e7, G=1.1e7, Gc=20.0)
    yield MaterialDeck(type="PDState", compute_from_classical=False, bond_potential_params=[2e7, 1e7, 1e-3])
    yield MaterialDeck(type="PMBBond", E=2.0e9, Gc=500.0)
    yield MaterialDeck(type="PMBBond", K=1.5e9, KIc=3.0e6)
    yield MaterialDeck(type="PDElasticBond", E=2.0e9)
    yield MaterialDeck(type="PDElasticBond", K=1.0e9)


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("inf", INFS)
def test_material_constants_match_python_setup(dim, inf):
    kind, params = inf
    for md in _decks():
        md.influence_fn_type, md.influence_fn_params = kind, list(params)
        fn = InfluenceFn(kind, params, dim)
        for horizon in (3.0e-4, 7.3e-4):
            mp = material_params(md, dim, horizon, fn)
            mat, kc = capi.material_constants(
                TYPES[md.type], dim, horizon, E=md.E, K=md.K, G=md.G, nu=md.nu, Gc=md.Gc, KIc=md.KIc,
                influence_fn_type=kind, influence_fn_params=params, check_sc_factor=md.check_sc_factor,
                compute_from_classical=md.compute_from_classical,
                bond_potential_params=md.bond_potential_params)
            assert mat == [float(v) for v in mp.mat], (md.type, dim, kind, params, mat, mp.mat)
            assert kc == mp.K_for_contact, (md.type, dim, kc, mp.K_for_contact)
            assert all(math.isfinite(v) for v in mat)


def test_material_constants_errors():
    with pytest.raises(ValueError, match="E or K"):
        capi.material_constants(capi.MAT_RNP, 2, 3e-4, Gc=500.0)
    with pytest.raises(ValueError, match="exactly two"):
        capi.material_constants(capi.MAT_PDSTATE, 3, 3e-4, K=1e7, G=1e7, E=1e7, Gc=1.0)
    with pytest.raises(ValueError, match="horizon"):
        capi.material_constants(capi.MAT_PMB, 2, 0.0, E=1e9, Gc=1.0)
    with pytest.raises(ValueError, match="not supported"):
        capi.material_constants(9, 2, 3e-4, E=1e9)
