"""Reference decks run unmodified on the GPU: FP64 parity with the oracle
(1e-12 of each field's maximum; bitwise in practice) on the two-particle demo
deck and the single-particle deck with region boundary conditions."""
import os

import pytest

from peridem_amd import capi, deck
import oracle

from parity_util import compare_state

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
DECKS = [os.path.join(HERE, "golden", "decks", "twoparticle_demo", "input_0.yaml"),
         os.path.join(HERE, "golden", "decks", "single_particle_circle", "input_0.yaml")]


@pytest.mark.parametrize("path", DECKS)
def test_deck_parity(path):
    spec = deck.load_deck(path)
    print(f"[deck] {os.path.basename(os.path.dirname(path))} nodes={spec.n_nodes}", flush=True)
    g = capi.GpuEngine(spec, capi.PREC_FP64)
    r = oracle.OracleEngine(spec, 8)
    g.integrate_begin()
    r.integrate_begin()
    for k in range(3):
        g.integrate_cd(100)
        r.integrate_cd(100)
        out = compare_state(g, r, tag=f"deck {path} block {k}")
        print(f"[deck] block {k}: {out}", flush=True)
