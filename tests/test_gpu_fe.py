"""Strain / stress post-processing on the GPU (pdem_strain_stress) against
the oracle (oracle/fe_oracle.py) on the engine's own displacement field.

Bar: the current quadrature points, strain and stress arrays equal to the
oracle's bit for bit (the kernel follows the reference's FP64 operation
order and rounds its FP32 accumulators at the same places), for linear
triangles (the shipped single-particle and two-particle decks, quadrature
orders 1, 3 and 5, plane stress and plane strain) and linear tetrahedra
(a cube of Kuhn tetrahedra, orders 1-3)."""
import itertools
import os

import numpy as np
import pytest

from peridem_amd import capi, configs, deck, fe
from peridem_amd.deck import _Region, fe_nodal_volumes
from peridem_amd.geometry import center_node
from peridem_amd.model import BcDeck, ModelBuilder, ZoneDeck
from oracle import fe_oracle

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
DEMO = os.path.join(HERE, "golden", "decks", "twoparticle_demo", "input_0.yaml")
SINGLE = os.path.join(HERE, "golden", "decks", "single_particle_circle", "input_0.yaml")


def _cube_tets(n=6, h=2.5e-4):
    g = np.arange(n + 1) * h - 0.5 * n * h
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    x = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1)
    idx = lambda i, j, k: (i * (n + 1) + j) * (n + 1) + k
    cells = []
    for i, j, k in itertools.product(range(n), repeat=3):
        for perm in itertools.permutations(range(3)):
            v = [0, 0, 0]
            path = [idx(i, j, k)]
            for ax in perm:
                v[ax] += 1
                path.append(idx(i + v[0], j + v[1], k + v[2]))
            cells.append(path)
    cells = np.array(cells, np.int64)
    return x, fe_nodal_volumes(x, cells, 3), cells


def _cube_spec():
    x, vol, cells = _cube_tets()
    R = np.abs(x).max()
    b = ModelBuilder(3, 2e-8, [ZoneDeck(configs.pdstate_material(ratio=2.0))], {}, multi_particle=False)
    b.add_particle(0, x, vol, center_node(x, (0, 0, 0), R), R, cells=cells)
    b.ubc.append(BcDeck([], [1, 3], capi.UFN_LINEAR, [-5.0],
                        region=_Region("cuboid", [-R, -R, -R, -0.6 * R, R, R], 3)))
    return b.build(name="kuhn_cube")


def _check(spec, g, order):
    fe.setup(g, spec, order)
    xq, eps, sig = g.strain_stress()
    cells, owner, elastic = fe.element_arrays(spec)
    u = g.field(capi.FIELD_U)
    xr = spec.x_ref.copy()
    if spec.dim == 2:
        xr[:, 2] = 0.0
        u[:, 2] = 0.0
    rx, re, rs = fe_oracle.strain_stress(spec.dim, order, xr, u, cells, owner, elastic)
    print(f"[fe] {spec.name} order {order}: {len(xq)} points, max |strain| {np.abs(eps).max():.3e}, "
          f"max |stress| {np.abs(sig).max():.3e}", flush=True)
    np.testing.assert_array_equal(xq, rx)
    np.testing.assert_array_equal(eps, re)
    np.testing.assert_array_equal(sig, rs)
    assert np.abs(u).max() > 0


@pytest.mark.parametrize("path,orders,plane", [(SINGLE, (1, 3), 0), (SINGLE, (2,), 1), (DEMO, (5,), 0)])
def test_strain_stress_triangles(path, orders, plane):
    spec = deck.load_deck(path)
    for p in spec.particles:
        p["plane_strain"] = plane
    print(f"[fe] {spec.name} nodes={spec.n_nodes}", flush=True)
    g = capi.GpuEngine(spec, capi.PREC_FP64)
    g.integrate_begin()
    g.integrate_cd(200)
    for order in orders:
        _check(spec, g, order)


def test_strain_stress_tetrahedra():
    spec = _cube_spec()
    print(f"[fe] cube nodes={spec.n_nodes}", flush=True)
    g = capi.GpuEngine(spec, capi.PREC_FP64)
    g.integrate_begin()
    g.integrate_cd(100)
    for order in (1, 2, 3):
        _check(spec, g, order)
