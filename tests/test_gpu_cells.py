"""The contact search's cell list has no per-cell capacity: at every rebuild
the nodes are sorted by cell key (counting sort: per-cell counts, a scan, a
scatter), so any crowding of nodes into a cell is held.  Checked on a model
whose particles are moved on top of one another: one search cell then holds
about six times the nodes of the densest cell of the reference
configuration, far beyond the fixed capacity (twice that count plus 32)
that cells had before; candidate sets, forces and the state after further
steps equal the oracle's."""
import numpy as np
import pytest

from peridem_amd import capi, configs
import oracle

from parity_util import compare_state, same_sets

pytestmark = pytest.mark.gpu


def _crowd(spec, movers, target):
    """Displacements that put the particles `movers` on top of particle
    `target`, each shifted by a fraction of the mesh size so that no two nodes
    coincide (a zero node distance has no contact direction)."""
    u = np.zeros_like(spec.x_ref)
    pt = spec.particles[target]
    c0 = spec.x_ref[pt["center_node"]]
    h = pt["mesh_size"]
    for k, p in enumerate(movers, start=1):
        q = spec.particles[p]
        shift = np.array([0.37 * k * h, 0.21 * k * h, 0.0])
        u[q["glob_start"]:q["glob_end"]] = c0 + shift - spec.x_ref[q["center_node"]]
    return u


def test_crowded_cells_beyond_the_old_capacity():
    spec = configs.compressive_2d(n_particles=9, nodes_per_particle=100, seed=4, interval=1, wall_speed=0.0,
                                  R_lo=1e-3, R_hi=1e-3, pitch_factor=1.03, dt=1e-9)
    u = _crowd(spec, movers=[1, 2, 3, 4, 5], target=0)
    # candidates per node grow with the stacking too (about five particles'
    # nodes within the search radius): a capacity above the geometric default
    g = capi.GpuEngine(spec, capi.PREC_FP64, contact_capacity=1024)
    r = oracle.OracleEngine(spec, 8)
    for e in (g, r):
        e.set_field(capi.FIELD_U, u)
        e.set_field(capi.FIELD_X, spec.x_ref + u)
        e.integrate_begin()
    a, b = g.neighbors(capi.NEIGH_CONTACT), r.neighbors(capi.NEIGH_CONTACT)
    cnt = np.diff(b[0])
    print(f"[cells] nodes={spec.n_nodes}: candidates {int(cnt.sum())}, at most {int(cnt.max())} per node",
          flush=True)
    assert cnt.max() > 0
    assert same_sets(*a, *b)
    compare_state(g, r, tag="crowded, n=0")
    # the densest cell, counted on the host at the engine's cell size (the
    # largest search radius the controller can pick: search factor x largest
    # contact radius, in 2D)
    x = spec.x_ref + u
    st = g.contact_state()
    cell = max(1.0, spec.search_factor) * st["max_contact_r"]
    keys = np.floor((x[:, :2] - x[:, :2].min(0)) / cell).astype(np.int64)
    _, per_cell = np.unique(keys[:, 0] * 100000 + keys[:, 1], return_counts=True)
    keys0 = np.floor((spec.x_ref[:, :2] - spec.x_ref[:, :2].min(0)) / cell).astype(np.int64)
    _, per_cell0 = np.unique(keys0[:, 0] * 100000 + keys0[:, 1], return_counts=True)
    print(f"[cells] densest cell: {per_cell.max()} nodes (reference configuration "
          f"{per_cell0.max()})", flush=True)
    assert per_cell.max() > 2 * per_cell0.max() + 32
    for k in range(3):
        g.integrate_cd(10)
        r.integrate_cd(10)
        compare_state(g, r, tag=f"crowded, chunk {k}")
        assert same_sets(*g.neighbors(capi.NEIGH_CONTACT), *r.neighbors(capi.NEIGH_CONTACT))
