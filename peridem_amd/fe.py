"""Strain / stress post-processing of a model whose particles carry finite
element meshes (the `Strain_Stress` output tag: DEMModel::output,
demModel.cpp:2026-2072; DEMModel::setupQuadratureData, :1602-1658).

The element lists of all particles are concatenated particle after particle
(the reference's d_globQuadStart order), so quadrature point k of element e
of the concatenation is row e * points_per_element + k of the outputs.  The
device computes the points, strain and stress (pdem_strain_stress)."""
from __future__ import annotations

import numpy as np


def element_arrays(spec):
    """(global element-node ids [ne][dim+1], owning particle [ne],
    per-particle {nu, lambda, mu, is_plane_strain} [np][4])."""
    nv = spec.dim + 1
    nodes, owner = [], []
    elastic = np.zeros((len(spec.particles), 4))
    for pid, p in enumerate(spec.particles):
        nu, lam, mu = p["elastic"]
        elastic[pid] = (nu, lam, mu, float(p.get("plane_strain", 0)))
        cells = p.get("fe_cells")
        if cells is None:
            raise ValueError(f"particle {pid} has no element-node connectivity (a .msh mesh is needed)")
        if cells.shape[1] != nv:
            raise ValueError(f"particle {pid}: strain / stress needs linear triangles (2D) or tetrahedra (3D); "
                             "the reference's quadrangle path reads past its gradient table (quadElem.cpp:51)")
        nodes.append(cells + p["glob_start"])
        owner.append(np.full(len(cells), pid, np.int32))
    return np.concatenate(nodes), np.concatenate(owner), elastic


def setup(eng, spec, quad_order=None):
    """Upload the element data of `spec` to a GpuEngine."""
    order = quad_order if quad_order is not None else spec.info.get("quad_order", 1)
    eng.fe_setup(order, *element_arrays(spec))
