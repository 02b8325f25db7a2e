"""Seeded particle meshes and placements for the synthetic workloads.

The reference meshes particles with gmsh (absent here); these generators give
node clouds of the same kind: each particle is a transformed copy of one
reference mesh per zone (ParticleTransform::apply, particleTransform.h:96-100:
x = t + s R(theta) p), with nodal volume V_ref s^d (baseParticle.cpp:146-152).
"""
from __future__ import annotations

import math

import numpy as np

SQ3 = math.sqrt(3.0)


def tri_lattice(extent: float, h: float):
    """Triangular lattice points (spacing h) covering [-extent, extent]^2, row-major."""
    ny = int(math.ceil(extent / (h * SQ3 / 2))) + 1
    nx = int(math.ceil(extent / h)) + 2
    pts = []
    for j in range(-ny, ny + 1):
        y = j * h * SQ3 / 2
        off = 0.5 * h if (j % 2) else 0.0
        for i in range(-nx, nx + 1):
            pts.append((i * h + off, y))
    return np.array(pts, np.float64)


def disk_mesh(R: float, h: float):
    """Circle of radius R on a triangular lattice; area per node (sqrt 3 / 2) h^2."""
    p = tri_lattice(R, h)
    keep = np.sqrt(p[:, 0] ** 2 + p[:, 1] ** 2) <= R * (1 + 1e-12)
    p = p[keep]
    x = np.zeros((len(p), 3))
    x[:, :2] = p
    vol = np.full(len(p), SQ3 / 2 * h * h)
    return x, vol


def hexagon_mesh(R: float, h: float):
    """Regular hexagon with circumradius R and a vertex on +x (reference Hexagon,
    axis (1,0,0)); triangular lattice nodes inside."""
    p = tri_lattice(R, h)
    ax, ay = np.abs(p[:, 0]), np.abs(p[:, 1])
    tol = 1 + 1e-12
    keep = (ay <= SQ3 / 2 * R * tol) & (SQ3 * ax + ay <= SQ3 * R * tol)
    p = p[keep]
    x = np.zeros((len(p), 3))
    x[:, :2] = p
    vol = np.full(len(p), SQ3 / 2 * h * h)
    return x, vol


def ball_mesh(R: float, h: float):
    """Sphere of radius R on a cubic lattice (volume h^3 per node)."""
    n = int(math.ceil(R / h)) + 1
    g = np.arange(-n, n + 1) * h
    Z, Y, X = np.meshgrid(g, g, g, indexing="ij")
    p = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)
    keep = np.sqrt((p ** 2).sum(1)) <= R * (1 + 1e-12)
    p = p[keep]
    return p.copy(), np.full(len(p), h ** 3)


def box_mesh(lo, hi, h: float, dim: int):
    """Rectangle / cuboid on a square (cubic) lattice with spacing h, node-centred
    cells: nodes at lo + h/2 + k h."""
    lo = np.asarray(lo, float)
    hi = np.asarray(hi, float)
    axes = []
    for d in range(dim):
        n = max(1, int(round((hi[d] - lo[d]) / h)))
        axes.append(lo[d] + h * (0.5 + np.arange(n)))
    if dim == 2:
        Y, X = np.meshgrid(axes[1], axes[0], indexing="ij")
        p = np.zeros((X.size, 3))
        p[:, 0] = X.ravel()
        p[:, 1] = Y.ravel()
        vol = np.full(len(p), h * h)
    else:
        Z, Y, X = np.meshgrid(axes[2], axes[1], axes[0], indexing="ij")
        p = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)
        vol = np.full(len(p), h ** 3)
    return p, vol


def transform(ref: np.ndarray, center, theta: float, scale: float):
    """t + s R_z(theta) p (rotation about the z axis)."""
    c, s = math.cos(theta), math.sin(theta)
    out = np.empty_like(ref)
    out[:, 0] = center[0] + scale * (c * ref[:, 0] - s * ref[:, 1])
    out[:, 1] = center[1] + scale * (s * ref[:, 0] + c * ref[:, 1])
    out[:, 2] = center[2] + scale * ref[:, 2]
    return out


def center_node(ref: np.ndarray, center, bounding_radius: float) -> int:
    """RefParticle: the first node strictly closer (util::isLess) than the running
    best, starting from the bounding radius (refParticle.cpp:36-45)."""
    best = bounding_radius
    idx = 0
    d = np.sqrt(((np.asarray(center)[None, :] - ref) ** 2).sum(1))
    for i, v in enumerate(d):
        if (best - v) > max(abs(v), abs(best)) * 1e-5:
            best = v
            idx = i
    return idx


def mesh_size(x: np.ndarray) -> float:
    """Smallest distance between two nodes of a particle (util::computeMeshSize,
    geom.cpp:578-600), evaluated as sqrt(dx^2 + dy^2 + dz^2)."""
    from scipy.spatial import cKDTree

    if len(x) < 2:
        return 0.0
    t = cKDTree(x)
    _, nn = t.query(x, k=2)
    j = nn[:, 1]
    d = x - x[j]
    dist = np.sqrt(d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1] + d[:, 2] * d[:, 2])
    return float(dist.min())
