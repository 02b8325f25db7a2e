"""Strain / stress post-processing: the oracle (oracle/fe_oracle.py) pinned
on the CPU.

* Element data against the reference's own TriElem / TetElem classes
  (golden vectors written by tests/golden/make_golden.py from
  src/fe/triElem.cpp, tetElem.cpp): mapped quadrature points, shape values
  and physical shape gradients, bit for bit, every supported order.
* The quadrature points of each rule, with the published weights, integrate
  every monomial up to the rule's degree exactly (a wrong point coordinate
  fails this).
* computeMaterialProperties (nu, lambda, mu) of all four materials against
  the reference's classes.
* Closed forms for linear displacement fields u = A x + c: the shear strains
  are the symmetric gradient (A_ij + A_ji) / 2; the normal strains are the
  reference's sum over the vertices of u_i + dN_i/dx_i, i.e. (dim + 1) times
  u at the centroid (the gradients sum to zero); the stress is the
  isotropic stiffness in Voigt form applied to that strain, with the plane
  stress / plane strain out-of-plane rules; the current quadrature points
  are x_q + u(x_q).
"""
import math
import os

import numpy as np
import pytest

from peridem_amd.material import InfluenceFn, MaterialDeck, elastic_constants, material_params
from oracle import fe_oracle

G = np.load(os.path.join(os.path.dirname(__file__), "golden", "ref_vectors.npz"))

# published weights (reference triangle area 1/2, reference tetrahedron volume 1/6)
WEIGHTS = {
    (2, 1): [0.5],
    (2, 2): [1 / 6] * 3,
    (2, 3): [-27 / 96, 25 / 96, 25 / 96, 25 / 96],
    (2, 4): [0.5 * 0.22338158967801] * 3 + [0.5 * 0.10995174365532] * 3,
    (2, 5): [0.5 * 0.225] + [0.5 * 0.13239415278851] * 3 + [0.5 * 0.12593918054483] * 3,
    (3, 1): [1 / 6],
    (3, 2): [1 / 24] * 4,
    (3, 3): [-2 / 15] + [0.075] * 4,
}


def _monomial_integral(exps):
    # int over the unit simplex of prod x_k^a_k = prod a_k! / (n + sum a_k)!
    n = len(exps)
    return math.prod(math.factorial(a) for a in exps) / math.factorial(n + sum(exps))


@pytest.mark.parametrize("dim,order", sorted(WEIGHTS))
def test_quadrature_rules_are_exact(dim, order):
    pts = fe_oracle.quad_points(dim, order)[:, :dim]
    w = np.array(WEIGHTS[(dim, order)])
    assert len(w) == len(pts)
    for total in range(order + 1):
        for exps in np.ndindex(*([total + 1] * dim)):
            if sum(exps) != total:
                continue
            val = float((w * np.prod(pts ** np.array(exps), axis=1)).sum())
            # the order-4 and -5 triangle rules are tabulated to 14 digits
            assert abs(val - _monomial_integral(exps)) < 1e-12, (dim, order, exps)


@pytest.mark.parametrize("dim,order", sorted(WEIGHTS))
def test_element_data_matches_reference(dim, order):
    V = G[f"fe{dim}_nodes"]
    pts, shp, ders = G[f"fe{dim}_o{order}_pts"], G[f"fe{dim}_o{order}_shapes"], G[f"fe{dim}_o{order}_ders"]
    dN = fe_oracle.gradients(dim, V)
    qp = fe_oracle.quad_points(dim, order)
    for q in range(len(qp)):
        N = fe_oracle.shapes(dim, qp[q])
        np.testing.assert_array_equal(np.broadcast_to(N, shp[:, q].shape), shp[:, q])
        np.testing.assert_array_equal(fe_oracle.mapped_points(dim, V, N, qp[q]), pts[:, q])
        np.testing.assert_array_equal(dN, ders[:, q])


def test_material_properties_match_reference():
    names = {0: "RNPBond", 1: "PDState", 2: "PMBBond", 3: "PDElasticBond"}
    for row in G["mat_props"]:
        law, E, K, Gs, Gc, hz, dim, inf = int(row[0]), row[1], row[2], row[3], row[4], row[5], int(row[6]), int(row[7])
        deck = MaterialDeck(type=names[law], E=E, K=K, G=Gs, Gc=Gc, density=1000.0, influence_fn_type=inf)
        fn = InfluenceFn(inf, [], dim)
        mp = material_params(deck, dim, hz, fn)
        nu, lam, mu = elastic_constants(mp, dim, hz, fn)
        ref = row[8:11]
        assert nu == ref[0], (law, dim)
        for a, b in ((lam, ref[1]), (mu, ref[2])):
            assert abs(a - b) <= 8 * np.spacing(abs(b)), (law, dim, inf, a, b)


def _mesh(dim, rng, n=40):
    """A few random well-shaped elements sharing nothing (global ids 0..)."""
    nv = dim + 1
    X = []
    for _ in range(n):
        c = rng.uniform(-1e-3, 1e-3, 3)
        if dim == 2:
            V = c + np.array([[0, 0, 0], [1, 0.1, 0], [0.2, 1, 0]]) * 1e-4 + rng.uniform(-1e-5, 1e-5, (3, 3))
            V[:, 2] = 0.0
        else:
            V = c + np.array([[0, 0, 0], [1, 0.1, 0], [0.2, 1, 0.1], [0.1, 0.2, 1]]) * 1e-4 + \
                rng.uniform(-1e-5, 1e-5, (4, 3))
        X.append(V)
    xr = np.concatenate(X)
    cells = np.arange(n * nv).reshape(n, nv)
    return xr, cells


@pytest.mark.parametrize("dim,plane_strain", [(2, 0), (2, 1), (3, 0)])
def test_linear_field_closed_forms(dim, plane_strain):
    rng = np.random.default_rng(7 + dim)
    xr, cells = _mesh(dim, rng)
    A = rng.uniform(-1e-3, 1e-3, (3, 3))
    c = rng.uniform(-1e-6, 1e-6, 3)
    if dim == 2:
        A[2, :] = 0.0
        A[:, 2] = 0.0
        c[2] = 0.0
    u = xr @ A.T + c
    nu, lam, mu = 0.3, 7.0e9, 3.0e9
    elastic = np.array([[nu, lam, mu, plane_strain]])
    part = np.zeros(len(cells), np.int64)
    order = 2
    xq, eps, sig = fe_oracle.strain_stress(dim, order, xr, u, cells, part, elastic)
    nq = len(fe_oracle.quad_points(dim, order))
    # current points: the mapped reference point displaced by the linear field
    Xe = xr[cells]
    for q, qp in enumerate(fe_oracle.quad_points(dim, order)):
        N = fe_oracle.shapes(dim, qp)
        xp = np.einsum("k,ekd->ed", N, Xe)
        np.testing.assert_allclose(xq[q::nq], xp + xp @ A.T + c, rtol=0, atol=1e-15)
    # strains (float32 storage)
    sym = 0.5 * (A + A.T)
    cent = Xe.mean(1)
    ucent = cent @ A.T + c
    e = eps.astype(np.float64).reshape(len(cells), nq, 6)
    tol = 2e-6
    for k, (i, j) in ((5, (0, 1)), (4, (0, 2)), (3, (1, 2))):
        if dim == 2 and k != 5:
            assert np.all(e[:, :, k] == 0.0)
            continue
        np.testing.assert_allclose(e[:, :, k], sym[i, j], rtol=tol, atol=tol * np.abs(sym).max())
    dN = fe_oracle.gradients(dim, Xe)
    Ue = u[cells]
    for k in range(dim):
        expect = (dim + 1) * ucent[:, k]
        # the float32 accumulator holds partial sums of size |dN| ~ 1 / h:
        # each of the dim + 1 additions rounds at that scale
        part = np.cumsum(Ue[:, :, k] + dN[:, :, k], axis=1)
        bound = (dim + 2) * np.spacing(np.abs(part).max(1).astype(np.float32)).astype(np.float64)
        err = np.abs(e[:, :, k] - expect[:, None]).max(1)
        assert np.all(err <= bound), (k, (err / bound).max())
    # stress: Voigt stiffness applied to the stored strain (engineering shear 2 eps)
    Cv = np.zeros((6, 6))
    Cv[:3, :3] = lam
    Cv[np.arange(3), np.arange(3)] += 2 * mu
    Cv[3:, 3:] = np.eye(3) * mu
    s = sig.astype(np.float64).reshape(len(cells), nq, 6)
    ev = e.copy()
    if dim == 2 and plane_strain:
        ev[:, :, 2] = -nu * (ev[:, :, 0] + ev[:, :, 1]) / (1 - nu)
        np.testing.assert_allclose(e[:, :, 2], ev[:, :, 2], rtol=1e-6, atol=1e-12)
    voigt = np.concatenate([ev[:, :, :3], 2 * ev[:, :, 3:]], axis=2)
    expect = voigt @ Cv.T
    if dim == 2 and not plane_strain:
        expect[:, :, 2] = nu * (expect[:, :, 0] + expect[:, :, 1])
    scale = np.abs(expect).max()
    np.testing.assert_allclose(s, expect, rtol=0, atol=1e-6 * scale)


def test_rigid_translation_normal_strain_is_vertex_count_times_shift():
    """u = c everywhere: the reference's normal strain is (dim + 1) c, shear 0."""
    rng = np.random.default_rng(3)
    xr, cells = _mesh(2, rng, 10)
    xr = xr * 1e4  # unit-size elements: the float32 sums of dN round at ~1e-7
    u = np.zeros_like(xr)
    u[:, 0], u[:, 1] = 0.2, -0.1
    _, eps, _ = fe_oracle.strain_stress(2, 1, xr, u, cells, np.zeros(10, np.int64), np.array([[0.25, 1e9, 1e9, 0]]))
    np.testing.assert_allclose(eps[:, 0], 0.6, atol=1e-6)
    np.testing.assert_allclose(eps[:, 1], -0.3, atol=1e-6)
    assert np.all(np.abs(eps[:, 5]) < 1e-7)
