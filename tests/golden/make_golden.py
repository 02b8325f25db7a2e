"""Write tests/golden/ref_vectors.npz from the reference's own sources.

Run in the container that holds the reference tree, after
`make -C oracle ref` has compiled oracle/_ref/libref.so from the reference's
material laws (src/material/mparticle/material.h, influenceFn.cpp), compare
functions (src/util/function.cpp), nanoflann radius searches
(src/nsearch/nsearch.h) and linear finite elements (src/fe/triElem.cpp,
tetElem.cpp).  The file holds inputs and the reference's outputs
only; the tests compare the oracle against it.
"""
import ctypes as C
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

D = C.c_double
PD = C.POINTER(C.c_double)
PI64 = C.POINTER(C.c_int64)


def lib():
    path = os.path.join(ROOT, "oracle", "_ref", "libref.so")
    L = C.CDLL(path)
    L.ref_rnp_params.argtypes = [D, D, D, D, C.c_int, C.c_int, PD, C.c_int, PD]
    L.ref_rnp_bond_ef.argtypes = [D, D, D, D, C.c_int, C.c_int, PD, C.c_int, D, D, C.c_int,
                                  C.c_int, PD]
    L.ref_rnp_bond_ef.restype = C.c_int
    L.ref_pdstate_params.argtypes = [D, D, D, D, D, D, D, C.c_int, C.c_int, PD, C.c_int, PD]
    L.ref_pdstate_bond_f.argtypes = [D, D, D, D, D, D, C.c_int, D, D, D, D]
    L.ref_pdstate_bond_f.restype = D
    L.ref_rnp_get_s.argtypes = [PD, PD]
    L.ref_rnp_get_s.restype = D
    L.ref_pdstate_get_s.argtypes = [PD, PD]
    L.ref_pdstate_get_s.restype = D
    L.ref_influence.argtypes = [C.c_int, PD, C.c_int, C.c_int, D]
    L.ref_influence.restype = D
    L.ref_influence_moment.argtypes = [C.c_int, PD, C.c_int, C.c_int, C.c_int]
    L.ref_influence_moment.restype = D
    L.ref_is_greater.argtypes = [D, D]
    L.ref_is_less.argtypes = [D, D]
    L.ref_neighbors.argtypes = [PD, C.c_int64, PI64, PD, C.POINTER(C.c_uint8), C.c_int, PI64, PI64]
    L.ref_bond_params.argtypes = [C.c_int, D, D, D, D, C.c_int, C.c_int, PD, C.c_int, PD]
    L.ref_bond_ef.argtypes = [C.c_int, D, D, D, D, C.c_int, C.c_int, PD, C.c_int, D, D, C.c_int, PD, PD, PD]
    L.ref_bond_ef.restype = C.c_int
    L.ref_material_properties.argtypes = [C.c_int, D, D, D, D, D, D, D, C.c_int, C.c_int, PD, C.c_int, PD]
    L.ref_fe_quad_datas.argtypes = [C.c_int, C.c_int, PD, PD, PD, PD]
    L.ref_fe_quad_datas.restype = C.c_int
    return L


def p(a):
    return a.ctypes.data_as(PD)


def main():
    L = lib()
    rng = np.random.default_rng(12345)
    out = {}
    empty = np.zeros(2)
    # ---- RNP parameters and bond law
    rows_p, rows_b = [], []
    for dim in (2, 3):
        for inf in (0, 1, 2):
            for (E, Gc, hor) in ((2.0e9, 500.0, 3.0e-4), (5.0e7, 50.0, 6.0e-4)):
                o = np.zeros(8)
                L.ref_rnp_params(E, Gc, 1200.0, hor, dim, inf, p(empty), 0, p(o))
                rows_p.append([dim, inf, E, Gc, hor] + list(o))
                rbar = o[2]
                for r_frac in (0.15, 0.4, 0.8, 0.999):
                    r = r_frac * hor
                    sc = rbar / np.sqrt(r)
                    for s in (0.0, 1e-7, -3e-5, 2e-3, 0.3 * sc, -0.999 * sc, 1.00002 * sc,
                              -1.5 * sc, 4.0 * sc):
                        for fs_in in (0, 1):
                            ef = np.zeros(2)
                            fs = L.ref_rnp_bond_ef(E, Gc, 1200.0, hor, dim, inf, p(empty), 0, r, s,
                                                   fs_in, 1, p(ef))
                            rows_b.append([dim, inf, E, Gc, hor, r, s, fs_in, fs, ef[0], ef[1]])
    out["rnp_params"] = np.array(rows_p)
    out["rnp_bond"] = np.array(rows_b)
    # ---- PDState parameters and bond law
    rows_p, rows_b = [], []
    for dim in (2, 3):
        for (E, K, G, nu) in ((-1.0, 2.16e7, 1.296e7, -1.0), (2.0e9, -1.0, -1.0, 0.25),
                              (-1.0, 1.0e9, -1.0, 0.3)):
            for hor in (6.0e-4, 4.8e-4):
                o = np.zeros(6)
                L.ref_pdstate_params(E, K, G, nu, 50.0, 1200.0, hor, dim, 1, p(empty), 0, p(o))
                rows_p.append([dim, E, K, G, nu, hor] + list(o))
                for _ in range(12):
                    r = rng.uniform(0.1, 1.0) * hor
                    s = rng.normal() * 1e-3
                    mx = rng.uniform(0.5, 2.0) * (hor ** (dim + 2))
                    th = rng.normal() * 1e-3
                    f = L.ref_pdstate_bond_f(E, K, G, nu, 50.0, hor, dim, r, s, mx, th)
                    rows_b.append([dim, E, K, G, nu, hor, r, s, mx, th, o[0], o[1], f])
    out["pdstate_params"] = np.array(rows_p)
    out["pdstate_bond"] = np.array(rows_b)
    # ---- PMB (law 2) and PdElastic (law 3) parameters and bond laws
    rows_p, rows_b = [], []
    for law in (2, 3):
        for dim in (2, 3):
            for inf in (0, 1):
                for (E, Gc, hor) in ((2.0e9, 500.0, 3.0e-4), (5.0e7, 50.0, 6.0e-4)):
                    o = np.zeros(4)
                    L.ref_bond_params(law, E, Gc, 1200.0, hor, dim, inf, p(empty), 0, p(o))
                    rows_p.append([law, dim, inf, E, Gc, hor] + list(o))
                    s0 = o[1]
                    for r_frac in (0.2, 0.7, 0.999):
                        r = r_frac * hor
                        for s in (0.0, 1e-6, -2e-4, 0.5 * s0, s0 + 0.5e-10, s0 * 1.00002, -1.3 * s0, 3e-2):
                            for fs_in in (0, 1):
                                dx = rng.normal(size=3) * hor * 0.5
                                du = rng.normal(size=3) * hor * 1e-3
                                ef = np.zeros(4)
                                fs = L.ref_bond_ef(law, E, Gc, 1200.0, hor, dim, inf, p(empty), 0, r, s, fs_in,
                                                   p(dx), p(du), p(ef))
                                rows_b.append([law, dim, inf, E, Gc, hor, r, s, fs_in, fs, ef[0], ef[1]]
                                              + list(dx) + list(du) + [ef[2], ef[3]])
    out["bond_params"] = np.array(rows_p)
    out["bond_law"] = np.array(rows_b)
    # ---- strains
    rows = []
    for _ in range(40):
        dx = rng.normal(size=3) * 1e-4
        du = rng.normal(size=3) * 1e-7 * rng.choice([1.0, 1e3])
        rows.append(list(dx) + list(du) + [L.ref_rnp_get_s(p(dx), p(du)), L.ref_pdstate_get_s(p(dx), p(du))])
    out["strain"] = np.array(rows)
    # ---- influence functions
    rows = []
    for dim in (1, 2, 3):
        for inf, prm in ((0, []), (1, []), (2, []), (0, [2.5]), (1, [7.0]), (1, [7.0, -3.0]),
                         (2, [1.5, 0.4])):
            a = np.array(prm + [0.0, 0.0])
            for r in (0.0, 0.25, 0.5, 0.999):
                rows.append([dim, inf, len(prm), a[0], a[1], r,
                             L.ref_influence(inf, p(a), len(prm), dim, r)]
                            + [L.ref_influence_moment(inf, p(a), len(prm), dim, i) for i in range(4)])
    out["influence"] = np.array(rows)
    # ---- relative compares
    rows = []
    base = [0.0, 1.0, -1.0, 1e-18, 3.0e-4, 1e10]
    for a in base:
        for rel in (0.0, 1e-6, 9.99e-6, 1e-5, 1.0001e-5, 2e-5, -1e-5, -1.0001e-5, 1e-3):
            b = a * (1 + rel) + (rel if a == 0 else 0.0)
            rows.append([a, b, L.ref_is_greater(a, b), L.ref_is_less(a, b),
                         L.ref_is_greater(b, a), L.ref_is_less(b, a)])
    out["compare"] = np.array(rows, dtype=np.float64)
    # ---- neighbor searches: jittered 2D and 3D lattices, three tags, some
    # points placed exactly at the search radius
    for dim in (2, 3):
        n1 = 9 if dim == 2 else 5
        h = 1.0e-4
        g = np.stack(np.meshgrid(*([np.arange(n1) * h] * dim), indexing="ij"), -1).reshape(-1, dim)
        pts = np.zeros((len(g), 3))
        pts[:, :dim] = g + rng.normal(size=g.shape) * 0.1 * h
        pts[: len(g) // 3, :dim] = g[: len(g) // 3]  # exact lattice part: ties at r
        tags = (np.arange(len(pts)) * 3 // len(pts)).astype(np.int64)
        r = 2.0 * h
        rad = np.full(len(pts), r)
        rp = np.zeros(len(pts) + 1, np.int64)
        L.ref_neighbors(p(pts), len(pts), tags.ctypes.data_as(PI64), p(rad), None, 1,
                        rp.ctypes.data_as(PI64), None)
        cols = np.zeros(max(1, rp[-1]), np.int64)
        L.ref_neighbors(p(pts), len(pts), tags.ctypes.data_as(PI64), p(rad), None, 1,
                        rp.ctypes.data_as(PI64), cols.ctypes.data_as(PI64))
        out[f"nbr{dim}_pts"], out[f"nbr{dim}_tags"], out[f"nbr{dim}_r"] = pts, tags, np.array([r])
        out[f"nbr{dim}_inc_rp"], out[f"nbr{dim}_inc_cols"] = rp, cols[: rp[-1]]
        rp2 = np.zeros(len(pts) + 1, np.int64)
        L.ref_neighbors(p(pts), len(pts), tags.ctypes.data_as(PI64), p(rad), None, 0,
                        rp2.ctypes.data_as(PI64), None)
        cols2 = np.zeros(max(1, rp2[-1]), np.int64)
        L.ref_neighbors(p(pts), len(pts), tags.ctypes.data_as(PI64), p(rad), None, 0,
                        rp2.ctypes.data_as(PI64), cols2.ctypes.data_as(PI64))
        out[f"nbr{dim}_exc_rp"], out[f"nbr{dim}_exc_cols"] = rp2, cols2[: rp2[-1]]
    # computeMaterialProperties(dim): law, E, K, G, nu, Gc, horizon, dim, inf -> nu, lambda, mu
    rows = []
    inf_p = np.zeros(2)
    for law in range(4):
        for dim in (2, 3):
            for inf in (0, 1, 2):
                for _ in range(3):
                    E = 10 ** rng.uniform(6, 11)
                    if law == 1:
                        K, G, Ein = E / 1.2, E / 2.6, -1.0
                    else:
                        K, G, Ein = -1.0, -1.0, E
                    Gc = 10 ** rng.uniform(0, 3)
                    hz = 10 ** rng.uniform(-4, -2)
                    o = np.zeros(3)
                    L.ref_material_properties(law, Ein, K, G, -1.0, Gc, 1000.0, hz, dim, inf, p(inf_p), 0, p(o))
                    rows.append([law, Ein, K, G, Gc, hz, dim, inf, *o])
    out["mat_props"] = np.array(rows)
    # linear element quadrature data: vertices -> mapped points, shapes, gradients
    for dim, orders in ((2, range(1, 6)), (3, range(1, 4))):
        nv = dim + 1
        verts = []
        for _ in range(6):
            base = rng.uniform(-1e-3, 1e-3, 3)
            V = base + rng.uniform(-2e-4, 2e-4, (nv, 3))
            if dim == 2:
                V[:, 2] = 0.0
            verts.append(V)
        out[f"fe{dim}_nodes"] = np.array(verts)
        for order in orders:
            P, S, G = [], [], []
            for V in verts:
                pts, sh, de = np.zeros(3 * 8), np.zeros(8 * nv), np.zeros(8 * nv * dim)
                nq = L.ref_fe_quad_datas(dim, order, p(np.ascontiguousarray(V)), p(pts), p(sh), p(de))
                P.append(pts[:3 * nq].reshape(nq, 3))
                S.append(sh[:nq * nv].reshape(nq, nv))
                G.append(de[:nq * nv * dim].reshape(nq, nv, dim))
            out[f"fe{dim}_o{order}_pts"] = np.array(P)
            out[f"fe{dim}_o{order}_shapes"] = np.array(S)
            out[f"fe{dim}_o{order}_ders"] = np.array(G)
    np.savez_compressed(os.path.join(HERE, "ref_vectors.npz"), **out)
    print({k: v.shape for k, v in out.items()})


if __name__ == "__main__":
    main()
