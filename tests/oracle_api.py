"""ctypes declarations of the oracle's known-answer helpers."""
import ctypes as C

import numpy as np

import oracle

D = C.c_double


def lib():
    L = oracle.library()
    L.oracle_rnp_bond_f.argtypes = [D, D, D, C.POINTER(D), C.c_int, C.c_int, C.POINTER(D)]
    L.oracle_rnp_bond_f.restype = C.c_int
    L.oracle_pdstate_bond_f.argtypes = [D] * 7
    L.oracle_pdstate_bond_f.restype = D
    L.oracle_influence.argtypes = [C.c_int, D, D, D]
    L.oracle_influence.restype = D
    L.oracle_bond_f.argtypes = [C.c_int, D, D, D, D, C.c_int, C.POINTER(D)]
    L.oracle_bond_f.restype = C.c_int
    return L


def rnp_bond_f(L, J, r, s, mat, irrev, fs_in):
    m = (D * 6)(*list(mat)[:6], *([0.0] * (6 - len(mat))))
    f = D()
    fs = L.oracle_rnp_bond_f(J, r, s, m, irrev, fs_in, C.byref(f))
    return fs, f.value


def bond_f(L, law, J, S, c, s0, fs_in):
    f = D()
    fs = L.oracle_bond_f(law, J, S, c, s0, fs_in, C.byref(f))
    return fs, f.value
