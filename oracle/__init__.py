"""CPU FP64 oracle of PeriDEM's explicit step -- test infrastructure only.

Only the tests, the smoke check and the benchmark's CPU-baseline leg import
this package.  It loads oracle/liboracle.so (built by oracle/Makefile) and,
for pinning the oracle itself, oracle/_ref/libref.so (the reference's own
sources, built only where the reference tree exists).
"""
import ctypes as _C
import os as _os
import subprocess as _sp

from peridem_amd import capi as _capi

HERE = _os.path.dirname(_os.path.abspath(__file__))
LIB = _os.path.join(HERE, "liboracle.so")
REF_LIB = _os.path.join(HERE, "_ref", "libref.so")
_lib = None


def build(ref: bool = False):
    targets = ["all"] + (["ref"] if ref else [])
    _sp.run(["make", "-s", "-C", HERE] + targets, check=True)


def library():
    global _lib
    if _lib is None:
        if not _os.path.exists(LIB):
            build()
        lib = _C.CDLL(LIB)
        _capi.declare(lib, "oracle_", extra_create_int=True)
        lib.oracle_is_greater.restype = _C.c_int
        lib.oracle_is_greater.argtypes = [_C.c_double, _C.c_double]
        lib.oracle_is_less.restype = _C.c_int
        lib.oracle_is_less.argtypes = [_C.c_double, _C.c_double]
        _lib = lib
    return _lib


def ref_library():
    """The reference-source harness, or None where it is not built."""
    if not _os.path.exists(REF_LIB):
        return None
    return _C.CDLL(REF_LIB)


class OracleEngine(_capi.Engine):
    def __init__(self, spec, nthreads: int = 0):
        super().__init__(library(), "oracle_", spec, _capi.PREC_FP64, nthreads)
