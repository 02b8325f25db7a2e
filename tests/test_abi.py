"""The C-ABI library: it loads without a GPU, exports every entry point that
include/pdem.h declares, and its structure layouts match the ctypes binding.
No compute call is made here."""
import ctypes as C
import os
import re

import pytest

from peridem_amd import capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "pdem.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pdem_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = capi.engine_library()
    syms = declared_symbols()
    assert len(syms) >= 20
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing


def test_struct_layouts_match_binding():
    lib = capi.engine_library()
    lib.pdem_abi_sizeof.restype = C.c_int64
    lib.pdem_abi_sizeof.argtypes = [C.c_int32]
    for which, st in enumerate([capi.Desc, capi.Particle, capi.ContactPair, capi.BcSet,
                                capi.ContactState, capi.MaterialDeckC]):
        assert lib.pdem_abi_sizeof(which) == C.sizeof(st), st.__name__
    assert lib.pdem_abi_sizeof(99) == -1


def test_create_rejects_bad_descriptors_before_touching_a_device():
    lib = capi.engine_library()
    ctx = C.c_void_p()
    assert lib.pdem_create(None, C.byref(ctx)) != 0
    assert b"null descriptor" in lib.pdem_last_error(None)
    d = capi.Desc()
    d.dim = 4
    assert lib.pdem_create(C.byref(d), C.byref(ctx)) != 0
    assert b"dim" in lib.pdem_last_error(None)
    d.dim = 2
    d.n_nodes = 0
    assert lib.pdem_create(C.byref(d), C.byref(ctx)) != 0
    assert not ctx.value


def test_context_calls_reject_null():
    lib = capi.engine_library()
    assert lib.pdem_integrate_cd(None, 1) != 0
    assert lib.pdem_compute_forces(None) != 0


def test_product_does_not_reference_the_oracle():
    # the engine package never imports or loads the checker
    pkg = os.path.join(ROOT, "peridem_amd")
    for dirpath, _, files in os.walk(pkg):
        for fn in files:
            if fn.endswith((".py", ".hip", ".hpp", ".cpp", ".h")):
                text = open(os.path.join(dirpath, fn)).read()
                assert "import oracle" not in text and "liboracle" not in text, fn
