"""Helpers shared by the parity tests: run the GPU engine and the oracle side
by side on one spec and compare every state field."""
import numpy as np

from peridem_amd import capi

VEC_FIELDS = {"f": capi.FIELD_F, "u": capi.FIELD_U, "v": capi.FIELD_V, "x": capi.FIELD_X}


def rel_err(a, b):
    """max |a - b| / max |b| (errors normalised by the field's largest entry)."""
    den = np.abs(b).max()
    if den == 0.0:
        return float(np.abs(a - b).max())
    return float(np.abs(a - b).max() / den)


def compare_state(gpu, ref, tol=1e-12, tag=""):
    out = {}
    for name, fid in VEC_FIELDS.items():
        out[name] = rel_err(gpu.field(fid), ref.field(fid))
    zg, zr = gpu.field(capi.FIELD_Z), ref.field(capi.FIELD_Z)
    # damage keeps the reference's 1e-5-slack maximum: equal to 1e-5 relative
    out["Z"] = float(np.max(np.abs(zg.astype(np.float64) - zr) / np.maximum(np.abs(zr), 1e-30)))
    out["bonds_equal"] = bool(np.array_equal(gpu.field(capi.FIELD_BONDS), ref.field(capi.FIELD_BONDS)))
    for name in VEC_FIELDS:
        assert out[name] <= tol, f"{tag} field {name} rel err {out[name]:.3e} > {tol}"
    assert out["Z"] <= 1e-5, f"{tag} damage mismatch {out['Z']}"
    assert out["bonds_equal"], f"{tag} bond states differ"
    return out


def same_sets(rp_a, col_a, rp_b, col_b):
    if not np.array_equal(rp_a, rp_b):
        return False
    for i in range(len(rp_a) - 1):
        a = np.sort(col_a[rp_a[i]:rp_a[i + 1]])
        b = np.sort(col_b[rp_b[i]:rp_b[i + 1]])
        if not np.array_equal(a, b):
            return False
    return True
