"""FP64 parity on the code paths the configuration workloads leave out:
3D bond-based RNP, the Gaussian influence function, reversible RNP bonds,
every displacement time function (constant, linear, quadratic, sin,
rotation) and force time function (constant, linear, linear_step,
linear_slow_fast, sin) on region selections, and the capacity error path.
Bar: 1e-12 of each field's maximum and identical bond states (exp, sin and
cos of the device library may differ from the host's in the last ulp, so
these cases are equal to ~1e-13 rather than bit for bit)."""
import math

import numpy as np
import pytest

from peridem_amd import capi, configs
from peridem_amd.deck import _Region
from peridem_amd.geometry import ball_mesh, center_node, disk_mesh
from peridem_amd.model import BcDeck, ModelBuilder, ZoneDeck
import oracle

from parity_util import compare_state

pytestmark = pytest.mark.gpu


def _single(dim, mat, ubc=(), fbc=(), gravity=(0.0, 0.0, 0.0), R=1e-3, h=1.0e-4):
    if dim == 2:
        x, vol = disk_mesh(R, h)
    else:
        x, vol = ball_mesh(R, 2.5 * h)
    b = ModelBuilder(dim, 2e-8, [ZoneDeck(mat)], {}, multi_particle=False, gravity=gravity)
    b.add_particle(0, x, vol, center_node(x, (0, 0, 0), R), R)
    b.ubc.extend(ubc)
    b.fbc.extend(fbc)
    return b.build(name=f"paths_{dim}d")


def _box(dim, lo, hi):
    if dim == 2:
        return _Region("rectangle", [lo[0], lo[1], 0.0, hi[0], hi[1], 0.0], 2)
    return _Region("cuboid", list(lo) + list(hi), 3)


def _cases():
    R = 1e-3
    out = {}
    # 3D RNP, gravity, sin displacement on the bottom cap, linear_step force on the top cap
    m = configs.rnp_material()
    out["rnp3d_sin_linear_step"] = _single(
        3, m, gravity=(0.0, 0.0, -9.81),
        ubc=[BcDeck([], [3], capi.UFN_SIN, [2e-6, 2e4], region=_box(3, (-R, -R, -R), (R, R, -0.7 * R)))],
        fbc=[BcDeck([], [1, 3], capi.FFN_LINEAR_STEP, [5e9, 0.0, 1e-6, 0.0],
                    region=_box(3, (-R, -R, 0.7 * R), (R, R, R)))])
    # 2D PDState, Gaussian influence, rotation + quadratic displacement, sin force
    ms = configs.pdstate_material(inf=capi.INF_GAUSSIAN)
    out["pdstate2d_gauss_rotation"] = _single(
        2, ms,
        ubc=[BcDeck([], [1, 2], capi.UFN_ROTATION, [200.0, 0.0, 0.0, 0.0],
                    region=_box(2, (-R, -R), (-0.6 * R, R))),
             BcDeck([], [2], capi.UFN_QUADRATIC, [1e-3, 5e2], region=_box(2, (0.7 * R, -R), (R, R)))],
        fbc=[BcDeck([], [1], capi.FFN_SIN, [1e8, 3e4], region=_box(2, (-R, 0.6 * R), (R, R)))])
    # 2D RNP with reversible bonds, linear displacement pull and linear_slow_fast force
    mr = configs.rnp_material()
    mr.irreversible = False
    out["rnp2d_reversible"] = _single(
        2, mr,
        ubc=[BcDeck([], [1], capi.UFN_LINEAR, [-5.0], region=_box(2, (-R, -R), (-0.7 * R, R))),
             BcDeck([], [1], capi.UFN_CONSTANT, [0.0], region=_box(2, (0.7 * R, -R), (R, R)), is_zero=True)],
        fbc=[BcDeck([], [2], capi.FFN_LINEAR_SLOW_FAST, [1e9, 2e-7, 1.0, 4.0],
                    region=_box(2, (-0.2 * R, -0.2 * R), (0.2 * R, 0.2 * R))),
             BcDeck([], [1], capi.FFN_CONSTANT, [1e8], region=_box(2, (-0.2 * R, 0.5 * R), (0.2 * R, R)))])
    return out


CASES = _cases()


@pytest.mark.parametrize("name", sorted(CASES))
def test_path_parity(name):
    spec = CASES[name]
    print(f"[paths] {name} nodes={spec.n_nodes}", flush=True)
    g = capi.GpuEngine(spec, capi.PREC_FP64)
    r = oracle.OracleEngine(spec, 8)
    g.integrate_begin()
    r.integrate_begin()
    for k in range(3):
        g.integrate_cd(100)
        r.integrate_cd(100)
        out = compare_state(g, r, tag=f"{name} block {k}")
        print(f"[paths] {name} block {k}: {out}", flush=True)
    assert np.abs(r.field(capi.FIELD_F)).max() > 0


def test_capacity_overflow_is_an_error():
    # a search radius of 10 contact radii gives ~120 candidates at the
    # touching nodes: a capacity of 32 per node overflows at the first search
    spec = configs.two_particle(gap=0.0, speed=0.1, dt=1.0e-7)
    spec.search_factor = 10.0
    with pytest.raises(RuntimeError, match="capacity"):
        g = capi.GpuEngine(spec, capi.PREC_FP64, contact_capacity=32)
        g.integrate_begin()
        g.integrate_cd(50)
