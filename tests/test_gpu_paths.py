"""FP64 parity on the code paths the configuration workloads leave out:
3D bond-based RNP, the Gaussian influence function, reversible RNP bonds,
bond-based PMB (with breaking) and PdElastic in 2D and 3D, a PDState and a
PMB particle in contact (FP64 and mixed precision),
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
from peridem_amd.model import BcDeck, ContactPairDeck, ModelBuilder, ZoneDeck
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
    # 2D PMB pulled fast on the left cap: bonds break near the pull
    mp = configs.pmb_material(Gc=1.0)
    out["pmb2d_break"] = _single(
        2, mp,
        ubc=[BcDeck([], [1], capi.UFN_LINEAR, [-20.0], region=_box(2, (-R, -R), (-0.7 * R, R))),
             BcDeck([], [1, 2], capi.UFN_CONSTANT, [0.0], region=_box(2, (0.7 * R, -R), (R, R)), is_zero=True)])
    # 3D PMB, Gaussian influence, sin displacement and sin force
    mg = configs.pmb_material(Gc=20.0, inf=capi.INF_GAUSSIAN)
    out["pmb3d_gauss_sin"] = _single(
        3, mg,
        ubc=[BcDeck([], [3], capi.UFN_SIN, [4e-6, 3e4], region=_box(3, (-R, -R, -R), (R, R, -0.7 * R)))],
        fbc=[BcDeck([], [1], capi.FFN_SIN, [1e9, 3e4], region=_box(3, (-R, -R, 0.7 * R), (R, R, R)))])
    # 2D and 3D PdElastic (never breaks), linear influence, fast pull
    me = configs.pdelastic_material(inf=capi.INF_LINEAR)
    out["pdelastic2d_linear"] = _single(
        2, me,
        ubc=[BcDeck([], [1], capi.UFN_LINEAR, [-20.0], region=_box(2, (-R, -R), (-0.7 * R, R)))],
        fbc=[BcDeck([], [2], capi.FFN_CONSTANT, [1e9], region=_box(2, (0.6 * R, -R), (R, R)))])
    out["pdelastic3d_const"] = _single(
        3, configs.pdelastic_material(), gravity=(0.0, 0.0, -9.81),
        ubc=[BcDeck([], [1, 2, 3], capi.UFN_LINEAR, [-10.0], region=_box(3, (-R, -R, -R), (R, R, -0.7 * R)))])
    return out


CASES = _cases()


def _state_pmb_pair(speed=3.0):
    """A PDState disk (zone 0, first particle) hits a PMB disk (zone 1)."""
    R, h = 1.0e-3, 1.0e-4
    ms = configs.pdstate_material()
    mp = configs.pmb_material(Gc=50.0)
    cp = ContactPairDeck(contact_radius_factor=0.95, Kn=18.0 * ms.K / (math.pi * (3 * h) ** 5), Kn_factor=1e-3,
                         damping_on=True, friction_on=True, friction_coeff=0.3, epsilon=0.8)
    b = ModelBuilder(2, 2e-8, [ZoneDeck(ms), ZoneDeck(mp)], {(0, 0): cp, (0, 1): cp, (1, 1): cp},
                     search_factor=5.0, neigh_update_interval=20)
    x, vol = disk_mesh(R, h)
    cl = center_node(x, (0, 0, 0), R)
    b.add_particle(0, x, vol, cl, R)
    b.add_particle(1, x + np.array([0.0, 2 * R, 0.0]), vol, cl, R)
    b.ic.append(([1], (0.0, -speed, 0.0)))
    return b.build(name="state_pmb_pair_2d")


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


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_state_and_pmb_particles_in_contact(precision):
    """Two laws in one model: the PDState particle runs the state kernels
    (FP32 tiles in mixed precision), the PMB particle the FP64 bond kernel."""
    spec = _state_pmb_pair()
    prec = capi.PREC_FP64 if precision == "fp64" else capi.PREC_MIXED
    print(f"[paths] state+pmb {precision} nodes={spec.n_nodes}", flush=True)
    g = capi.GpuEngine(spec, prec)
    r = oracle.OracleEngine(spec, 8)
    g.integrate_begin()
    r.integrate_begin()
    g.integrate_cd(600)
    r.integrate_cd(600)
    if precision == "fp64":
        out = compare_state(g, r, tag="state+pmb")
        print(f"[paths] state+pmb fp64: {out}", flush=True)
    else:
        f, fr = g.field(capi.FIELD_F), r.field(capi.FIELD_F)
        n1 = spec.particles[1]["glob_start"]
        err = float(np.abs(f - fr).max() / np.abs(fr).max())
        err_pmb = float(np.abs(f[n1:] - fr[n1:]).max() / np.abs(fr[n1:]).max())
        print(f"[paths] state+pmb mixed: force err {err:.3e}, PMB particle {err_pmb:.3e}", flush=True)
        assert err < 1e-4 and err_pmb < 1e-4
    # the particles touched: contact forces acted on both
    v = r.field(capi.FIELD_V)
    cn = [p["center_node"] for p in spec.particles]
    assert v[cn[0], 1] < 0.0 and v[cn[1], 1] > -3.0


def test_capacity_overflow_is_an_error():
    # a search radius of 10 contact radii gives ~120 candidates at the
    # touching nodes: a capacity of 32 per node overflows at the first search
    spec = configs.two_particle(gap=0.0, speed=0.1, dt=1.0e-7)
    spec.search_factor = 10.0
    with pytest.raises(RuntimeError, match="capacity"):
        g = capi.GpuEngine(spec, capi.PREC_FP64, contact_capacity=32)
        g.integrate_begin()
        g.integrate_cd(50)


@pytest.mark.parametrize("case", ["compressive", "rnp2d_pull_break", "rnp3d"])
def test_rnp_kernels_agree(case):
    """The lane-per-node RNP kernel (workgroups of 64 and 256) and the
    pair-tile kernel (each bond scalar once per pair, pd_tpb 512) give the
    same step bit for bit: forces, displacements, velocities, bond bits and
    damage, with bonds breaking."""
    R = 1e-3
    if case == "compressive":
        spec = configs.compressive_2d(n_particles=12, nodes_per_particle=150, seed=3, init_speed=1.0,
                                      pitch_factor=1.01, R_lo=1.2e-3, R_hi=1.2e-3, friction=True)
    elif case == "rnp2d_pull_break":
        m = configs.rnp_material(Gc=5.0)
        spec = _single(2, m, ubc=[BcDeck([], [1], capi.UFN_LINEAR, [-20.0],
                                         region=_box(2, (-R, -R), (-0.7 * R, R)))])
    else:
        spec = CASES["rnp3d_sin_linear_step"]
    runs = {}
    for tpb in (64, 256, 512):
        g = capi.GpuEngine(spec, capi.PREC_FP64)
        g.set_launch_config(tpb)
        assert g.launch_config()[0] == tpb
        g.integrate_begin()
        g.integrate_cd_graph(300)
        runs[tpb] = {f: g.field(f) for f in (capi.FIELD_F, capi.FIELD_U, capi.FIELD_V, capi.FIELD_Z,
                                              capi.FIELD_BONDS)}
        g.close()
    nb = int(runs[512][capi.FIELD_BONDS].sum())
    print(f"[paths] rnp kernels {case}: nodes={spec.n_nodes} broken bonds {nb}", flush=True)
    for tpb in (64, 256):
        for f, a in runs[tpb].items():
            np.testing.assert_array_equal(a, runs[512][f])
    if case == "rnp2d_pull_break":
        assert nb > 0
