"""ctypes mirror of include/pdem.h, the C-ABI of the GPU engine.

The structures below must stay byte-identical to the C declarations; the
tests check their sizes against the values the library reports.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

MAT_RNP = 0
MAT_PDSTATE = 1
MAT_PMB = 2
MAT_PDELASTIC = 3

INF_CONST = 0
INF_LINEAR = 1
INF_GAUSSIAN = 2

PREC_FP64 = 0
PREC_MIXED = 1

UFN_CONSTANT, UFN_LINEAR, UFN_QUADRATIC, UFN_SIN, UFN_ROTATION, UFN_NONE = range(6)
FFN_CONSTANT, FFN_LINEAR, FFN_LINEAR_STEP, FFN_LINEAR_SLOW_FAST, FFN_SIN = range(5)

FIELD_X, FIELD_U, FIELD_V, FIELD_F, FIELD_XREF, FIELD_VOL, FIELD_Z, FIELD_THETA, \
    FIELD_MX, FIELD_VMAG, FIELD_BONDS = range(11)
STOP_MAX_PARTICLE_DIST, STOP_MAX_NODE_DIST = 1, 2
NEIGH_PD = 0
NEIGH_CONTACT = 1


class Particle(C.Structure):
    _fields_ = [
        ("glob_start", C.c_int64), ("glob_end", C.c_int64), ("center_node", C.c_int64),
        ("zone", C.c_int32), ("type_index", C.c_int32),
        ("all_dofs_constrained", C.c_int32), ("compute_force", C.c_int32),
        ("material_type", C.c_int32), ("irreversible", C.c_int32),
        ("density", C.c_double), ("horizon", C.c_double), ("mesh_size", C.c_double),
        ("bounding_radius", C.c_double), ("Kn", C.c_double), ("Rc", C.c_double),
        ("mat", C.c_double * 6),
    ]


class ContactPair(C.Structure):
    _fields_ = [("contact_r", C.c_double), ("Kn", C.c_double), ("mu", C.c_double),
                ("betan", C.c_double), ("kappa", C.c_double)]


class BcSet(C.Structure):
    _fields_ = [("time_fn", C.c_int32), ("dir_mask", C.c_int32), ("is_zero", C.c_int32),
                ("reserved", C.c_int32), ("time_params", C.c_double * 4),
                ("n_entries", C.c_int64), ("nodes", C.POINTER(C.c_int64)),
                ("amp", C.POINTER(C.c_double))]


class Desc(C.Structure):
    _fields_ = [
        ("dim", C.c_int32), ("precision", C.c_int32), ("multi_particle", C.c_int32),
        ("influence_fn_type", C.c_int32), ("influence_fn_params", C.c_double * 2),
        ("dt", C.c_double), ("n_nodes", C.c_int64),
        ("x_ref", C.POINTER(C.c_double)), ("u", C.POINTER(C.c_double)),
        ("v", C.POINTER(C.c_double)), ("vol", C.POINTER(C.c_double)),
        ("fix", C.POINTER(C.c_uint8)), ("pt_id", C.POINTER(C.c_int64)),
        ("n_particles", C.c_int32), ("n_zones", C.c_int32),
        ("particles", C.POINTER(Particle)), ("contact", C.POINTER(ContactPair)),
        ("gravity", C.c_double * 3), ("search_factor", C.c_double),
        ("neigh_update_interval", C.c_int64),
        ("n_ubc", C.c_int32), ("n_fbc", C.c_int32),
        ("ubc", C.POINTER(BcSet)), ("fbc", C.POINTER(BcSet)),
        ("contact_capacity", C.c_int64),
    ]


class ContactState(C.Structure):
    _fields_ = [("interval", C.c_int64), ("counter", C.c_int64),
                ("search_radius", C.c_double), ("max_velocity", C.c_double),
                ("max_contact_r", C.c_double), ("n_rebuilds", C.c_int64),
                ("overflow", C.c_int64), ("capacity", C.c_int64)]


class MaterialDeckC(C.Structure):
    """pdem_material_deck: the deck values of pdem_material_constants."""
    _fields_ = [("material_type", C.c_int32), ("dim", C.c_int32), ("compute_from_classical", C.c_int32),
                ("influence_fn_type", C.c_int32), ("n_influence_fn_params", C.c_int32),
                ("n_bond_potential_params", C.c_int32), ("influence_fn_params", C.c_double * 2),
                ("E", C.c_double), ("K", C.c_double), ("G", C.c_double), ("nu", C.c_double),
                ("Gc", C.c_double), ("KIc", C.c_double), ("horizon", C.c_double),
                ("check_sc_factor", C.c_double), ("bond_potential_params", C.c_double * 3)]


def material_constants(material_type, dim, horizon, E=-1.0, K=-1.0, G=-1.0, nu=-1.0, Gc=-1.0, KIc=-1.0,
                       influence_fn_type=INF_CONST, influence_fn_params=(), check_sc_factor=1.0,
                       compute_from_classical=True, bond_potential_params=()):
    """pdem_material_constants: (mat[6], K_contact) of one particle, computed by
    the engine library's host code (what a C++ DEMModel subclass calls)."""
    lib = engine_library()
    d = MaterialDeckC()
    d.material_type, d.dim, d.compute_from_classical = material_type, dim, int(compute_from_classical)
    d.influence_fn_type = influence_fn_type
    d.n_influence_fn_params = len(influence_fn_params)
    for k, v in enumerate(influence_fn_params):
        d.influence_fn_params[k] = v
    d.n_bond_potential_params = len(bond_potential_params)
    for k, v in enumerate(bond_potential_params):
        d.bond_potential_params[k] = v
    d.E, d.K, d.G, d.nu, d.Gc, d.KIc = E, K, G, nu, Gc, KIc
    d.horizon, d.check_sc_factor = horizon, check_sc_factor
    mat = (C.c_double * 6)()
    kc = C.c_double()
    if lib.pdem_material_constants(C.byref(d), mat, C.byref(kc)) != 0:
        raise ValueError(lib.pdem_material_last_error().decode())
    return list(mat), kc.value


def _ptr(a: np.ndarray, ctype):
    return a.ctypes.data_as(C.POINTER(ctype))


class DescHolder:
    """Keeps the numpy buffers a Desc points into alive."""

    def __init__(self, spec, precision: int = PREC_FP64, contact_capacity: int = 0):
        self._keep = []
        s = spec

        def keep(a, dtype):
            a = np.ascontiguousarray(a, dtype=dtype)
            self._keep.append(a)
            return a

        n = s.n_nodes
        xr = keep(s.x_ref, np.float64)
        u = keep(s.u, np.float64)
        v = keep(s.v, np.float64)
        vol = keep(s.vol, np.float64)
        fix = keep(s.fix, np.uint8)
        pt = keep(s.pt_id, np.int64)
        parts = (Particle * len(s.particles))()
        for k, p in enumerate(s.particles):
            for name, _ in Particle._fields_:
                if name == "mat":
                    for q in range(6):
                        parts[k].mat[q] = float(p["mat"][q]) if q < len(p["mat"]) else 0.0
                else:
                    setattr(parts[k], name, p[name])
        self._keep.append(parts)
        nz = s.n_zones
        ct = (ContactPair * max(1, nz * nz))()
        for a in range(nz):
            for b in range(nz):
                c = s.contact[a][b]
                ct[a * nz + b] = ContactPair(c["contact_r"], c["Kn"], c["mu"], c["betan"],
                                             c["kappa"])
        self._keep.append(ct)

        def bcs(sets):
            arr = (BcSet * max(1, len(sets)))()
            for k, b in enumerate(sets):
                nodes = keep(b["nodes"], np.int64)
                amp = keep(b["amp"], np.float64)
                tp = list(b["time_params"]) + [0.0] * (4 - len(b["time_params"]))
                arr[k] = BcSet(b["time_fn"], b["dir_mask"], int(b.get("is_zero", 0)), 0,
                               (C.c_double * 4)(*tp[:4]), len(nodes), _ptr(nodes, C.c_int64),
                               _ptr(amp, C.c_double))
            self._keep.append(arr)
            return arr

        ubc = bcs(s.ubc)
        fbc = bcs(s.fbc)
        d = Desc()
        d.dim = s.dim
        d.precision = precision
        d.multi_particle = int(s.multi_particle)
        d.influence_fn_type = s.influence_fn_type
        d.influence_fn_params[0] = s.influence_fn_params[0]
        d.influence_fn_params[1] = s.influence_fn_params[1]
        d.dt = s.dt
        d.n_nodes = n
        d.x_ref = _ptr(xr, C.c_double)
        d.u = _ptr(u, C.c_double)
        d.v = _ptr(v, C.c_double)
        d.vol = _ptr(vol, C.c_double)
        d.fix = _ptr(fix, C.c_uint8)
        d.pt_id = _ptr(pt, C.c_int64)
        d.n_particles = len(s.particles)
        d.n_zones = nz
        d.particles = C.cast(parts, C.POINTER(Particle))
        d.contact = C.cast(ct, C.POINTER(ContactPair))
        for q in range(3):
            d.gravity[q] = s.gravity[q]
        d.search_factor = s.search_factor
        d.neigh_update_interval = s.neigh_update_interval
        d.n_ubc = len(s.ubc)
        d.n_fbc = len(s.fbc)
        d.ubc = C.cast(ubc, C.POINTER(BcSet))
        d.fbc = C.cast(fbc, C.POINTER(BcSet))
        d.contact_capacity = contact_capacity
        self.desc = d


def declare(lib: C.CDLL, prefix: str, extra_create_int: bool = False):
    """Declare argtypes/restype for the pdem_* family (or the oracle_* twin)."""
    vp = C.c_void_p

    def f(name, res, args):
        fn = getattr(lib, prefix + name)
        fn.restype = res
        fn.argtypes = args
        return fn

    if extra_create_int:
        f("create", C.c_int, [C.POINTER(Desc), C.c_int32, C.POINTER(vp)])
    else:
        f("create", C.c_int, [C.POINTER(Desc), C.POINTER(vp)])
    f("destroy", None, [vp])
    f("last_error", C.c_char_p, [vp])
    for name in ("integrate_begin", "compute_forces", "compute_peridynamic_forces"):
        f(name, C.c_int, [vp])
    f("integrate_cd", C.c_int, [vp, C.c_int64])
    f("integrate_verlet", C.c_int, [vp, C.c_int64])
    f("get_field", C.c_int, [vp, C.c_int32, vp])
    f("set_field", C.c_int, [vp, C.c_int32, vp])
    f("get_time", C.c_int, [vp, C.POINTER(C.c_int64), C.POINTER(C.c_double)])
    f("get_neighbor_count", C.c_int, [vp, C.c_int32, C.POINTER(C.c_int64)])
    f("get_neighbors", C.c_int, [vp, C.c_int32, C.POINTER(C.c_int64), C.POINTER(C.c_int64)])
    f("get_contact_state", C.c_int, [vp, C.POINTER(ContactState)])
    f("pp_compressive_test", C.c_int,
      [vp, C.c_int32, C.c_int32, C.POINTER(C.c_double), C.POINTER(C.c_double)])
    f("broken_bonds", C.c_int, [vp, C.POINTER(C.c_int64), C.POINTER(C.c_int64)])
    f("pp_two_particle_test", C.c_int, [vp, C.POINTER(C.c_double)])
    f("restart", C.c_int, [vp, C.c_int64])
    f("check_stop", C.c_int, [vp, C.c_int32, C.c_double, C.POINTER(C.c_int32), C.POINTER(C.c_double),
                              C.POINTER(C.c_int64)])
    return lib


class Engine:
    """Thin object wrapper over one context of a pdem_*-shaped library."""

    def __init__(self, lib, prefix, spec, precision=PREC_FP64, nthreads=None,
                 contact_capacity=0):
        self.lib = lib
        self.p = prefix
        self.spec = spec
        self.holder = DescHolder(spec, precision, contact_capacity)
        self.ctx = C.c_void_p()
        if nthreads is not None:
            rc = self._f("create")(C.byref(self.holder.desc), nthreads, C.byref(self.ctx))
        else:
            rc = self._f("create")(C.byref(self.holder.desc), C.byref(self.ctx))
        if rc != 0:
            raise RuntimeError(f"{prefix}create failed: {self._f('last_error')(None).decode()}")

    def _f(self, name):
        return getattr(self.lib, self.p + name)

    def _check(self, rc, what):
        if rc != 0:
            msg = self._f("last_error")(self.ctx).decode()
            raise RuntimeError(f"{self.p}{what} failed: {msg}")

    def close(self):
        if self.ctx:
            self._f("destroy")(self.ctx)
            self.ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ------------------------------------------------------------ stepping
    def integrate_begin(self):
        self._check(self._f("integrate_begin")(self.ctx), "integrate_begin")

    def integrate_cd(self, nsteps=1):
        self._check(self._f("integrate_cd")(self.ctx, nsteps), "integrate_cd")

    def integrate_verlet(self, nsteps=1):
        self._check(self._f("integrate_verlet")(self.ctx, nsteps), "integrate_verlet")

    def integrate(self, nsteps=1):
        """DEMModel::integrateStep (demModel.cpp:342-347) nsteps times: the
        model's time scheme, central difference or velocity Verlet."""
        if self.spec.time_scheme == "velocity_verlet":
            self.integrate_verlet(nsteps)
        else:
            self.integrate_cd(nsteps)

    def compute_forces(self):
        self._check(self._f("compute_forces")(self.ctx), "compute_forces")

    def compute_peridynamic_forces(self):
        self._check(self._f("compute_peridynamic_forces")(self.ctx), "compute_peridynamic_forces")

    # ------------------------------------------------------------ queries
    def field(self, fid):
        n = self.spec.n_nodes
        if fid in (FIELD_X, FIELD_U, FIELD_V, FIELD_F, FIELD_XREF):
            a = np.zeros((n, 3), np.float64)
        elif fid == FIELD_Z:
            a = np.zeros(n, np.float32)
        elif fid == FIELD_BONDS:
            a = np.zeros(self.neighbor_count(NEIGH_PD), np.uint8)
        else:
            a = np.zeros(n, np.float64)
        self._check(self._f("get_field")(self.ctx, fid, a.ctypes.data_as(C.c_void_p)), "get_field")
        return a

    def set_field(self, fid, a):
        a = np.ascontiguousarray(a)
        self._check(self._f("set_field")(self.ctx, fid, a.ctypes.data_as(C.c_void_p)), "set_field")

    def time(self):
        n = C.c_int64()
        t = C.c_double()
        self._check(self._f("get_time")(self.ctx, C.byref(n), C.byref(t)), "get_time")
        return n.value, t.value

    def neighbor_count(self, kind):
        m = C.c_int64()
        self._check(self._f("get_neighbor_count")(self.ctx, kind, C.byref(m)), "get_neighbor_count")
        return m.value

    def neighbors(self, kind):
        m = self.neighbor_count(kind)
        rp = np.zeros(self.spec.n_nodes + 1, np.int64)
        cols = np.zeros(max(1, m), np.int64)
        self._check(self._f("get_neighbors")(self.ctx, kind, rp.ctypes.data_as(C.POINTER(C.c_int64)),
                                             cols.ctypes.data_as(C.POINTER(C.c_int64))),
                    "get_neighbors")
        return rp, cols[:m]

    def contact_state(self):
        st = ContactState()
        self._check(self._f("get_contact_state")(self.ctx, C.byref(st)), "get_contact_state")
        return {k: getattr(st, k) for k, _ in ContactState._fields_}

    def pp_two_particle_test(self):
        """DEMModel::ppTwoParticleTest key data: pen_dist, contact_area_radius,
        max_y, max_dist, max_y_loc."""
        out = (C.c_double * 5)()
        self._check(self._f("pp_two_particle_test")(self.ctx, out), "pp_two_particle_test")
        return dict(zip(("pen_dist", "contact_area_radius", "max_y", "max_dist", "max_y_loc"), list(out)))

    def restart(self, step):
        """DEMModel::restart: step counter, time and the controller's restart rule."""
        self._check(self._f("restart")(self.ctx, int(step)), "restart")

    def check_stop(self, criterion, param):
        """(stop, value, node) of a checkStop criterion (STOP_MAX_PARTICLE_DIST
        or STOP_MAX_NODE_DIST)."""
        st, v, nd = C.c_int32(), C.c_double(), C.c_int64()
        self._check(self._f("check_stop")(self.ctx, criterion, param, C.byref(st), C.byref(v), C.byref(nd)),
                    "check_stop")
        return bool(st.value), v.value, nd.value

    def pp_compressive_test(self, wall_id, f_dir):
        pen = C.c_double()
        reac = C.c_double()
        self._check(self._f("pp_compressive_test")(self.ctx, wall_id, f_dir, C.byref(pen),
                                                   C.byref(reac)), "pp_compressive_test")
        return pen.value, reac.value

    def broken_bonds(self):
        b = C.c_int64()
        t = C.c_int64()
        self._check(self._f("broken_bonds")(self.ctx, C.byref(b), C.byref(t)), "broken_bonds")
        return b.value, t.value


_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libpdem_hip.so")

_engine_lib = None


def engine_library():
    """Load the HIP engine library; raises if it has not been built."""
    global _engine_lib
    if _engine_lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"HIP engine library missing: {LIB_PATH} (run __graft_entry__.build())")
        lib = C.CDLL(LIB_PATH)
        declare(lib, "pdem_")
        lib.pdem_integrate_cd_graph.restype = C.c_int
        lib.pdem_integrate_cd_graph.argtypes = [C.c_void_p, C.c_int64]
        lib.pdem_integrate_verlet_graph.restype = C.c_int
        lib.pdem_integrate_verlet_graph.argtypes = [C.c_void_p, C.c_int64]
        lib.pdem_fe_setup.restype = C.c_int
        lib.pdem_fe_setup.argtypes = [C.c_void_p, C.c_int32, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p]
        lib.pdem_fe_quad_count.restype = C.c_int
        lib.pdem_fe_quad_count.argtypes = [C.c_void_p, C.POINTER(C.c_int64)]
        lib.pdem_strain_stress.restype = C.c_int
        lib.pdem_strain_stress.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        lib.pdem_synchronize.restype = C.c_int
        lib.pdem_synchronize.argtypes = [C.c_void_p]
        for nm in ("update_contact_neighborlist", "compute_contact_forces",
                   "compute_external_forces", "compute_external_displacement_bc"):
            fn = getattr(lib, "pdem_" + nm)
            fn.restype = C.c_int
            fn.argtypes = [C.c_void_p]
        lib.pdem_device_count.restype = C.c_int
        lib.pdem_device_count.argtypes = []
        lib.pdem_profile_enable.restype = C.c_int
        lib.pdem_profile_enable.argtypes = [C.c_void_p, C.c_int32]
        lib.pdem_profile_get.restype = C.c_int
        lib.pdem_profile_get.argtypes = [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int64)]
        lib.pdem_closure_setup.restype = C.c_int
        lib.pdem_closure_setup.argtypes = [C.c_void_p, C.POINTER(C.c_float)]
        lib.pdem_closure_enable.restype = C.c_int
        lib.pdem_closure_enable.argtypes = [C.c_void_p, C.c_void_p, C.c_double, C.c_double, C.c_int32]
        lib.pdem_closure_disable.restype = C.c_int
        lib.pdem_closure_disable.argtypes = [C.c_void_p]
        lib.pdem_closure_stats.restype = C.c_int
        lib.pdem_closure_stats.argtypes = [C.c_void_p, C.POINTER(C.c_int64), C.POINTER(C.c_int64),
                                           C.POINTER(C.c_double), C.POINTER(C.c_double),
                                           C.POINTER(C.c_double)]
        lib.pdem_pair_log.restype = C.c_int
        lib.pdem_pair_log.argtypes = [C.c_void_p, C.c_double, C.c_int64, C.POINTER(C.c_float),
                                      C.POINTER(C.c_float), C.POINTER(C.c_float), C.POINTER(C.c_int64)]
        lib.pdem_material_constants.restype = C.c_int
        lib.pdem_material_constants.argtypes = [C.POINTER(MaterialDeckC), C.POINTER(C.c_double),
                                                C.POINTER(C.c_double)]
        lib.pdem_material_last_error.restype = C.c_char_p
        lib.pdem_material_last_error.argtypes = []
        lib.pdem_pair_log_noalloc.restype = C.c_int
        lib.pdem_pair_log_noalloc.argtypes = lib.pdem_pair_log.argtypes
        lib.pdem_checkpoint_size.restype = C.c_int
        lib.pdem_checkpoint_size.argtypes = [C.c_void_p, C.POINTER(C.c_int64)]
        lib.pdem_checkpoint_save.restype = C.c_int
        lib.pdem_checkpoint_save.argtypes = [C.c_void_p, C.c_void_p, C.c_int64]
        lib.pdem_checkpoint_load.restype = C.c_int
        lib.pdem_checkpoint_load.argtypes = [C.c_void_p, C.c_void_p, C.c_int64]
        lib.pdem_autotune.restype = C.c_int
        lib.pdem_autotune.argtypes = [C.c_void_p, C.c_int32, C.POINTER(C.c_double)]
        lib.pdem_get_launch_config.restype = C.c_int
        lib.pdem_get_launch_config.argtypes = [C.c_void_p] + [C.POINTER(C.c_int32)] * 3
        lib.pdem_set_launch_config.restype = C.c_int
        lib.pdem_set_launch_config.argtypes = [C.c_void_p] + [C.c_int32] * 3
        lib.pdem_get_tile_counts.restype = C.c_int
        lib.pdem_set_contact_light_max.restype = C.c_int
        lib.pdem_set_contact_light_max.argtypes = [C.c_void_p, C.c_int32]
        lib.pdem_get_tile_counts.argtypes = [C.c_void_p] + [C.POINTER(C.c_int32)] * 2
        lib.pdem_set_step_fusion.restype = C.c_int
        lib.pdem_set_step_fusion.argtypes = [C.c_void_p, C.c_int32]
        lib.pdem_set_state_fusion.restype = C.c_int
        lib.pdem_set_state_fusion.argtypes = [C.c_void_p, C.c_int32]
        _engine_lib = lib
    return _engine_lib


def device_count() -> int:
    """HIP devices visible to the engine library (0 without a GPU)."""
    return int(engine_library().pdem_device_count())


class GpuEngine(Engine):
    """The MI355X engine: one pdem_ctx on the current HIP device."""

    def __init__(self, spec, precision=PREC_FP64, contact_capacity=0):
        super().__init__(engine_library(), "pdem_", spec, precision, None, contact_capacity)

    def integrate_cd_graph(self, nsteps):
        self._check(self.lib.pdem_integrate_cd_graph(self.ctx, nsteps), "integrate_cd_graph")

    def fe_setup(self, quad_order, elem_nodes, elem_particle, elastic):
        """Element data for strain_stress (pdem_fe_setup)."""
        en = np.ascontiguousarray(elem_nodes, np.int64)
        ep = np.ascontiguousarray(elem_particle, np.int32)
        el = np.ascontiguousarray(elastic, np.float64)
        self._check(self.lib.pdem_fe_setup(self.ctx, int(quad_order), len(en), en.ctypes.data_as(C.c_void_p),
                                           ep.ctypes.data_as(C.c_void_p), el.ctypes.data_as(C.c_void_p)),
                    "fe_setup")

    def strain_stress(self):
        """Current quadrature points [nq][3], strain and stress [nq][6]
        (xx yy zz yz xz xy, float32) at the current displacement."""
        n = C.c_int64()
        self._check(self.lib.pdem_fe_quad_count(self.ctx, C.byref(n)), "fe_quad_count")
        xq = np.zeros((n.value, 3))
        eps = np.zeros((n.value, 6), np.float32)
        sig = np.zeros((n.value, 6), np.float32)
        self._check(self.lib.pdem_strain_stress(self.ctx, xq.ctypes.data_as(C.c_void_p),
                                                eps.ctypes.data_as(C.c_void_p), sig.ctypes.data_as(C.c_void_p)),
                    "strain_stress")
        return xq, eps, sig

    def integrate_verlet_graph(self, nsteps):
        self._check(self.lib.pdem_integrate_verlet_graph(self.ctx, nsteps), "integrate_verlet_graph")

    def integrate_graph(self, nsteps):
        """integrate() through the captured step of the model's time scheme."""
        if self.spec.time_scheme == "velocity_verlet":
            self.integrate_verlet_graph(nsteps)
        else:
            self.integrate_cd_graph(nsteps)

    def synchronize(self):
        self._check(self.lib.pdem_synchronize(self.ctx), "synchronize")

    def profile_enable(self, on=True):
        self._check(self.lib.pdem_profile_enable(self.ctx, int(on)), "profile_enable")

    def profile_get(self):
        ms = C.c_double()
        n = C.c_int64()
        self._check(self.lib.pdem_profile_get(self.ctx, C.byref(ms), C.byref(n)), "profile_get")
        return ms.value, n.value

    def checkpoint(self) -> bytes:
        """The complete step state; restore() on a context of the same model
        continues bit for bit."""
        n = C.c_int64()
        self._check(self.lib.pdem_checkpoint_size(self.ctx, C.byref(n)), "checkpoint_size")
        buf = C.create_string_buffer(n.value)
        self._check(self.lib.pdem_checkpoint_save(self.ctx, buf, n.value), "checkpoint_save")
        return buf.raw

    def restore(self, blob: bytes):
        self._check(self.lib.pdem_checkpoint_load(self.ctx, blob, len(blob)), "checkpoint_load")

    def autotune(self, reps=5):
        """Pick the fastest launch shape of the force pass; returns its ms."""
        ms = C.c_double()
        self._check(self.lib.pdem_autotune(self.ctx, reps, C.byref(ms)), "autotune")
        return ms.value

    def launch_config(self):
        """(pd_tpb, mixed_tpb, tile_nodes) currently in use."""
        v = [C.c_int32() for _ in range(3)]
        self._check(self.lib.pdem_get_launch_config(self.ctx, *[C.byref(x) for x in v]), "launch_config")
        return tuple(x.value for x in v)

    def set_launch_config(self, pd_tpb, mixed_tpb=1024, tile_nodes=1024):
        """Apply a launch shape (pd_tpb 512: the RNP pair-tile kernel)."""
        self._check(self.lib.pdem_set_launch_config(self.ctx, pd_tpb, mixed_tpb, tile_nodes), "set_launch_config")

    def set_contact_light_max(self, light_max):
        """Nodes with at most light_max contact candidates take the lane-per-node kernel."""
        self._check(self.lib.pdem_set_contact_light_max(self.ctx, int(light_max)), "set_contact_light_max")

    def tile_counts(self):
        """(tiles staged in LDS, tiles gathering from HBM) of the mixed bond kernels."""
        v = [C.c_int32() for _ in range(2)]
        self._check(self.lib.pdem_get_tile_counts(self.ctx, *[C.byref(x) for x in v]), "tile_counts")
        return tuple(x.value for x in v)

    def set_step_fusion(self, on=True):
        """Fused step kernels: False never; True (default) for multi-particle
        models of up to 2^20 nodes, and the fused node update and controller
        for larger models without sine or rotation displacement BCs; 2 all
        of it at any size (results unchanged)."""
        self._check(self.lib.pdem_set_step_fusion(self.ctx, int(on)), "set_step_fusion")

    def set_state_fusion(self, on=True):
        """Fused dilation and force passes for particles that fit one LDS tile
        (mixed precision; results unchanged)."""
        self._check(self.lib.pdem_set_state_fusion(self.ctx, int(on)), "set_state_fusion")

    def closure_setup(self, codes=None):
        """Closure data; codes: one float per particle (shape/material class)."""
        if codes is None:
            ptr = None
        else:
            self._codes = np.ascontiguousarray(codes, dtype=np.float32)
            ptr = self._codes.ctypes.data_as(C.POINTER(C.c_float))
        self._check(self.lib.pdem_closure_setup(self.ctx, ptr), "closure_setup")

    def closure_enable(self, net=None, pen_max_h=0.05, dz_max=0.5, check_every=100):
        self._closure_net = net  # keep the network alive while the context uses it
        h = net.h if net is not None else None
        self._check(self.lib.pdem_closure_enable(self.ctx, h, float(pen_max_h), float(dz_max), int(check_every)),
                    "closure_enable")

    def closure_disable(self):
        self._check(self.lib.pdem_closure_disable(self.ctx), "closure_disable")
        self._closure_net = None

    def closure_stats(self):
        modes = (C.c_int64 * 3)()
        chk = C.c_int64()
        em, ex, pm = C.c_double(), C.c_double(), C.c_double()
        self._check(self.lib.pdem_closure_stats(self.ctx, modes, C.byref(chk), C.byref(em), C.byref(ex),
                                                C.byref(pm)), "closure_stats")
        return {"no_contact": modes[0], "closure": modes[1], "plain": modes[2], "checked": chk.value,
                "check_err_mean": em.value, "check_err_max": ex.value, "check_plain_mean": pm.value}

    PAIR_LOG_NOALLOC_MAX_ROWS = 65536

    def pair_log(self, band_h=0.5, cap=1 << 20, baseline=False, noalloc=False):
        """Training rows of the current state: features [n][16], plain targets
        [n][4] and, with baseline=True, the normal-contact baseline [n][3].
        noalloc=True stages the rows in the library's fixed device array
        (pdem_pair_log_noalloc, cap <= PAIR_LOG_NOALLOC_MAX_ROWS) instead of
        device buffers allocated for the call."""
        f = np.empty((cap, 16), np.float32)
        t = np.empty((cap, 4), np.float32)
        b = np.empty((cap, 3), np.float32)
        n = C.c_int64()
        P = C.POINTER(C.c_float)
        fn = self.lib.pdem_pair_log_noalloc if noalloc else self.lib.pdem_pair_log
        self._check(fn(self.ctx, float(band_h), int(cap), f.ctypes.data_as(P),
                       t.ctypes.data_as(P), b.ctypes.data_as(P), C.byref(n)),
                    "pair_log_noalloc" if noalloc else "pair_log")
        m = min(n.value, cap)
        if baseline:
            return f[:m].copy(), t[:m].copy(), b[:m].copy()
        return f[:m].copy(), t[:m].copy()

    def update_contact_neighborlist(self):
        self._check(self.lib.pdem_update_contact_neighborlist(self.ctx), "update_contact_neighborlist")

    def compute_contact_forces(self):
        self._check(self.lib.pdem_compute_contact_forces(self.ctx), "compute_contact_forces")

    def compute_external_forces(self):
        self._check(self.lib.pdem_compute_external_forces(self.ctx), "compute_external_forces")

    def compute_external_displacement_bc(self):
        self._check(self.lib.pdem_compute_external_displacement_bc(self.ctx),
                    "compute_external_displacement_bc")


class Mlp:
    """The closure network (pdem_mlp_*): BF16-MFMA forward/backward, Adam on the GPU."""

    NIN, NH, NOUT = 16, 128, 4

    def __init__(self, seed=0):
        lib = engine_library()
        self.lib = lib
        P = C.POINTER
        f32p = P(C.c_float)
        sig = {
            "create": ([C.c_uint64, P(C.c_void_p)], C.c_int),
            "destroy": ([C.c_void_p], C.c_int),
            "last_error": ([C.c_void_p], C.c_char_p),
            "param_count": ([], C.c_int64),
            "get_params": ([C.c_void_p, f32p], C.c_int),
            "set_params": ([C.c_void_p, f32p], C.c_int),
            "set_normalization": ([C.c_void_p, f32p, f32p, f32p, f32p], C.c_int),
            "get_normalization": ([C.c_void_p, f32p], C.c_int),
            "forward": ([C.c_void_p, f32p, C.c_int64, f32p], C.c_int),
            "gradient": ([C.c_void_p, f32p, f32p, C.c_int64, f32p, f32p], C.c_int),
            "train": ([C.c_void_p, f32p, f32p, C.c_int64, C.c_int32, C.c_int32, C.c_float, C.c_uint64,
                       f32p], C.c_int),
        }
        for nm, (args, res) in sig.items():
            fn = getattr(lib, "pdem_mlp_" + nm)
            fn.argtypes = args
            fn.restype = res
        self.h = C.c_void_p()
        if lib.pdem_mlp_create(int(seed), C.byref(self.h)) != 0:
            raise RuntimeError("pdem_mlp_create: " + lib.pdem_mlp_last_error(None).decode())
        self.nparam = int(lib.pdem_mlp_param_count())

    def _check(self, rc, what):
        if rc != 0:
            raise RuntimeError(f"pdem_mlp_{what}: " + self.lib.pdem_mlp_last_error(self.h).decode())

    @staticmethod
    def _f32(a, shape=None):
        a = np.ascontiguousarray(a, dtype=np.float32)
        if shape is not None:
            a = a.reshape(shape)
        return a, a.ctypes.data_as(C.POINTER(C.c_float))

    def close(self):
        """Free the network; raises while an engine context still has it enabled."""
        if self.h:
            self._check(self.lib.pdem_mlp_destroy(self.h), "destroy")
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def params(self):
        out = np.empty(self.nparam, np.float32)
        self._check(self.lib.pdem_mlp_get_params(self.h, self._f32(out)[1]), "get_params")
        return out

    def set_params(self, p):
        a, ptr = self._f32(p)
        assert a.size == self.nparam
        self._check(self.lib.pdem_mlp_set_params(self.h, ptr), "set_params")

    def set_normalization(self, in_mean, in_scale, out_scale, out_weight=(1.0, 1.0, 1.0, 1.0)):
        bufs = [self._f32(v) for v in (in_mean, in_scale, out_scale, out_weight)]
        self._check(self.lib.pdem_mlp_set_normalization(self.h, *[b[1] for b in bufs]), "set_normalization")

    def normalization(self):
        out = np.empty(40, np.float32)
        self._check(self.lib.pdem_mlp_get_normalization(self.h, self._f32(out)[1]), "get_normalization")
        return out[:16], out[16:32], out[32:36], out[36:40]

    def forward(self, x):
        x, px = self._f32(x)
        n = x.shape[0]
        y = np.empty((n, self.NOUT), np.float32)
        self._check(self.lib.pdem_mlp_forward(self.h, px, n, self._f32(y)[1]), "forward")
        return y

    def gradient(self, x, t):
        x, px = self._f32(x)
        t, pt = self._f32(t)
        g = np.empty(self.nparam, np.float32)
        loss = np.empty(1, np.float32)
        self._check(self.lib.pdem_mlp_gradient(self.h, px, pt, x.shape[0], self._f32(g)[1],
                                               self._f32(loss)[1]), "gradient")
        return float(loss[0]), g

    def train(self, x, t, iters, batch=1024, lr=1e-3, seed=0):
        x, px = self._f32(x)
        t, pt = self._f32(t)
        curve = np.empty(iters, np.float32)
        self._check(self.lib.pdem_mlp_train(self.h, px, pt, x.shape[0], int(iters), int(batch), float(lr),
                                            int(seed), self._f32(curve)[1]), "train")
        return curve
