"""Model setup: particles, zones, contact table and boundary conditions.

This is the host-side setup that precedes the time step (the reference's
DEMModel::init, demModel.cpp:109-266): it resolves every per-particle and
per-zone constant the step needs and hands them to the engine as one
descriptor (include/pdem.h).
  particle constants   BaseParticle ctor        src/particle/baseParticle.cpp:60-189
  contact table        DEMModel::setupContact   demModel.cpp:1531-1600
  fixity               ParticleULoading::setFixity  particleULoading.cpp:90-113
  BC node selection    needToProcessParticle / needToComputeDof  particleULoading.cpp:45-88
  force-off particles  two_particle test        demModel.cpp:218-220
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

from . import capi
from .geometry import mesh_size
from .material import InfluenceFn, MaterialDeck, elastic_constants, material_params, particle_contact_constants


@dataclass
class ZoneDeck:
    material: MaterialDeck
    is_wall: bool = False
    all_dofs_constrained: bool = False


@dataclass
class ContactPairDeck:
    """inp::ContactPairDeck (src/inp/pdecks/contactDeck.h:20-80) as read by
    Input::setContactDeck (input.cpp:665-833)."""
    contact_radius_factor: Optional[float] = 0.95
    contact_radius: Optional[float] = None
    Kn: float = 0.0
    Kn_factor: float = 1.0
    epsilon: float = 1.0
    beta_n_factor: float = 1.0
    friction_coeff: float = 0.0
    damping_on: bool = True
    friction_on: bool = True


@dataclass
class BcDeck:
    """One Displacement_BC / Force_BC set (inp::PBCData, input.cpp:487-597):
    a particle list, a region, or a region with include / exclude lists."""
    particles: Sequence[int]
    directions: Sequence[int]        # 1-based, as in the deck
    time_fn: int
    time_params: Sequence[float]
    spatial_fn: str = "constant"
    spatial_params: Sequence[float] = ()
    is_zero: bool = False
    region: Optional[object] = None  # object with inside(x_ref) -> bool
    exclude: Sequence[int] = ()

    def selection(self) -> str:
        if self.region is None:
            return "particle" if len(self.particles) else ""
        s = "region"
        if len(self.particles):
            s += "_with_include_list"
        if len(self.exclude):
            s += "_with_exclude_list"
        return s

    def process_particle(self, pid: int) -> bool:
        """ParticleULoading::needToProcessParticle (particleULoading.cpp:40-57)."""
        sel = self.selection()
        if sel in ("particle", "region_with_include_list") and pid not in self.particles:
            return False
        if sel == "region_with_exclude_list" and pid in self.exclude:
            return False
        if sel == "region_with_include_list_with_exclude_list" and (pid in self.exclude or
                                                                    pid not in self.particles):
            return False
        return True

    def compute_dof(self, x, pid: int) -> bool:
        """ParticleULoading::needToComputeDof (particleULoading.cpp:59-88)."""
        sel = self.selection()
        if self.region is None:
            return sel == "particle" and pid in self.particles
        if not self.region.inside(x):
            return False
        if sel == "region":
            return True
        if sel == "region_with_include_list":
            return pid in self.particles
        if sel == "region_with_exclude_list":
            return pid not in self.exclude
        return pid in self.particles and pid not in self.exclude


@dataclass
class ModelSpec:
    dim: int
    dt: float
    n_nodes: int
    x_ref: np.ndarray
    u: np.ndarray
    v: np.ndarray
    vol: np.ndarray
    fix: np.ndarray
    pt_id: np.ndarray
    particles: List[dict]
    n_zones: int
    contact: List[List[dict]]
    gravity: Tuple[float, float, float]
    search_factor: float
    neigh_update_interval: int
    ubc: List[dict]
    fbc: List[dict]
    multi_particle: bool
    influence_fn_type: int
    influence_fn_params: Tuple[float, float]
    name: str = ""
    n_steps: int = 0
    wall_id: int = -1          # compressive test wall (Particle_Id)
    f_dir: int = 1             # compressive test direction, 0-based
    h_min: float = 0.0
    max_contact_r: float = 0.0
    info: dict = field(default_factory=dict)
    time_scheme: str = "central_difference"   # or "velocity_verlet"

    @property
    def n_particles(self):
        return len(self.particles)

    def compute_nodes(self):
        """Number of nodes integrated per step (|d_fPdCompNodes|)."""
        cf = np.array([p["compute_force"] for p in self.particles], bool)
        return int(cf[self.pt_id].sum())


def _is_greater(a, b):
    return (a - b) > max(abs(a), abs(b)) * 1e-5


class ModelBuilder:
    def __init__(self, dim: int, dt: float, zones: List[ZoneDeck],
                 contact: Dict[Tuple[int, int], ContactPairDeck],
                 search_factor: float = 1.0, neigh_update_interval: int = 1,
                 gravity=(0.0, 0.0, 0.0), multi_particle: bool = True):
        self.dim = dim
        self.dt = dt
        self.zones = zones
        self.contact = contact
        self.search_factor = search_factor
        self.interval = neigh_update_interval
        self.gravity = tuple(gravity)
        self.multi = multi_particle
        self.parts = []          # (zone, x (n,3), vol (n,), center_local, bounding_radius)
        self.ubc: List[BcDeck] = []
        self.fbc: List[BcDeck] = []
        self.ic = []             # (particle ids, velocity)
        self.force_off: List[int] = []
        self.cells: List[Optional[np.ndarray]] = []

    def add_particle(self, zone: int, x: np.ndarray, vol: np.ndarray, center_local: int,
                     bounding_radius: float, cells: Optional[np.ndarray] = None) -> int:
        """cells: the particle mesh's element-node connectivity (local node
        ids), kept for the strain / stress post-processing."""
        self.parts.append((zone, np.asarray(x, float), np.asarray(vol, float), int(center_local),
                           float(bounding_radius)))
        self.cells.append(None if cells is None else np.asarray(cells, np.int64))
        return len(self.parts) - 1

    def build(self, name="", n_steps=0, wall_id=-1, f_dir=1, h_override=None) -> ModelSpec:
        dim = self.dim
        counts = [len(p[1]) for p in self.parts]
        starts = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
        N = int(starts[-1])
        x_ref = np.concatenate([p[1] for p in self.parts]) if self.parts else np.zeros((0, 3))
        if dim == 2:
            x_ref[:, 2] = 0.0
        vol = np.concatenate([p[2] for p in self.parts])
        pt_id = np.repeat(np.arange(len(self.parts), dtype=np.int64), counts)
        # the influence function is process-global: the first material built wins
        first = self.zones[self.parts[0][0]].material
        inf = InfluenceFn(first.influence_fn_type, first.influence_fn_params, dim)
        particles = []
        for pid, (z, x, vv, cl, br) in enumerate(self.parts):
            zd = self.zones[z]
            md = zd.material
            h = h_override[pid] if h_override is not None and h_override[pid] is not None else mesh_size(x)
            horizon = md.horizon
            if md.horizon_mesh_ratio > 0.0:
                horizon = md.horizon_mesh_ratio * h
            mp = material_params(md, dim, horizon, inf)
            Kn, Rc = particle_contact_constants(mp.K_for_contact, horizon, h)
            compute_force = not zd.all_dofs_constrained and pid not in self.force_off
            particles.append(dict(
                glob_start=int(starts[pid]), glob_end=int(starts[pid + 1]),
                center_node=int(starts[pid] + cl), zone=int(z), type_index=1 if zd.is_wall else 0,
                all_dofs_constrained=int(zd.all_dofs_constrained), compute_force=int(compute_force),
                material_type=mp.material_type, irreversible=int(mp.irreversible),
                density=float(md.density), horizon=float(horizon), mesh_size=float(h),
                bounding_radius=br, Kn=float(Kn), Rc=float(Rc), mat=list(mp.mat),
                elastic=list(elastic_constants(mp, dim, horizon, inf)), plane_strain=int(md.plane_strain),
                fe_cells=self.cells[pid] if pid < len(self.cells) else None))
        # DEMModel::setupContact (demModel.cpp:1531-1600)
        nz = len(self.zones)
        h_min = h_max = 0.0
        for c, p in enumerate(particles):
            h = p["mesh_size"]
            if c == 0:
                h_min = h_max = h
            if _is_greater(h_min, h):
                h_min = h
            if _is_greater(h, h_max):
                h_max = h
        contact = [[None] * nz for _ in range(nz)]
        max_r = 0.0
        for a in range(nz):
            for b in range(nz):
                cd = self.contact.get((min(a, b), max(a, b)), ContactPairDeck())
                if cd.contact_radius is not None and cd.contact_radius_factor is None:
                    r, compute_r = cd.contact_radius, False
                else:
                    r, compute_r = cd.contact_radius_factor or 0.0, True
                if compute_r:
                    r *= h_min
                max_r = max(max_r, r)
                Ka = self.zones[a].material.K
                Kb = self.zones[b].material.K
                if self.multi and (Ka < 0 or Kb < 0):
                    raise ValueError("contact needs the bulk modulus K of every zone in the deck")
                kappa = 2.0 * Ka * Kb / (Ka + Kb) if self.multi else 1.0
                Kn = cd.Kn * cd.Kn_factor
                eps = cd.epsilon if cd.damping_on else 1.0
                bfac = cd.beta_n_factor if cd.damping_on else 0.0
                log_e = math.log(eps)
                betan = bfac * (-2.0 * log_e * math.sqrt(1.0 / (math.pi * math.pi + log_e * log_e)))
                mu = cd.friction_coeff if cd.friction_on else 0.0
                contact[a][b] = dict(contact_r=r, Kn=Kn, mu=mu, betan=betan, kappa=kappa)
        # initial velocity (DEMModel::applyInitialCondition, demModel.cpp:1068-1097)
        v = np.zeros((N, 3))
        for plist, vel in self.ic:
            for pid in plist:
                v[starts[pid]:starts[pid + 1], :] = vel
        # fixity and BC node lists
        fix = np.zeros(N, np.uint8)
        ubc, fbc = [], []
        for bc in self.ubc:
            mask = 0
            for d in bc.directions:
                mask |= 1 << (d - 1)
            nodes, amp = self._bc_nodes(bc, starts, x_ref, is_force=False)
            # setFixity marks every selected node, zero sets included
            fix[nodes] |= np.uint8(mask)
            if bc.is_zero:
                # ParticleULoading::apply marks a zero set applied on its first call,
                # which is for particle 0; only particle 0's nodes are ever written
                keep = nodes < starts[1]
                nodes, amp = nodes[keep], amp[keep]
            ubc.append(dict(time_fn=bc.time_fn, dir_mask=mask, is_zero=int(bc.is_zero),
                            time_params=list(bc.time_params), nodes=nodes, amp=amp))
        for bc in self.fbc:
            mask = 0
            for d in bc.directions:
                mask |= 1 << (d - 1)
            nodes, amp = self._bc_nodes(bc, starts, x_ref, is_force=True)
            fbc.append(dict(time_fn=bc.time_fn, dir_mask=mask, is_zero=0,
                            time_params=list(bc.time_params), nodes=nodes, amp=amp))
        spec = ModelSpec(dim=dim, dt=self.dt, n_nodes=N, x_ref=x_ref, u=np.zeros((N, 3)), v=v,
                         vol=vol, fix=fix, pt_id=pt_id, particles=particles, n_zones=nz,
                         contact=contact, gravity=self.gravity, search_factor=self.search_factor,
                         neigh_update_interval=self.interval, ubc=ubc, fbc=fbc,
                         multi_particle=self.multi, influence_fn_type=inf.kind,
                         influence_fn_params=tuple(inf.p), name=name, n_steps=n_steps,
                         wall_id=wall_id, f_dir=f_dir, h_min=h_min, max_contact_r=max_r)
        return spec

    def _bc_nodes(self, bc: BcDeck, starts, x_ref, is_force):
        nodes, amp = [], []
        for pid in range(len(self.parts)):
            if not bc.process_particle(pid):
                continue
            idx = np.arange(starts[pid], starts[pid + 1], dtype=np.int64)
            if bc.region is not None or pid not in bc.particles:
                keep = np.array([bc.compute_dof(x_ref[i], pid) for i in idx], bool)
                idx = idx[keep]
            if len(idx) == 0:
                continue
            x = x_ref[idx]
            base = 1.0 if is_force else float(bc.time_params[0])
            a = np.full(len(idx), base)
            sp, prm = bc.spatial_fn, list(bc.spatial_params)
            if sp in ("sin_x", "sin_y"):
                c = x[:, 0] if sp == "sin_x" else x[:, 1]
                s = np.array([math.sin(math.pi * prm[0] * t) for t in c])
                a = (prm[0] if is_force else base) * s
            elif sp in ("linear_x", "linear_y"):
                c = x[:, 0] if sp == "linear_x" else x[:, 1]
                a = ((prm[0] if is_force else base) * prm[0]) * c
            elif sp != "constant":
                raise ValueError(f"spatial function {sp} is not supported by the synthetic decks")
            nodes.append(idx)
            amp.append(a)
        if not nodes:
            return np.zeros(0, np.int64), np.zeros(0)
        return np.concatenate(nodes), np.concatenate(amp)
