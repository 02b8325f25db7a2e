"""PeriDEM input decks: the reference's YAML deck and mesh files to a ModelSpec.

A reference deck (YAML, read as the reference's Input class does,
src/inp/input.cpp:189-1395) plus its mesh files (gmsh .msh 2.x ASCII or
.csv `id,x,y[,z],volume`) becomes the ModelSpec the engine takes, so decks
written for the reference run unmodified:

  Restart          File, Step (restart file resolved; applied by output.restart_from_vtu)
  Model            Dimension, Final_Time / Time_Steps -> dt, Particle_Sim_Type,
                   Discretization_Type.Time (central_difference / velocity_verlet)
  Zone             Zones, Zone_k.Is_Wall
  Particle         Zone_k geometry (Type, Parameters), All_Dofs_Constrained,
                   Test_Name, Compressive_Test.{Particle_Id, Particle_Force_Direction}
  Particle_Generation  From_File (loc_rad / loc_rad_orient CSV) or one particle
                   per zone from the zone's own mesh
  Mesh             Zone_k.File, Mesh_Size
  Material         Zone_k: Type (RNPBond / PMBBond / PDElasticBond / PDState), E K G Poisson_Ratio Gc KIc,
                   Density, Horizon / Horizon_Mesh_Ratio, Influence_Function,
                   Bond_Potential.{Check_Sc_Factor, Irreversible_Bond_Fracture},
                   Compute_From_Classical, Copy_Material_Data
  Contact          Zone_ij: Contact_Radius(_Factor), Kn, Kn_Factor, Epsilon,
                   Beta_n_Factor, Friction_Coeff, Friction_On, Damping_On,
                   Copy_Contact_Data
  Neighbor         Search_Factor, Search_Interval
  Force_BC         Gravity, Sets (particle / region selections)
  Displacement_BC  Sets (particle / region selections, Zero_Displacement)
  IC               Constant_Velocity.{Velocity_Vector, Particle_List}

Output, restart and post-processing keys are accepted and ignored.  Nodal
volumes of .msh meshes are the finite-element ones of fe::Mesh::computeVol
(mesh.cpp:250-313): the integral of each node's linear shape function over
its elements (area / 3 per triangle, volume / 4 per tetrahedron, the 2x2
Gauss integral for quadrangles).
"""
from __future__ import annotations

import csv
import math
import os
from typing import Dict, List, Optional, Tuple

import numpy as np
import yaml

from . import capi
from .geometry import center_node
from .material import MaterialDeck
from .model import BcDeck, ContactPairDeck, ModelBuilder, ModelSpec, ZoneDeck

COMPARE_EPS = 1e-5


def _is_less(a, b):
    return (a - b) < -max(abs(a), abs(b)) * COMPARE_EPS


def _is_greater(a, b):
    return (a - b) > max(abs(a), abs(b)) * COMPARE_EPS


# --------------------------------------------------------------- meshes
def read_msh(path: str, dim: int, with_cells: bool = False):
    """gmsh 2.x ASCII mesh (rw/mshReader.cpp:22-230): nodes and finite-element
    nodal volumes of the triangle / quadrangle (2D) or tetrahedron (3D) cells;
    with `with_cells`, also the cells (0-based node ids, file order)."""
    with open(path) as fh:
        tok = fh.read().split("\n")
    nodes = None
    cells: List[List[int]] = []
    ctype = None
    i = 0
    n = len(tok)
    while i < n:
        line = tok[i].strip()
        if line.startswith("$MeshFormat"):
            ver, fmt = tok[i + 1].split()[:2]
            if float(ver) not in (2.0, 2.1, 2.2) or int(fmt) != 0:
                raise ValueError(f"{path}: only ASCII .msh 2.x is supported")
            i += 2
        elif line.startswith("$Nodes") or line.startswith("$NOD") or line.startswith("$NOE"):
            cnt = int(tok[i + 1])
            nodes = np.zeros((cnt, 3))
            for k in range(cnt):
                f = tok[i + 2 + k].split()
                nodes[int(f[0]) - 1] = [float(f[1]), float(f[2]), float(f[3])]
            i += 2 + cnt
        elif line.startswith("$Elements") or line.startswith("$ELM"):
            cnt = int(tok[i + 1])
            for k in range(cnt):
                f = [int(v) for v in tok[i + 2 + k].split()]
                etype, ntags = f[1], f[2]
                conn = [v - 1 for v in f[3 + ntags:]]
                want = (dim == 2 and etype in (2, 3)) or (dim == 3 and etype == 4)
                if want:
                    if ctype is not None and ctype != etype:
                        raise ValueError(f"{path}: mixed triangle and quadrangle cells")
                    ctype = etype
                    cells.append(conn)
            i += 2 + cnt
        else:
            i += 1
    if nodes is None or not cells:
        raise ValueError(f"{path}: no nodes or no supported cells")
    enc = np.array(cells, dtype=np.int64)
    vol = fe_nodal_volumes(nodes, enc, dim)
    return (nodes, vol, enc) if with_cells else (nodes, vol)


_GP = 1.0 / math.sqrt(3.0)


def fe_nodal_volumes(x: np.ndarray, cells: np.ndarray, dim: int) -> np.ndarray:
    vol = np.zeros(len(x))
    if dim == 2 and cells.shape[1] == 3:
        a, b, c = x[cells[:, 0], :2], x[cells[:, 1], :2], x[cells[:, 2], :2]
        area = 0.5 * np.abs((b[:, 0] - a[:, 0]) * (c[:, 1] - a[:, 1]) - (c[:, 0] - a[:, 0]) * (b[:, 1] - a[:, 1]))
        for k in range(3):
            np.add.at(vol, cells[:, k], area / 3.0)
    elif dim == 2 and cells.shape[1] == 4:
        # bilinear quadrangle, 2x2 Gauss points: sum_q N_i(q) |det J(q)| w_q
        for xi, eta in ((-_GP, -_GP), (_GP, -_GP), (_GP, _GP), (-_GP, _GP)):
            N = 0.25 * np.array([(1 - xi) * (1 - eta), (1 + xi) * (1 - eta), (1 + xi) * (1 + eta), (1 - xi) * (1 + eta)])
            dNxi = 0.25 * np.array([-(1 - eta), (1 - eta), (1 + eta), -(1 + eta)])
            dNeta = 0.25 * np.array([-(1 - xi), -(1 + xi), (1 + xi), (1 - xi)])
            P = x[cells][:, :, :2]
            J00 = P[:, :, 0] @ dNxi
            J01 = P[:, :, 0] @ dNeta
            J10 = P[:, :, 1] @ dNxi
            J11 = P[:, :, 1] @ dNeta
            det = np.abs(J00 * J11 - J01 * J10)
            for k in range(4):
                np.add.at(vol, cells[:, k], N[k] * det)
    elif dim == 3 and cells.shape[1] == 4:
        a, b, c, d = (x[cells[:, k]] for k in range(4))
        v6 = np.abs(np.einsum("ij,ij->i", b - a, np.cross(c - a, d - a)))
        for k in range(4):
            np.add.at(vol, cells[:, k], v6 / 24.0)
    else:
        raise ValueError("unsupported cell type")
    return vol


def read_csv_mesh(path: str, dim: int) -> Tuple[np.ndarray, np.ndarray]:
    """`id,x,y[,z],volume` node file (rw/reader.cpp:16-56)."""
    xs, vs = [], []
    with open(path) as fh:
        rd = csv.DictReader(fh, skipinitialspace=True)
        for row in rd:
            p = [float(row["x"]), float(row.get("y", 0.0)) if dim > 1 else 0.0,
                 float(row.get("z", 0.0)) if dim > 2 else 0.0]
            xs.append(p)
            vs.append(float(row["volume"]))
    return np.array(xs), np.array(vs)


def read_mesh(path: str, dim: int):
    """Nodes, nodal volumes and cells (None for node files) of a mesh file."""
    ext = os.path.splitext(path)[1].lower()
    if ext == ".msh":
        return read_msh(path, dim, with_cells=True)
    if ext == ".csv":
        return read_csv_mesh(path, dim) + (None,)
    raise ValueError(f"{path}: only .msh and .csv meshes are read (.vtu needs VTK)")


# ------------------------------------------------------------- geometry
def _geom_center_radius(gtype: str, prm: List[float]) -> Tuple[np.ndarray, float]:
    """Centre and bounding radius of a zone's representative geometry
    (util::geometry objects; parameters as in the deck)."""
    p = list(prm) + [0.0] * 8
    if gtype in ("circle", "sphere", "hexagon", "triangle", "square", "cube"):
        r = p[0]
        if gtype == "square":
            r = p[0] * math.sqrt(0.5)
        elif gtype == "cube":
            r = p[0] * math.sqrt(0.75)
        return np.array(p[1:4]), r
    if gtype in ("rectangle", "cuboid"):
        lo, hi = np.array(p[0:3]), np.array(p[3:6])
        return 0.5 * (lo + hi), 0.5 * float(np.linalg.norm(hi - lo))
    if gtype == "drum2d":
        return np.array(p[2:5]), p[0]
    raise ValueError(f"geometry type {gtype} is not supported by the deck reader")


class _Region:
    """Region test of a BC set (isInside of util::geometry objects with the
    reference's 1e-12 slack and relative compares)."""

    def __init__(self, gtype: str, prm: List[float], dim: int):
        self.t, self.p, self.dim = gtype, list(prm) + [0.0] * 8, dim

    def inside(self, x) -> bool:
        p = self.p
        if self.t == "rectangle":
            return not (_is_less(x[0], p[0] - 1e-12) or _is_less(x[1], p[1] - 1e-12) or
                        _is_greater(x[0], p[3] + 1e-12) or _is_greater(x[1], p[4] + 1e-12))
        if self.t == "cuboid":
            return not (_is_less(x[0], p[0] - 1e-12) or _is_less(x[1], p[1] - 1e-12) or
                        _is_less(x[2], p[2] - 1e-12) or _is_greater(x[0], p[3] + 1e-12) or
                        _is_greater(x[1], p[4] + 1e-12) or _is_greater(x[2], p[5] + 1e-12))
        if self.t in ("circle", "sphere"):
            d = math.sqrt(sum((x[k] - p[1 + k]) ** 2 for k in range(3)))
            return _is_less(d, p[0] + 1e-12)
        raise ValueError(f"region geometry {self.t} is not supported")


# ------------------------------------------------------------ deck keys
_UFN = {"constant": capi.UFN_CONSTANT, "linear": capi.UFN_LINEAR, "quadratic": capi.UFN_QUADRATIC,
        "sin": capi.UFN_SIN, "rotation": capi.UFN_ROTATION}
_FFN = {"constant": capi.FFN_CONSTANT, "linear": capi.FFN_LINEAR, "linear_step": capi.FFN_LINEAR_STEP,
        "linear_slow_fast": capi.FFN_LINEAR_SLOW_FAST, "sin": capi.FFN_SIN}


def _material(e: dict) -> MaterialDeck:
    t = e.get("Type")
    if t not in ("RNPBond", "PMBBond", "PDElasticBond", "PDState"):
        raise ValueError(f"material type {t}: the step supports RNPBond, PMBBond, PDElasticBond and PDState "
                         "(baseParticle.cpp:156-166)")
    if "Density" not in e:
        raise ValueError("material density missing")
    m = MaterialDeck(type=t, density=float(e["Density"]))
    for key, attr in (("E", "E"), ("K", "K"), ("G", "G"), ("Poisson_Ratio", "nu"), ("Gc", "Gc"), ("KIc", "KIc"),
                      ("Horizon", "horizon"), ("Horizon_Mesh_Ratio", "horizon_mesh_ratio")):
        if key in e:
            setattr(m, attr, float(e[key]))
    if "Compute_From_Classical" in e:
        m.compute_from_classical = bool(e["Compute_From_Classical"])
    if "Is_Plane_Strain" in e:
        m.plane_strain = bool(e["Is_Plane_Strain"])
    inf = e.get("Influence_Function") or {}
    if "Type" in inf:
        m.influence_fn_type = int(inf["Type"])
    if "Parameters" in inf:
        m.influence_fn_params = [float(v) for v in inf["Parameters"]]
    bp = e.get("Bond_Potential") or {}
    if "Check_Sc_Factor" in bp:
        m.check_sc_factor = float(bp["Check_Sc_Factor"])
    if "Irreversible_Bond_Fracture" in bp:
        m.irreversible = bool(bp["Irreversible_Bond_Fracture"])
    if "Parameters" in bp:
        m.bond_potential_params = [float(v) for v in bp["Parameters"]]
    return m


def _contact(e: dict) -> ContactPairDeck:
    c = ContactPairDeck(contact_radius_factor=None, contact_radius=None)
    if "Contact_Radius_Factor" in e:
        c.contact_radius_factor = float(e["Contact_Radius_Factor"])
    elif "Contact_Radius" in e:
        c.contact_radius = float(e["Contact_Radius"])
    else:
        raise ValueError("contact pair needs Contact_Radius_Factor or Contact_Radius")
    if "Kn" not in e:
        raise ValueError("contact pair needs Kn")
    c.Kn = float(e["Kn"])
    c.Kn_factor = float(e.get("Kn_Factor", 1.0))
    c.epsilon = float(e.get("Epsilon", 1.0))
    c.beta_n_factor = float(e.get("Beta_n_Factor", 1.0))
    c.friction_coeff = float(e.get("Friction_Coeff", 0.0))
    c.friction_on = bool(e.get("Friction_On", True))
    c.damping_on = bool(e.get("Damping_On", True))
    return c


def _resolve(path: str, deck_dir: str) -> str:
    """Mesh / particle files: as given (relative to the working directory),
    relative to the deck's folder or its parent, or next to the deck."""
    for cand in (path, os.path.join(deck_dir, path), os.path.join(os.path.dirname(deck_dir), path),
                 os.path.join(deck_dir, os.path.basename(path))):
        if os.path.exists(cand):
            return cand
    raise FileNotFoundError(path)


def load_deck(path: str, name: Optional[str] = None) -> ModelSpec:
    """Build the ModelSpec of a reference input deck."""
    with open(path) as fh:
        cfg = yaml.safe_load(fh)
    deck_dir = os.path.dirname(os.path.abspath(path))
    model = cfg["Model"]
    dim = int(model["Dimension"])
    if dim not in (2, 3):
        raise ValueError("Dimension must be 2 or 3")
    dt = float(model["Final_Time"]) / float(model["Time_Steps"])
    multi = model.get("Particle_Sim_Type", "Single_Particle") == "Multi_Particle"
    nz = int((cfg.get("Zone") or {}).get("Zones", 1))
    pblock = cfg.get("Particle") or {}
    mblock = cfg.get("Material") or {}
    meshblock = cfg.get("Mesh") or {}

    zones: List[ZoneDeck] = []
    zone_geom = []
    zone_mesh = []
    for z in range(1, nz + 1):
        zk = f"Zone_{z}"
        zinfo = (cfg.get("Zone") or {}).get(zk) or {}
        pz = pblock.get(zk) or {}
        me = mblock.get(zk) if zk in mblock else mblock
        if me and "Copy_Material_Data" in me:
            me = mblock[f"Zone_{int(me['Copy_Material_Data'])}"]
        zones.append(ZoneDeck(_material(me), is_wall=bool(zinfo.get("Is_Wall", False)),
                              all_dofs_constrained=bool(pz.get("All_Dofs_Constrained", False))))
        gtype = pz.get("Type", "circle")
        zone_geom.append((gtype, [float(v) for v in pz.get("Parameters", [])]))
        mk = meshblock.get(zk) or meshblock
        if "File" not in mk:
            raise ValueError(f"{zk}: mesh file missing (mesh generation is not part of the reader)")
        x, vol, cells = read_mesh(_resolve(str(mk["File"]).strip(), deck_dir), dim)
        zone_mesh.append((x, vol, float(mk["Mesh_Size"]) if "Mesh_Size" in mk else None, cells))

    contact: Dict[Tuple[int, int], ContactPairDeck] = {}
    cblock = cfg.get("Contact") or {}
    for a in range(nz):
        for b in range(a, nz):
            key = f"Zone_{a + 1}{b + 1}"
            if key not in cblock:
                continue
            e = cblock[key]
            if "Copy_Contact_Data" in e:
                i, j = [int(v) for v in e["Copy_Contact_Data"]]
                e = cblock[f"Zone_{i}{j}"]
            contact[(a, b)] = _contact(e)

    nb = cfg.get("Neighbor") or {}
    if multi and not nb:
        raise ValueError("Neighbor block missing in a multi-particle deck")
    gravity = tuple(float(v) for v in ((cfg.get("Force_BC") or {}).get("Gravity") or [0.0, 0.0, 0.0]))
    b = ModelBuilder(dim, dt, zones, contact, search_factor=float(nb.get("Search_Factor", 1.0)),
                     neigh_update_interval=int(nb.get("Search_Interval", 1)), gravity=gravity,
                     multi_particle=multi)

    # particles, zone by zone (DEMModel::createParticles, demModel.cpp:1160-1264)
    gen = cfg.get("Particle_Generation") or {}
    h_over: List[Optional[float]] = []
    for z in range(nz):
        gtype, prm = zone_geom[z]
        center, R = _geom_center_radius(gtype, prm)
        x_ref, vol_ref, hdeck, cells = zone_mesh[z]
        cl = center_node(x_ref, center, R)
        from_file = "From_File" in gen and not zones[z].is_wall and multi
        pz = pblock.get(f"Zone_{z + 1}") or {}
        if pz.get("Create_Particle_Using_ParticleZone_GeomObject", False):
            from_file = False
        if from_file:
            sites, rads, orients = _particle_file(_resolve(str(gen["From_File"]).strip(), deck_dir),
                                                  str(gen.get("File_Data_Type", "loc_rad")), z,
                                                  pblock.get("Test_Name", ""))
            for site, rad, th in zip(sites, rads, orients):
                s = rad / R
                c, sn = math.cos(th), math.sin(th)
                rot = np.array([[c, -sn, 0.0], [sn, c, 0.0], [0.0, 0.0, 1.0]])
                x = np.asarray(site)[None, :] + s * (x_ref @ rot.T)
                b.add_particle(z, x, vol_ref * s ** dim, cl, rad, cells=cells)
                h_over.append(hdeck * s if hdeck else None)
        else:
            b.add_particle(z, x_ref.copy(), vol_ref.copy(), cl, R, cells=cells)
            h_over.append(hdeck)
    if pblock.get("Test_Name", "") == "two_particle" and b.parts:
        b.force_off.append(0)

    ic = (cfg.get("IC") or {}).get("Constant_Velocity")
    if ic:
        vel = [float(v) for v in ic["Velocity_Vector"]]
        b.ic.append(([int(p) for p in ic.get("Particle_List", [])], tuple(vel + [0.0] * (3 - len(vel)))))

    for tag, table, dest in (("Displacement_BC", _UFN, b.ubc), ("Force_BC", _FFN, b.fbc)):
        blk = cfg.get(tag) or {}
        for s in range(1, int(blk.get("Sets", 0)) + 1):
            e = blk[f"Set_{s}"]
            tf = e.get("Time_Function") or {}
            ttype = tf.get("Type", "constant")
            fn = table.get(ttype, capi.UFN_NONE if tag == "Displacement_BC" else None)
            if fn is None:
                raise ValueError(f"{tag} time function {ttype} is not supported")
            sf = e.get("Spatial_Function") or {}
            region = None
            if "Region" in e:
                g = e["Region"]["Geometry"]
                region = _Region(g["Type"], [float(v) for v in g["Parameters"]], dim)
            dest.append(BcDeck(particles=[int(p) for p in e.get("Particle_List", [])],
                               directions=[int(d) for d in e["Direction"]], time_fn=fn,
                               time_params=[float(v) for v in tf.get("Parameters", [0.0])],
                               spatial_fn=sf.get("Type", "constant"),
                               spatial_params=[float(v) for v in sf.get("Parameters", [])],
                               is_zero=bool(e.get("Zero_Displacement", False)) and tag == "Displacement_BC",
                               region=region,
                               exclude=[int(p) for p in e.get("Particle_Exclude_List", [])]))

    ct = pblock.get("Compressive_Test") or {}
    wall_id = int(ct.get("Particle_Id", -1)) if ct else -1
    f_dir = int(ct.get("Particle_Force_Direction", 2)) - 1 if ct else 1
    hover = None if all(h is None for h in h_over) else h_over
    spec = b.build(name=name or os.path.splitext(os.path.basename(path))[0], n_steps=int(model["Time_Steps"]),
                   wall_id=wall_id, f_dir=f_dir, h_override=hover)
    spec.info.update(deck=os.path.abspath(path), test_name=pblock.get("Test_Name", ""))
    # Model.Discretization_Type.Time (input.cpp:205-217); integrateStep
    # (demModel.cpp:342-347) steps nothing for any other name, so those are errors here
    scheme = (model.get("Discretization_Type") or {}).get("Time", "central_difference")
    if scheme not in ("central_difference", "velocity_verlet"):
        raise ValueError(f"time discretization {scheme}: central_difference or velocity_verlet")
    spec.time_scheme = scheme
    # strain / stress post-processing: quadrature order (modelDeck.h:107, input.cpp:222-223)
    spec.info["quad_order"] = int(model.get("Quad_Approximation_Order", 1))
    # restart (input.cpp:835-860): file and step, applied by output.restart_from_vtu
    rs = cfg.get("Restart")
    if rs is not None:
        if "File" not in rs or "Step" not in rs:
            raise ValueError("Restart needs File and Step")
        f = str(rs["File"]).strip()
        try:
            f = _resolve(f, deck_dir)
        except FileNotFoundError:
            f = os.path.join(deck_dir, f)  # written later by an earlier run
        spec.info["restart"] = {"file": f, "step": int(rs["Step"])}
    return spec


def _particle_file(path: str, kind: str, zone: int, test_name: str):
    """Sites, radii and orientations of one zone (rw/reader.cpp:58-123):
    rows whose first column equals the 0-based zone id."""
    sites, rads, orients = [], [], []
    with open(path) as fh:
        rd = csv.DictReader(fh, skipinitialspace=True)
        for row in rd:
            if int(float(row["i"])) != zone:
                continue
            sites.append([float(row["x"]), float(row["y"]), float(row.get("z", 0.0) or 0.0)])
            rads.append(float(row["r"]))
            if kind == "loc_rad_orient":
                orients.append(float(row["o"]))
    if kind == "loc_rad":
        if test_name != "two_particle":
            raise ValueError("loc_rad particle files draw random orientations from the reference's "
                             "seeded generator; use loc_rad_orient")
        orients = [k * math.pi for k in range(len(rads))]
    elif kind != "loc_rad_orient":
        raise ValueError(f"File_Data_Type {kind} is not supported")
    return sites, rads, orients
