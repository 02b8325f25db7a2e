"""The synthetic workloads: seeded generators for the five configurations.

Every generator returns a ModelSpec; sizes are parameters so the tests can
build reduced versions of the same decks.  Deck deviations from the
configuration text (contact stiffness factor, box size) are stated in each
generator's docstring and applied identically to the oracle and the GPU.
"""
from __future__ import annotations

import math

import numpy as np

from . import capi
from .geometry import ball_mesh, box_mesh, center_node, disk_mesh, hexagon_mesh, transform
from .material import MaterialDeck
from .model import BcDeck, ContactPairDeck, ModelBuilder, ZoneDeck


def rnp_material(E=2.0e9, nu=0.25, rho=1200.0, Gc=500.0, ratio=3.0, inf=capi.INF_CONST):
    K = E / (3.0 * (1.0 - 2.0 * nu))
    return MaterialDeck(type="RNPBond", E=E, K=K, Gc=Gc, density=rho, horizon_mesh_ratio=ratio,
                        influence_fn_type=inf)


def pdstate_material(K=2.16e7, G=1.296e7, rho=1200.0, Gc=50.0, ratio=3.0, inf=capi.INF_LINEAR):
    return MaterialDeck(type="PDState", K=K, G=G, Gc=Gc, density=rho, horizon_mesh_ratio=ratio,
                        influence_fn_type=inf)


def pmb_material(E=2.0e9, rho=1200.0, Gc=500.0, ratio=3.0, inf=capi.INF_CONST):
    """Bond-based PMB (nu fixed at 1/4 by the model)."""
    return MaterialDeck(type="PMBBond", E=E, K=E / 1.5, Gc=Gc, density=rho, horizon_mesh_ratio=ratio,
                        influence_fn_type=inf)


def pdelastic_material(E=2.0e9, rho=1200.0, ratio=3.0, inf=capi.INF_CONST):
    """Bond-based linear elastic PD (no fracture; nu fixed at 1/4)."""
    return MaterialDeck(type="PDElasticBond", E=E, K=E / 1.5, density=rho, horizon_mesh_ratio=ratio,
                        influence_fn_type=inf)


def _kn(mat: MaterialDeck, horizon: float) -> float:
    """Deck Kn = 18 K / (pi delta^5), as the reference's deck generators write it
    (test/test_data/peridem/twop_circ/inp/problem_setup.py:195-197)."""
    return 18.0 * mat.K / (math.pi * horizon ** 5)


def two_particle(R=1.0e-3, h=1.0e-4, gap=1.0e-4, speed=0.1, dt=1.0e-7, n_steps=20000,
                 kn_factor=1.0e-3, precision_note=""):
    """Config 0: two R = 1 mm circles, h = 0.1 mm, RNP, particle 2 at 0.1 m/s
    toward particle 1 across a 0.1 mm gap; no damping, no friction.

    Deviation: Kn_Factor = 1e-3 on the deck Kn = 18K/(pi delta^5).  With the
    full value the explicit step is unstable at dt = 1e-7 (per-pair contact
    stiffness Kn V / rho gives dt_crit ~ 1e-8 s)."""
    mat = rnp_material()
    zones = [ZoneDeck(mat)]
    horizon = 3.0 * h
    cp = ContactPairDeck(contact_radius_factor=0.95, Kn=_kn(mat, horizon), Kn_factor=kn_factor,
                         damping_on=False, friction_on=False)
    b = ModelBuilder(2, dt, zones, {(0, 0): cp}, search_factor=5.0, neigh_update_interval=40)
    ref, vol = disk_mesh(R, h)
    cl = center_node(ref, (0, 0, 0), R)
    b.add_particle(0, transform(ref, (0.0, 0.0, 0.0), 0.0, 1.0), vol, cl, R)
    b.add_particle(0, transform(ref, (0.0, 2 * R + gap, 0.0), math.pi, 1.0), vol, cl, R)
    b.ic.append(([1], (0.0, -speed, 0.0)))
    return b.build(name="two_particle_rnp_2d", n_steps=n_steps)


def compressive_2d(n_particles=500, R_lo=0.8e-3, R_hi=1.2e-3, nodes_per_particle=400,
                   box=None, seed=0, dt=5.0e-8, n_steps=50000, wall_speed=0.1,
                   kn_factor=1.0e-3, epsilon=0.7, mu=0.5, interval=40, pitch_factor=1.08,
                   init_speed=0.0, friction=False):
    """Config 1: 2D compressive test, 50/50 circles and hexagons, R ~ U[0.8, 1.2] mm,
    random orientation, seeded jittered grid; ~400 nodes per particle; RNP
    (E = 2 GPa, nu = 0.25, rho = 1200, Gc = 500), horizon 3h; contact with
    restitution 0.7 (Beta_n_Factor 1) and friction 0.5; the top wall moves down
    at 0.1 m/s (linear Displacement_BC); compressive_test observables on it.

    Deviations: the box is enlarged from 40 x 40 mm to the grid the particles
    need without overlap (the stated box would be 90% full); Kn_Factor = 1e-3
    for explicit stability at dt = 5e-8 (see two_particle); friction off
    (Friction_Coeff 0.5 with Friction_On false, as the reference's own
    compressive deck examples/PeriDEM/compressive_test/n500_circ_hex sets it):
    the reference's friction term pushes node i against the relative sliding
    velocity (demModel.cpp:896-897), which adds energy to every sliding
    contact, and the wall-driven packing then blows up within tens of
    microseconds of contact at any time step.  friction=True turns it on."""
    rng = np.random.default_rng(seed)
    mat = rnp_material()
    wall_mat = rnp_material()
    zones = [ZoneDeck(mat), ZoneDeck(mat), ZoneDeck(wall_mat, is_wall=True, all_dofs_constrained=True),
             ZoneDeck(wall_mat, is_wall=True)]
    # reference meshes at R = 1 mm with ~nodes_per_particle nodes
    h_ref = math.sqrt(math.pi / (nodes_per_particle * math.sqrt(3) / 2)) * 1.0e-3
    circ, cvol = disk_mesh(1.0e-3, h_ref)
    hexa, hvol = hexagon_mesh(1.0e-3, h_ref)
    c_cl = center_node(circ, (0, 0, 0), 1.0e-3)
    h_cl = center_node(hexa, (0, 0, 0), 1.0e-3)
    pitch = 2.0 * R_hi * pitch_factor
    ncol = int(math.ceil(math.sqrt(n_particles)))
    nrow = int(math.ceil(n_particles / ncol))
    W = ncol * pitch
    H = nrow * pitch
    if box is not None:
        W, H = max(W, box[0]), max(H, box[1])
    h_wall = 0.8 * h_ref * R_lo / 1.0e-3
    t_wall = 3.5 * h_wall
    delta_ref = 3.0 * h_ref
    cp = ContactPairDeck(contact_radius_factor=0.95, Kn=_kn(mat, delta_ref * R_lo / 1e-3),
                         Kn_factor=kn_factor, epsilon=epsilon, beta_n_factor=1.0,
                         friction_coeff=mu, damping_on=True, friction_on=friction)
    contact = {(a, b): cp for a in range(4) for b in range(a, 4)}
    b = ModelBuilder(2, dt, zones, contact, search_factor=5.0, neigh_update_interval=interval,
                     gravity=(0.0, -10.0, 0.0))
    k = 0
    for r in range(nrow):
        for c in range(ncol):
            if k >= n_particles:
                break
            rad = rng.uniform(R_lo, R_hi)
            jit = (pitch / 2 - rad) * 0.9
            cx = (c + 0.5) * pitch + rng.uniform(-jit, jit)
            cy = (r + 0.5) * pitch + rng.uniform(-jit, jit)
            th = rng.uniform(0, 2 * math.pi)
            s = rad / 1.0e-3
            if k % 2 == 0:
                b.add_particle(0, transform(circ, (cx, cy, 0.0), th, s), cvol * s * s, c_cl, rad)
            else:
                b.add_particle(1, transform(hexa, (cx, cy, 0.0), th, s), hvol * s * s, h_cl, rad)
            if init_speed > 0.0:
                a = rng.uniform(0, 2 * math.pi)
                b.ic.append(([k], (init_speed * math.cos(a), init_speed * math.sin(a), 0.0)))
            k += 1
    gap = 0.5 * h_wall
    # fixed container: bottom, left and right strips (one wall particle)
    parts = []
    for lo, hi in (((-t_wall, -t_wall - gap), (W + t_wall, -gap)),
                   ((-t_wall - gap, 0.0), (-gap, H + 2 * pitch)),
                   ((W + gap, 0.0), (W + gap + t_wall, H + 2 * pitch))):
        p, vv = box_mesh(lo, hi, h_wall, 2)
        parts.append((p, vv))
    wx = np.concatenate([p for p, _ in parts])
    wv = np.concatenate([v for _, v in parts])
    fixed_id = b.add_particle(2, wx, wv, 0, 0.5 * math.hypot(W, H))
    top, tv = box_mesh((0.0, H + gap), (W, H + gap + t_wall), h_wall, 2)
    top_id = b.add_particle(3, top, tv, 0, 0.5 * W)
    b.ubc.append(BcDeck(particles=[fixed_id], directions=[1, 2], time_fn=capi.UFN_CONSTANT,
                        time_params=[0.0], is_zero=True))
    b.ubc.append(BcDeck(particles=[top_id], directions=[2], time_fn=capi.UFN_LINEAR,
                        time_params=[-wall_speed]))
    spec = b.build(name=f"compressive_2d_n{n_particles}", n_steps=n_steps, wall_id=top_id, f_dir=1)
    spec.info.update(box=(W, H), seed=seed, shape_of_zone=[0.0, 1.0, 0.0, 0.0], wall_speed=wall_speed,
                     args=dict(n_particles=n_particles, R_lo=R_lo, R_hi=R_hi,
                               nodes_per_particle=nodes_per_particle, box=box, seed=seed, dt=dt,
                               n_steps=n_steps, wall_speed=wall_speed, kn_factor=kn_factor,
                               epsilon=epsilon, mu=mu, interval=interval, pitch_factor=pitch_factor,
                               init_speed=init_speed, friction=friction))
    return spec


def packing_3d(n_spheres=2000, R_lo=0.9e-3, R_hi=1.1e-3, nodes_per_sphere=1000, L=30.0e-3,
               seed=0, dt=2.0e-8, n_steps=10000, wall_speed=0.05, kn_factor=1.0e-2,
               max_tries=200000, gap_factor=1.04, init_speed=0.0, placement="auto"):
    """Config 2: 3D packing, state-based PD: spheres R ~ U[0.9, 1.1] mm placed by
    seeded sequential random insertion in an L^3 cube, ~1000 nodes per sphere,
    horizon 3h, gravity, a fixed bottom wall and a top wall moving down at
    0.05 m/s.  Material (not given by the configuration): the reference's own
    PDState deck values K = 2.16e7, G = 1.296e7, Gc = 50, rho = 1200, linear
    influence function; dt chosen for stability with Kn_Factor 1e-2.

    Deviation: sequential random insertion stalls near 1800 spheres at this
    volume fraction (0.31); for n_spheres >= 1000 the centres are a random
    subset of the sites of a seeded jittered body-centred cubic lattice
    (placement "lattice"), still without overlap."""
    rng = np.random.default_rng(seed)
    mat = pdstate_material()
    zones = [ZoneDeck(mat), ZoneDeck(mat, is_wall=True, all_dofs_constrained=True),
             ZoneDeck(mat, is_wall=True)]
    h_ref = 1.0e-3 * (4.0 / 3.0 * math.pi / nodes_per_sphere) ** (1.0 / 3.0)
    ball, bvol = ball_mesh(1.0e-3, h_ref)
    b_cl = center_node(ball, (0, 0, 0), 1.0e-3)
    delta = 3.0 * h_ref * R_lo / 1e-3
    cp = ContactPairDeck(contact_radius_factor=0.95, Kn=_kn(mat, delta), Kn_factor=kn_factor,
                         epsilon=0.7, beta_n_factor=1.0, friction_coeff=0.5)
    contact = {(a, b): cp for a in range(3) for b in range(a, 3)}
    b = ModelBuilder(3, dt, zones, contact, search_factor=5.0, neigh_update_interval=40,
                     gravity=(0.0, 0.0, -10.0))
    centers, radii = [], []
    margin = R_hi * 1.02
    tries = 0
    if placement == "lattice" or (placement == "auto" and n_spheres >= 1000):
        # body-centred cubic sites: nearest-neighbour distance a sqrt(3) / 2
        m = 1
        while 2 * m ** 3 < n_spheres:
            m += 1
        a = (L - 2 * margin) / m
        nn = a * math.sqrt(3.0) / 2.0
        if nn < 2 * R_hi * gap_factor:
            raise RuntimeError(f"{n_spheres} spheres do not fit a body-centred lattice in L = {L}")
        g = np.stack(np.meshgrid(*([np.arange(m)] * 3), indexing="ij"), -1).reshape(-1, 3)
        sites = np.concatenate([g + 0.25, g + 0.75]) * a + margin
        pick = np.sort(rng.permutation(len(sites))[:n_spheres])
        for sidx in pick:
            r = rng.uniform(R_lo, R_hi)
            jit = max(0.0, (nn - 2 * R_hi * gap_factor) / 2.0 * 0.9 / math.sqrt(3.0))
            centers.append(sites[sidx] + rng.uniform(-jit, jit, size=3))
            radii.append(r)
        max_tries = 0
    grid = {}
    cell = 2.0 * R_hi * 1.05
    while len(centers) < n_spheres and tries < max_tries:
        tries += 1
        r = rng.uniform(R_lo, R_hi)
        c = rng.uniform(margin, L - margin, size=3)
        key = tuple((c // cell).astype(int))
        ok = True
        for dx in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dz in (-1, 0, 1):
                    for j in grid.get((key[0] + dx, key[1] + dy, key[2] + dz), ()):
                        if np.linalg.norm(c - centers[j]) < (r + radii[j]) * gap_factor:
                            ok = False
                            break
                    if not ok:
                        break
                if not ok:
                    break
            if not ok:
                break
        if ok:
            grid.setdefault(key, []).append(len(centers))
            centers.append(c)
            radii.append(r)
    if len(centers) < n_spheres:
        raise RuntimeError(f"could only place {len(centers)} spheres")
    for c, r in zip(centers, radii):
        s = r / 1.0e-3
        th = rng.uniform(0, 2 * math.pi)
        pid = b.add_particle(0, transform(ball, c, th, s), bvol * s ** 3, b_cl, r)
        if init_speed > 0.0:
            d = rng.normal(size=3)
            d *= init_speed / np.linalg.norm(d)
            b.ic.append(([pid], tuple(d)))
    h_wall = 2.0 * h_ref
    t_wall = 1.5 * h_wall
    bot, bv = box_mesh((0.0, 0.0, -t_wall - 0.5 * h_wall), (L, L, -0.5 * h_wall), h_wall, 3)
    top, tv = box_mesh((0.0, 0.0, L + 0.5 * h_wall), (L, L, L + 0.5 * h_wall + t_wall), h_wall, 3)
    bid = b.add_particle(1, bot, bv, 0, L)
    tid = b.add_particle(2, top, tv, 0, L)
    b.ubc.append(BcDeck(particles=[bid], directions=[1, 2, 3], time_fn=capi.UFN_CONSTANT,
                        time_params=[0.0], is_zero=True))
    b.ubc.append(BcDeck(particles=[tid], directions=[3], time_fn=capi.UFN_LINEAR,
                        time_params=[-wall_speed]))
    spec = b.build(name=f"packing_3d_n{n_spheres}", n_steps=n_steps, wall_id=tid, f_dir=2)
    spec.info.update(seed=seed)
    return spec


def single_particle_3d(R=1.0e-3, nodes=1000, state=True, stretch=0.0):
    """One sphere, no contact: used by the closed-form tests."""
    mat = pdstate_material() if state else rnp_material()
    zones = [ZoneDeck(mat)]
    h = 1.0e-3 * (4.0 / 3.0 * math.pi / nodes) ** (1.0 / 3.0)
    ball, bvol = ball_mesh(R, h)
    cl = center_node(ball, (0, 0, 0), R)
    b = ModelBuilder(3, 1e-8, zones, {}, multi_particle=False)
    b.add_particle(0, ball, bvol, cl, R)
    return b.build(name="single_3d")


def ensemble_2d(n_instances=512, seed=0, dt=5.0e-8, n_steps=5000, R_lo=0.8e-3, R_hi=1.2e-3,
                n_min=2, n_max=64, ratios=(6, 8, 10), speed_lo=0.05, speed_hi=1.0,
                angle_max_deg=60.0, kn_factor=1.0e-2, pitch_factor=1.06, mu=0.5):
    """Config 3: a batch of independent 2D simulations in one model.

    Instance i (seed + i): particle count ~ U[n_min, n_max], mesh ratio
    h = R / r with r drawn from `ratios`, circles and hexagons alternating,
    R ~ U[R_lo, R_hi], seeded jittered grid, every particle moving toward the
    instance centre at speed ~ U[speed_lo, speed_hi] along a direction turned
    by an angle ~ U[0, angle_max] (impact angle).  State-based material (the
    reference's PDState deck values), mixed precision as config 2.

    Particles sit on a jittered grid whose gaps are at least 1.2 contact
    radii (no contact force at t = 0); pitch_factor widens the grid.

    Instances occupy disjoint regions of one model, separated by far more
    than the contact search radius plus the distance travelled, so no
    contact, damping or candidate pair ever joins two instances: the batch is
    the union of the independent simulations, advanced in one launch per
    kernel.  Zones: one per (shape, mesh ratio); a zone pair's contact radius
    is 0.95 times the finest mesh size of that ratio class, the value the
    reference would compute for the instance alone."""
    import numpy as _np

    mat = pdstate_material()
    zones = []
    meshes = []
    for r in ratios:
        h_ref = 1.0e-3 / r
        for shape in ("circle", "hexagon"):
            zones.append(ZoneDeck(mat))
            x, v = disk_mesh(1.0e-3, h_ref) if shape == "circle" else hexagon_mesh(1.0e-3, h_ref)
            meshes.append((x, v, center_node(x, (0, 0, 0), 1.0e-3)))
    nz = len(zones)
    contact = {}
    for a in range(nz):
        for b in range(a, nz):
            ra, rb = ratios[a // 2], ratios[b // 2]
            h = 1.0e-3 / max(ra, rb) * R_lo / 1.0e-3
            delta = 3.0 * h
            contact[(a, b)] = ContactPairDeck(contact_radius_factor=None, contact_radius=0.95 * h,
                                              Kn=_kn(mat, delta), Kn_factor=kn_factor, epsilon=0.7,
                                              beta_n_factor=1.0, friction_coeff=mu)
    b = ModelBuilder(2, dt, zones, contact, search_factor=5.0, neigh_update_interval=40)
    # start every pair out of contact: gaps of at least 1.2 contact radii
    rc_max = max(c.contact_radius for c in contact.values())
    pitch = 2.0 * R_hi * pitch_factor + 1.2 * rc_max
    box = int(math.ceil(math.sqrt(n_max))) * pitch
    spacing = box + 10.0e-3
    ncol_inst = int(math.ceil(math.sqrt(n_instances)))
    instances = []
    if n_instances < 1:
        raise ValueError("an ensemble needs at least one instance")
    for k in range(n_instances):
        rng = _np.random.default_rng(seed + k)
        n_p = int(rng.integers(n_min, n_max + 1))
        ri = int(rng.integers(0, len(ratios)))
        ox = (k % ncol_inst) * spacing
        oy = (k // ncol_inst) * spacing
        ncol = int(math.ceil(math.sqrt(n_p)))
        cxy = _np.array([ox + 0.5 * ncol * pitch, oy + 0.5 * int(math.ceil(n_p / ncol)) * pitch])
        first = len(b.parts)
        for q in range(n_p):
            rad = rng.uniform(R_lo, R_hi)
            jit = max(0.0, (pitch / 2 - rad - 0.6 * rc_max) * 0.9)
            cx = ox + (q % ncol + 0.5) * pitch + rng.uniform(-jit, jit)
            cy = oy + (q // ncol + 0.5) * pitch + rng.uniform(-jit, jit)
            z = 2 * ri + (q % 2)
            x, vol, cl = meshes[z]
            s = rad / 1.0e-3
            pid = b.add_particle(z, transform(x, (cx, cy, 0.0), rng.uniform(0, 2 * math.pi), s),
                                 vol * s * s, cl, rad)
            d = cxy - _np.array([cx, cy])
            nd = _np.linalg.norm(d)
            d = d / nd if nd > 0 else _np.array([1.0, 0.0])
            ang = math.radians(rng.uniform(0.0, angle_max_deg)) * (1 if rng.random() < 0.5 else -1)
            c, sn = math.cos(ang), math.sin(ang)
            sp = rng.uniform(speed_lo, speed_hi)
            vx, vy = sp * (c * d[0] - sn * d[1]), sp * (sn * d[0] + c * d[1])
            b.ic.append(([pid], (vx, vy, 0.0)))
        instances.append((first, len(b.parts), ratios[ri]))
    spec = b.build(name=f"ensemble_2d_x{n_instances}", n_steps=n_steps)
    spec.info.update(instances=instances, seed=seed, shape_of_zone=[float(z % 2) for z in range(nz)])
    return spec


def ensemble_instance(spec, k):
    """The k-th instance of an ensemble as a model of its own (same nodes,
    constants and contact table), for running it separately."""
    import copy

    import numpy as _np

    first, last, _ = spec.info["instances"][k]
    p0, p1 = spec.particles[first], spec.particles[last - 1]
    n0, n1 = p0["glob_start"], p1["glob_end"]
    s = copy.copy(spec)
    s.x_ref = spec.x_ref[n0:n1].copy()
    s.u = spec.u[n0:n1].copy()
    s.v = spec.v[n0:n1].copy()
    s.vol = spec.vol[n0:n1].copy()
    s.fix = spec.fix[n0:n1].copy()
    s.pt_id = spec.pt_id[n0:n1] - first
    s.n_nodes = int(n1 - n0)
    parts = []
    for p in spec.particles[first:last]:
        q = dict(p)
        q["glob_start"] -= n0
        q["glob_end"] -= n0
        q["center_node"] -= n0
        parts.append(q)
    s.particles = parts
    s.ubc, s.fbc = [], []
    s.info = dict(spec.info, instance=k)
    s.name = f"{spec.name}_i{k}"
    return s


# Prepared state of the 2D compressive test: the particles packed against the
# top wall, with contacts everywhere and the first bonds breaking.  The
# reference's own compressive example runs in stages restarted from an
# initial configuration (examples/PeriDEM/compressive_test/n500_circ_hex:
# init_config, run1, run2); this is that first stage, driven faster (a wall
# at 0.5 m/s stays stable; at 2 m/s it crushes the packing within 0.5 ms).
COMPACT_PREP_SPEED = 0.5      # m/s: top-wall speed of the preparation run
COMPACT_PREP_STEPS = 160000   # preparation steps at the deck's dt


def compacted_state(spec, prep_speed=COMPACT_PREP_SPEED, prep_steps=COMPACT_PREP_STEPS, precision=None,
                    chunk=20000):
    """Drive the top wall of a compressive test (built by compressive_2d) at
    prep_speed for prep_steps steps on the GPU (FP64 unless precision says
    otherwise), from the same seeded configuration, and return the state to
    restart the deck's own run from: positions, displacements, velocities,
    bond states and the restart step at which the deck's wall (speed
    spec.info['wall_speed']) stands where the preparation left it.  Applied
    with apply_state() to a GPU engine or the oracle alike."""
    ws = spec.info["wall_speed"]
    ratio = prep_speed / ws
    if abs(ratio - round(ratio)) > 1e-12:
        raise ValueError("the preparation speed must be a whole multiple of the deck's wall speed")
    prep = compressive_2d(**dict(spec.info["args"], wall_speed=prep_speed))
    e = capi.GpuEngine(prep, capi.PREC_FP64 if precision is None else precision)
    e.integrate_begin()
    done = 0
    while done < prep_steps:
        k = min(chunk, prep_steps - done)
        e.integrate_cd_graph(k)
        done += k
    st = {"step": int(round(prep_steps * ratio)), "prep_speed": prep_speed, "prep_steps": prep_steps,
          "x": e.field(capi.FIELD_X), "u": e.field(capi.FIELD_U), "v": e.field(capi.FIELD_V),
          "bonds": e.field(capi.FIELD_BONDS)}
    st["penetration"], st["reaction"] = e.pp_compressive_test(prep.wall_id, prep.f_dir)
    st["broken_bonds"] = e.broken_bonds()
    e.close()
    return st


def apply_state(engine, st):
    """Load a compacted_state() into an engine of the deck's own model, as a
    restart at st['step'] (call before integrate_begin)."""
    engine.set_field(capi.FIELD_X, st["x"])
    engine.set_field(capi.FIELD_U, st["u"])
    engine.set_field(capi.FIELD_V, st["v"])
    engine.set_field(capi.FIELD_BONDS, st["bonds"])
    engine.restart(st["step"])
