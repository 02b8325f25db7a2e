"""Benchmark: mesh-node updates per second of the explicit PeriDEM step on one MI355X.

A step is one central-difference time step (node update, displacement BC,
peridynamic forces, contact search at its natural cadence, contact and
damping forces, external forces) over the whole synthetic model, with all
state resident in HBM.

Default workload (the JSON line's value): the 3D packing of 2000 spheres
(~1000 nodes each, 2.08M nodes), state-based PD, mixed precision (FP32 bond
arithmetic, FP64 dilation sums and integration).  Carried alongside under
"compressive_2d": the 2D compressive test (500 particles, ~400 nodes each,
198k nodes, FP64 parity mode) timed from its compacted state (particles
packed against the top wall, contacts everywhere), and under "closure" the
learned contact closure against the plain method on that state and on a
batched ensemble.  Other workloads: --workload ensemble_2d | compressive_2d |
two_particle.

value = (nodes integrated per step) x steps / seconds of the timed region
(HIP-graph replay of exactly --steps steps).  "steady_state" repeats the
measurement over a whole number of contact-rebuild intervals.  The roofline
entry is for the peridynamic bond passes: the algorithm's bytes per launch
divided by the average duration of the passes, timed with HIP events over a
second, eagerly launched region of the same step (around the kernel on the
engine stream and around the tiles that run beside it on the side stream;
per step the longer of the two).
Before the warm-up the launch shape of the bond passes is autotuned (outside
the timed regions).
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "mesh-node updates/sec at 1 GPU, % HBM roofline; closure vs plain at fixed error"
HBM_PEAK_GBS = 8000.0
DEFAULT_PRECISION = {"packing_3d": "mixed", "ensemble_2d": "mixed", "compressive_2d": "fp64",
                     "two_particle": "fp64"}


def workload(name: str):
    from peridem_amd import configs

    if name == "compressive_2d":
        return configs.compressive_2d(n_particles=500, nodes_per_particle=400, seed=0)
    if name == "packing_3d":
        return configs.packing_3d(n_spheres=2000, nodes_per_sphere=1000, seed=0)
    if name == "two_particle":
        return configs.two_particle()
    if name == "ensemble_2d":
        return configs.ensemble_2d(n_instances=512, seed=0)
    raise SystemExit(f"unknown workload {name}")


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def pd_bytes_per_launch(spec, eng, precision):
    """Algorithmic HBM bytes of one pass over the bonds (every array element
    the passes must touch, counted once; neighbor data reused on chip), per
    force-computing node with N_i bonds:
      canonical (FP64 state, 4-byte ids, 1-bit bonds):
        bond-based RNP                     24 d + 12 + 4.25 N_i  (xRef, u, f: 8d each; V 8; Z 4)
        state-based, dilation and force    24 d + 20 + 4.25 N_i  (adds m_x; theta kept on chip)
      this build's layout:
        FP64 RNP kernel                    24 d + 16 + 4.25 N_i  (adds the 4-byte degree)
        FP64 RNP pair-tile kernel          24 d + 16 + 8.25 N_i  (adds the 2-byte pair slot of each
                                                                   bond and the 4-byte pair entry)
        mixed state-based passes           24 d + 36 + 4.375 N_i (theta written, degree, thread slot;
                                                                   the 2-byte ids and the bond words
                                                                   read by both passes)
    Returns (bytes in this build's layout, canonical bytes, mean N_i)."""
    import numpy as np
    from peridem_amd import capi

    d = spec.dim
    rp, _ = eng.neighbors(capi.NEIGH_PD)
    deg = np.diff(rp)
    cf = np.array([p["compute_force"] for p in spec.particles], bool)[spec.pt_id]
    state = np.array([p["material_type"] == capi.MAT_PDSTATE for p in spec.particles], bool)[spec.pt_id]
    nb = deg[cf].astype(np.float64)
    st = state[cf]
    canon = np.where(st, 24 * d + 20 + 4.25 * nb, 24 * d + 12 + 4.25 * nb)
    if precision == capi.PREC_MIXED:
        layout = np.where(st, 24 * d + 36 + 4.375 * nb, 24 * d + 16 + 4.25 * nb)
    else:
        rnp = 24 * d + 16 + (8.25 if eng.launch_config()[0] == 512 else 4.25) * nb
        layout = np.where(st, 24 * d + 32 + 4.125 * nb, rnp)
    return float(layout.sum()), float(canon.sum()), float(nb.mean()) if len(nb) else 0.0


def read_traffic(kernel_tag: str):
    """HBM bytes per launch of the bond passes from the committed rocprofv3
    --pmc summary of this benchmark's command (FETCH_SIZE + WRITE_SIZE, scaled;
    see the file), or None."""
    rel = os.path.join("profiles", f"{kernel_tag}_traffic.txt")
    path = os.path.join(ROOT, rel)
    if not os.path.exists(path):
        return None, None
    for line in open(path):
        if line.startswith("hbm_bytes_per_launch"):
            return float(line.split("=")[1]), rel
    return None, None


def kernel_names(spec, eng, prec):
    from peridem_amd import capi

    state = any(p["material_type"] == capi.MAT_PDSTATE and p["compute_force"] for p in spec.particles)
    if prec == capi.PREC_MIXED and state:
        return ("k_mixed_state (+ k_mixed_dilation, k_mixed_force of open tiles)", "mixed_state",
                "FP32 VALU issue (about one wave instruction per CU cycle, 57 per bond and pass)")
    if state:
        return "k_pd_dilation + k_pd_state", "pd_state", "FP64 VALU"
    if eng.launch_config()[0] == 512:
        return "k_pd_rnp_tile", "pd_rnp_tile", "FP64 VALU"
    return "k_pd_rnp", "pd_rnp", "FP64 VALU"


def measure(spec, prec, steps, warmup, prof_steps, autotune=True, state=None, steady_cycles=10):
    """Time one workload on the GPU.  Returns (open engine, result dict)."""
    from peridem_amd import capi, configs

    t0 = time.perf_counter()
    eng = capi.GpuEngine(spec, prec)
    if state is not None:
        configs.apply_state(eng, state)
    eng.integrate_begin()
    t_setup = time.perf_counter() - t0
    tms = round(eng.autotune(reps=5), 5) if autotune else None
    tuned = dict(zip(("pd_tpb", "mixed_tpb", "tile_nodes"), eng.launch_config()), bond_passes_ms=tms,
                 autotuned=autotune)
    if prec == capi.PREC_MIXED:
        tuned["tiles_lds"], tuned["tiles_hbm"] = eng.tile_counts()
    print(f"# {spec.name}: {spec.n_nodes} nodes, setup {t_setup:.1f} s, launch {tuned}", file=sys.stderr,
          flush=True)
    eng.integrate_cd_graph(warmup)
    eng.synchronize()
    nodes = spec.compute_nodes()
    r0 = eng.contact_state()["n_rebuilds"]
    ta = time.perf_counter()
    eng.integrate_cd_graph(steps)
    eng.synchronize()
    secs = time.perf_counter() - ta
    r1 = eng.contact_state()["n_rebuilds"]
    # steady state: a whole number of the deck's contact-rebuild intervals
    deck_interval = max(1, int(spec.neigh_update_interval))
    m = steady_cycles * deck_interval
    ta = time.perf_counter()
    eng.integrate_cd_graph(m)
    eng.synchronize()
    msecs = time.perf_counter() - ta
    r2 = eng.contact_state()["n_rebuilds"]
    # eager region with HIP events around the bond passes
    eng.profile_enable(True)
    eng.synchronize()
    ta = time.perf_counter()
    eng.integrate_cd(prof_steps)
    eng.synchronize()
    t_prof = time.perf_counter() - ta
    kms, nlaunch = eng.profile_get()
    eng.profile_enable(False)
    bytes_layout, bytes_canon, mean_nbr = pd_bytes_per_launch(spec, eng, prec)
    kname, ktag, limiter = kernel_names(spec, eng, prec)
    ktag = f"{ktag}_{spec.name}"
    achieved = bytes_canon / (kms * 1e-3) / 1e9 if kms > 0 else 0.0
    achieved_layout = bytes_layout / (kms * 1e-3) / 1e9 if kms > 0 else 0.0
    traffic, traffic_src = read_traffic(ktag)
    roof = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
            "traffic_source": (f"copied from {traffic_src} (rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE runs of "
                               f"this workload, scaled as the file states); not measured in this run")
            if traffic_src else None,
            "kernel": kname, "limiter": limiter, "kernel_ms_avg": round(kms, 5), "launches_timed": nlaunch,
            "bytes_per_launch": bytes_canon, "bytes_per_launch_layout": bytes_layout,
            "achieved_layout": round(achieved_layout, 2), "frac_layout": round(achieved_layout / HBM_PEAK_GBS, 4),
            "mean_neighbors": round(mean_nbr, 2),
            "timing": f"HIP events around the bond passes (dilation and force) of {prof_steps} eager steps "
                      f"({t_prof:.3f} s): around the kernel on the engine stream and, where tiles run beside it "
                      f"on the side stream, around those; per step the longer of the two"}
    n, t = eng.time()
    st = eng.contact_state()
    broken, total = eng.broken_bonds()
    res = {"value": nodes * steps / secs, "ms_per_step": secs / steps * 1e3, "nodes": spec.n_nodes,
           "nodes_integrated": nodes, "rebuilds_in_timed_region": r1 - r0,
           "steady_state": {"steps": m, "deck_rebuild_interval": deck_interval, "rebuilds": r2 - r1,
                            "ms_per_step": round(msecs / m * 1e3, 5), "value": round(nodes * m / msecs, 1)},
           "launch": tuned, "roofline": roof, "sim_step_end": n, "sim_time_end": t,
           "contact_rebuilds": st["n_rebuilds"], "contact_interval": st["interval"],
           "contact_candidates": eng.neighbor_count(capi.NEIGH_CONTACT),
           "broken_bond_fraction": broken / max(1, total)}
    print(f"# {spec.name}: {steps} steps {secs:.4f} s ({res['ms_per_step']:.4f} ms/step, "
          f"{r1 - r0} rebuilds); steady {res['steady_state']}; bond passes {kms:.4f} ms "
          f"({roof['frac']:.3f} of HBM peak)", file=sys.stderr, flush=True)
    return eng, res


def cpu_baseline(spec, prec, nthreads, budget_s=12.0, max_steps=-1, state=None):
    """The FP64 oracle on the host cores: the same workload from the same
    state, timed over a bounded sample of steps (chunks of steps until about
    budget_s seconds have run).  The same number of steps is then run on the
    GPU from the same state and its forces are compared with the oracle's
    (max and RMS over nodes of |f - f_oracle|, over the largest oracle force);
    in mixed precision also on the oracle's state with every bond loaded by a
    smooth displacement (a state at rest only carries gravity)."""
    import numpy as np
    import oracle
    from peridem_amd import capi, configs

    t0 = time.perf_counter()
    o = oracle.OracleEngine(spec, nthreads)
    if state is not None:
        configs.apply_state(o, state)
    o.integrate_begin()
    o.integrate_cd(2)  # warm-up
    t1 = time.perf_counter()
    steps, chunk = 0, 2
    while True:
        o.integrate_cd(chunk)
        steps += chunk
        el = time.perf_counter() - t1
        if el >= budget_s or (max_steps > 0 and steps >= max_steps):
            break
        chunk = max(1, min(chunk * 2, int(chunk * (budget_s - el) / max(el, 1e-3)) + 1))
    t2 = time.perf_counter()
    fo = o.field(capi.FIELD_F)
    g = capi.GpuEngine(spec, prec)
    if state is not None:
        configs.apply_state(g, state)
    g.integrate_begin()
    g.integrate_cd(2 + steps)
    fg = g.field(capi.FIELD_F)

    def rel(fa, fb):
        den = float(np.abs(fb).max())
        dd = np.abs(fa - fb)
        return {"max": float(dd.max() / den) if den > 0 else float(dd.max()),
                "rms": float(np.sqrt((dd ** 2).sum(1).mean()) / den) if den > 0 else 0.0, "max_force": den}

    err = dict(rel(fg, fo), steps=2 + steps)
    if prec == capi.PREC_MIXED:
        # the oracle's state with a smooth displacement added (peak strain
        # 1e-4, wavelength 4 mesh spacings), so that every bond carries load,
        # set on both engines; the forces of that state compared
        xr = o.field(capi.FIELD_XREF)
        h = min(p["mesh_size"] for p in spec.particles)
        ls, eps = 4.0 * h, 1e-4
        du = np.zeros_like(xr)
        for d in range(spec.dim):
            du[:, d] = eps * ls * np.sin(xr[:, (d + 1) % spec.dim] / ls + d)
        x, u, v, b = (o.field(f) for f in (capi.FIELD_X, capi.FIELD_U, capi.FIELD_V, capi.FIELD_BONDS))
        for e in (o, g):
            e.set_field(capi.FIELD_X, x + du)
            e.set_field(capi.FIELD_U, u + du)
            e.set_field(capi.FIELD_V, v)
            e.set_field(capi.FIELD_BONDS, b)
            e.integrate_begin()
        err["deformed_state"] = dict(rel(g.field(capi.FIELD_F), o.field(capi.FIELD_F)), peak_strain=eps,
                                     wavelength_m=ls)
    o.close()
    g.close()
    nodes = spec.compute_nodes()
    return ({"value": round(nodes * steps / (t2 - t1), 1), "unit": "node-updates/s",
             "cores": nthreads, "kind": "port", "cpu_model": cpu_model(),
             "sample": f"{steps} steps ({t2 - t1:.1f} s) of the same workload from the same state after setup "
                       f"({t1 - t0:.1f} s) and 2 warm-up steps; FP64 oracle, {nthreads} threads"}, err)


def _closure_run(spec, precision, net, steps, sample_every, reaction, state=None):
    """Plain (net None) or closure run of `steps` steps; returns ms/step, the
    sampled observable curve and the closure statistics."""
    import numpy as np
    from peridem_amd import capi, closure, configs

    e = capi.GpuEngine(spec, precision)
    e.closure_setup(closure.particle_codes(spec))
    if net is not None:
        e.closure_enable(net, pen_max_h=0.05, dz_max=0.5, check_every=100)
    if state is not None:
        configs.apply_state(e, state)
    e.integrate_begin()
    e.autotune(reps=3)  # both forms at their tuned bond-kernel shape
    e.integrate_cd_graph(10)
    e.synchronize()
    curve, tt, done = [], 0.0, 0
    mass = spec.vol * np.array([p["density"] for p in spec.particles])[spec.pt_id]
    while done < steps:
        k = min(sample_every, steps - done)
        ta = time.perf_counter()
        e.integrate_cd_graph(k)
        e.synchronize()
        tt += time.perf_counter() - ta
        done += k
        if reaction:
            curve.append(e.pp_compressive_test(spec.wall_id, spec.f_dir)[1])
        else:  # kinetic energy of the model
            v = e.field(capi.FIELD_V)
            curve.append(float(0.5 * (v ** 2).sum(1) @ mass))
    st = e.closure_stats() if net is not None else None
    e.close()
    return tt / steps * 1e3, np.array(curve), st


def closure_comparison(c1_spec, c1_state, steps, sample_every, train_instances, train_steps, eval_steps):
    """Learned closure vs plain method: train the network on a batched ensemble
    (plain method, mixed precision), then time both forms on the 2D
    compressive test from its compacted state (FP64; reaction-force curve
    compared) and on a fresh ensemble (mixed; kinetic-energy curve compared)."""
    import numpy as np
    from peridem_amd import capi, closure, configs

    t0 = time.perf_counter()
    train = configs.ensemble_2d(n_instances=train_instances, seed=1000, pitch_factor=1.0)
    X, T = closure.collect(train, steps=train_steps, log_every=10, band_h=0.5, cap=1 << 18)
    Xs, Ts = closure.select(X, T)
    t1 = time.perf_counter()
    net, curve = closure.fit(Xs, Ts, iters=4000, batch=1024, lr=2e-3, seed=0)
    t2 = time.perf_counter()
    print(f"# closure: {len(X)} logged rows, {len(Xs)} in the band ({t1 - t0:.1f} s); "
          f"training {t2 - t1:.1f} s, loss {curve[:10].mean():.4f} -> {curve[-50:].mean():.5f}",
          file=sys.stderr, flush=True)
    out = {"train": {"ensemble_instances": train_instances, "ensemble_steps": train_steps,
                     "rows_logged": int(len(X)), "rows_used": int(len(Xs)),
                     "held_out_rel_rms": [round(float(r), 4) for r in closure.fit_report(net, Xs[::7], Ts[::7])],
                     "collect_s": round(t1 - t0, 2), "train_s": round(t2 - t1, 2)}}
    cases = (("compressive_2d", c1_spec, c1_state, capi.PREC_FP64, steps, sample_every, True),
             ("ensemble_2d", configs.ensemble_2d(n_instances=train_instances, seed=2000, pitch_factor=1.0), None,
              capi.PREC_MIXED, eval_steps, max(1, eval_steps // 20), False))
    for name, spec, state, prec, n, every, reaction in cases:
        mp, cp, _ = _closure_run(spec, prec, None, n, every, reaction, state)
        mc, cc, st = _closure_run(spec, prec, net, n, every, reaction, state)
        err = float(np.abs(cc - cp).max() / max(np.abs(cp).max(), 1e-300))
        out[name] = {"workload": spec.name + (" (compacted state)" if state is not None else ""), "steps": n,
                     "speedup": round(mp / mc, 4), "ms_per_step_plain": round(mp, 5),
                     "ms_per_step_closure": round(mc, 5),
                     ("reaction_curve_err" if reaction else "kinetic_energy_curve_err"): err, "pairs": st}
        print(f"# closure {name}: {out[name]}", file=sys.stderr, flush=True)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=400)
    ap.add_argument("--warmup", type=int, default=80)
    ap.add_argument("--workload", default="packing_3d")
    ap.add_argument("--precision", default=None, choices=["fp64", "mixed"])
    ap.add_argument("--cpu-steps", type=int, default=-1)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--prof-steps", type=int, default=200)
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--no-c1", action="store_true", help="skip the compressive-test line of the default run")
    ap.add_argument("--c1-steps", type=int, default=2000)
    ap.add_argument("--no-closure", action="store_true")
    ap.add_argument("--closure-steps", type=int, default=5000)
    ap.add_argument("--closure-train-instances", type=int, default=512)
    ap.add_argument("--closure-train-steps", type=int, default=5000)
    ap.add_argument("--closure-eval-steps", type=int, default=5000)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only; the multi-GPU (particle-sharded) "
              "form is a design, not part of this benchmark.", file=sys.stderr)
        raise SystemExit(2)

    from peridem_amd import capi, configs

    assert capi.device_count() > 0, "bench.py needs a GPU"
    pname = args.precision or DEFAULT_PRECISION.get(args.workload, "fp64")
    prec = capi.PREC_FP64 if pname == "fp64" else capi.PREC_MIXED
    nthreads = args.cpu_threads or min(16, os.cpu_count() or 1)
    spec = workload(args.workload)
    state = None
    if args.workload == "compressive_2d":
        state = configs.compacted_state(spec)
    eng, res = measure(spec, prec, args.steps, args.warmup, args.prof_steps, not args.no_autotune, state)
    eng.close()
    cpu, err = None, None
    if not args.no_cpu:
        cpu, err = cpu_baseline(spec, prec, nthreads, args.cpu_seconds, args.cpu_steps, state)
        print(f"# cpu baseline {cpu}; GPU vs oracle {err}", file=sys.stderr, flush=True)

    c1 = None
    c1_spec = c1_state = None
    if args.workload == "packing_3d" and not args.no_c1:
        c1_spec = workload("compressive_2d")
        tp = time.perf_counter()
        c1_state = configs.compacted_state(c1_spec)
        prep_s = time.perf_counter() - tp
        e1, r1 = measure(c1_spec, capi.PREC_FP64, args.c1_steps, 200, args.prof_steps, not args.no_autotune,
                         c1_state)
        e1.close()
        c1 = dict(r1, workload=c1_spec.name, precision="fp64", dtype="f64",
                  compacted_state={"prep_wall_speed": c1_state["prep_speed"], "prep_steps": c1_state["prep_steps"],
                                   "restart_step": c1_state["step"], "prep_s": round(prep_s, 2),
                                   "broken_bonds": list(c1_state["broken_bonds"])})
        if not args.no_cpu:
            c1["cpu_baseline"], c1["parity_vs_oracle"] = cpu_baseline(c1_spec, capi.PREC_FP64, nthreads,
                                                                      args.cpu_seconds, args.cpu_steps, c1_state)

    closure_res = None
    if not args.no_closure and (c1_spec is not None or args.workload == "compressive_2d"):
        if c1_spec is None:
            c1_spec, c1_state = spec, state
        closure_res = closure_comparison(c1_spec, c1_state, args.closure_steps, 250, args.closure_train_instances,
                                         args.closure_train_steps, args.closure_eval_steps)

    cfg = {"workload": spec.name, "nodes": spec.n_nodes, "nodes_integrated": res["nodes_integrated"],
           "particles": spec.n_particles, "dim": spec.dim, "dt": spec.dt, "precision": pname,
           "step_mode": "hip graph replay", "rebuilds_in_timed_region": res["rebuilds_in_timed_region"],
           "sim_step_end": res["sim_step_end"], "contact_rebuilds": res["contact_rebuilds"],
           "contact_interval": res["contact_interval"], "contact_candidates": res["contact_candidates"],
           "broken_bond_fraction": res["broken_bond_fraction"], "launch": res["launch"]}
    if state is not None:
        cfg["compacted_state"] = {"restart_step": state["step"], "prep_wall_speed": state["prep_speed"],
                                  "prep_steps": state["prep_steps"]}
    out = {
        "metric": METRIC, "value": round(res["value"], 1), "unit": "node-updates/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(res["ms_per_step"], 5),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if prec == capi.PREC_FP64 else "f32/f64",
        "data": "synthetic (seeded generator, reference deck semantics)",
        "config": cfg, "steady_state": res["steady_state"], "roofline": res["roofline"], "cpu_baseline": cpu,
    }
    if err is not None:
        out["force_error_vs_oracle" if prec == capi.PREC_MIXED else "parity_vs_oracle"] = err
    if c1 is not None:
        out["compressive_2d"] = c1
    if closure_res is not None:
        out["closure"] = closure_res
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
