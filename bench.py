"""Benchmark: mesh-node updates per second of the explicit PeriDEM step on one MI355X.

A step is one central-difference time step (node update, displacement BC,
peridynamic forces, contact search at its natural cadence, contact and
damping forces, external forces) over the whole synthetic model, with all
state resident in HBM.  Default workload: the 2D compressive test with 500
particles (~400 nodes each, ~200k nodes), FP64.  Other workloads: the 3D
packing of 2000 spheres (mixed precision) and a batched ensemble of 512
independent 2D simulations in one model.  Before the warm-up the launch
shape of the bond-force pass is autotuned (outside the timed regions).

Prints one JSON line.  value = (nodes integrated per step) x steps / seconds
of the timed region (HIP-graph replay of the step).  The roofline entry is
for the peridynamic bond-force kernel: algorithmic bytes per launch divided
by its average duration, timed with HIP events on the engine's stream over a
second, eagerly launched timed region of the same step.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "mesh-node updates/sec at 1 GPU, % HBM roofline; closure vs plain at fixed error"
HBM_PEAK_GBS = 8000.0


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


def pd_bytes_per_launch(spec, eng, precision):
    """Algorithmic HBM bytes of one bond-force launch: every array element the
    kernel must touch, counted once per launch (neighbor data assumed reused
    on chip).  Returns (bytes in this build's layout, canonical bytes).
      canonical (FP64 state, 4-byte ids, 1-bit bonds):
        bond-based RNP            24 d + 12 + 4.25 N_i   (xRef, u, f: 8d each; V 8; Z 4)
        state-based force pass    24 d + 28 + 4.125 N_i  (adds m_x, theta; bonds read)
      this build:
        FP64 RNP kernel           24 d + 16 + 4.25 N_i   (adds the 4-byte degree)
        FP64 RNP pair-tile kernel 24 d + 16 + 8.25 N_i   (adds the 2-byte pair slot of
                                                           each bond and the 4-byte pair
                                                           list entry of each pair; the
                                                           per-pair scratch it writes and
                                                           reads back in L2 is not counted)
        mixed state force kernel  24 d + 28 + 2.125 N_i  (2-byte ids; alpha, b terms
                                                           FP32 instead of m_x, theta; the
                                                           4-byte thread slot -> node entry)
    """
    import numpy as np
    from peridem_amd import capi

    d = spec.dim
    rp, _ = eng.neighbors(capi.NEIGH_PD)
    deg = np.diff(rp)
    cf = np.array([p["compute_force"] for p in spec.particles], bool)[spec.pt_id]
    state = np.array([p["material_type"] == capi.MAT_PDSTATE for p in spec.particles], bool)[spec.pt_id]
    nb = deg[cf].astype(np.float64)
    st = state[cf]
    canon = np.where(st, 24 * d + 28 + 4.125 * nb, 24 * d + 12 + 4.25 * nb)
    if precision == capi.PREC_MIXED:
        layout = np.where(st, 24 * d + 28 + 2.125 * nb, 24 * d + 16 + 4.25 * nb)
    else:
        rnp = 24 * d + 16 + (8.25 if eng.launch_config()[0] == 512 else 4.25) * nb
        layout = np.where(st, 24 * d + 32 + 4.125 * nb, rnp)
    return float(layout.sum()), float(canon.sum()), float(nb.mean()) if len(nb) else 0.0


def force_error_vs_fp64(spec, eng):
    """Mixed precision against the FP64 parity path (itself equal to the
    oracle bit for bit in the tests) at full size: an FP64 engine takes the
    mixed run's final state (x, u, v, bond states) plus a smooth displacement
    perturbation (peak strain ~1e-4, so that every bond carries force even
    where bodies still move rigidly), both recompute the peridynamic forces,
    and the difference is reported as max and RMS over nodes, normalised by
    the largest FP64 force."""
    import numpy as np
    from peridem_amd import capi

    h = min(p["mesh_size"] for p in spec.particles)
    lam = 10.0 * h
    amp = 1e-4 * lam / (2 * np.pi)
    xr = spec.x_ref
    du = np.zeros_like(xr)
    for d in range(spec.dim):
        du[:, d] = amp * np.sin(2 * np.pi * xr[:, (d + 1) % spec.dim] / lam + d)
    u = eng.field(capi.FIELD_U) + du
    x = eng.field(capi.FIELD_X) + du
    ref = capi.GpuEngine(spec, capi.PREC_FP64)
    ref.integrate_begin()
    for e in (eng, ref):
        e.set_field(capi.FIELD_U, u)
        e.set_field(capi.FIELD_X, x)
    ref.set_field(capi.FIELD_V, eng.field(capi.FIELD_V))
    ref.set_field(capi.FIELD_BONDS, eng.field(capi.FIELD_BONDS))
    eng.compute_peridynamic_forces()
    ref.compute_peridynamic_forces()
    fm, fr = eng.field(capi.FIELD_F), ref.field(capi.FIELD_F)
    ref.close()
    den = float(np.abs(fr).max())
    d = np.abs(fm - fr)
    return {"max": float(d.max() / den), "rms": float(np.sqrt((d ** 2).sum(1).mean()) / den),
            "state": "final state of the eager region plus a smooth displacement perturbation "
                     "(wavelength 10 h, peak strain 1e-4); peridynamic forces of both recomputed on it",
            "max_force": den}


def read_traffic(kernel_tag: str):
    path = os.path.join(ROOT, "profiles", f"{kernel_tag}_traffic.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("hbm_bytes_per_launch"):
            return float(line.split("=")[1])
    return None


def cpu_baseline(spec, nthreads, budget_s=12.0, max_steps=-1):
    """The FP64 oracle on the host cores: the same workload, timed over a bounded
    sample of steps (chunks of steps until about budget_s seconds have run)."""
    import oracle

    t0 = time.perf_counter()
    o = oracle.OracleEngine(spec, nthreads)
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
    o.close()
    nodes = spec.compute_nodes()
    return {"value": round(nodes * steps / (t2 - t1), 1), "unit": "node-updates/s",
            "cores": nthreads, "kind": "port",
            "sample": f"{steps} steps ({t2 - t1:.1f} s) of the same workload after setup "
                      f"({t1 - t0:.1f} s) and 2 warm-up steps; FP64 oracle, {nthreads} threads"}


def _closure_run(spec, precision, net, steps, sample_every, reaction):
    """Plain (net None) or closure run of `steps` steps; returns ms/step, the
    sampled observable curve and the closure statistics."""
    import numpy as np
    from peridem_amd import capi, closure

    e = capi.GpuEngine(spec, precision)
    e.closure_setup(closure.particle_codes(spec))
    if net is not None:
        e.closure_enable(net, pen_max_h=0.05, dz_max=0.5, check_every=100)
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


def closure_comparison(c1_spec, steps, sample_every, train_instances, train_steps, eval_steps, verbose=True):
    """Learned closure vs plain method: train the network on a batched ensemble
    (plain method, mixed precision), then time both forms on the 2D
    compressive test (FP64; reaction-force curve compared) and on a fresh
    ensemble (mixed; kinetic-energy curve compared)."""
    import numpy as np
    from peridem_amd import capi, closure, configs

    t0 = time.perf_counter()
    train = configs.ensemble_2d(n_instances=train_instances, seed=1000, pitch_factor=1.0)
    X, T = closure.collect(train, steps=train_steps, log_every=10, band_h=0.5, cap=1 << 18)
    Xs, Ts = closure.select(X, T)
    t1 = time.perf_counter()
    net, curve = closure.fit(Xs, Ts, iters=4000, batch=1024, lr=2e-3, seed=0)
    t2 = time.perf_counter()
    if verbose:
        print(f"# closure: {len(X)} logged rows, {len(Xs)} in the band ({t1 - t0:.1f} s); "
              f"training {t2 - t1:.1f} s, loss {curve[:10].mean():.4f} -> {curve[-50:].mean():.5f}",
              file=sys.stderr, flush=True)
    out = {"train": {"ensemble_instances": train_instances, "ensemble_steps": train_steps,
                     "rows_logged": int(len(X)), "rows_used": int(len(Xs)),
                     "held_out_rel_rms": [round(float(r), 4) for r in closure.fit_report(net, Xs[::7], Ts[::7])],
                     "collect_s": round(t1 - t0, 2), "train_s": round(t2 - t1, 2)}}
    cases = (("compressive_2d", c1_spec, capi.PREC_FP64, steps, sample_every, True),
             ("ensemble_2d", configs.ensemble_2d(n_instances=train_instances, seed=2000, pitch_factor=1.0),
              capi.PREC_MIXED, eval_steps, max(1, eval_steps // 20), False))
    for name, spec, prec, n, every, reaction in cases:
        mp, cp, _ = _closure_run(spec, prec, None, n, every, reaction)
        mc, cc, st = _closure_run(spec, prec, net, n, every, reaction)
        err = float(np.abs(cc - cp).max() / max(np.abs(cp).max(), 1e-300))
        out[name] = {"workload": spec.name, "steps": n, "speedup": round(mp / mc, 4),
                     "ms_per_step_plain": round(mp, 5), "ms_per_step_closure": round(mc, 5),
                     ("reaction_curve_err" if reaction else "kinetic_energy_curve_err"): err,
                     "pairs": st}
        if verbose:
            print(f"# closure {name}: {out[name]}", file=sys.stderr, flush=True)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--workload", default="compressive_2d")
    ap.add_argument("--precision", default="fp64", choices=["fp64", "mixed"])
    ap.add_argument("--cpu-steps", type=int, default=-1)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--prof-steps", type=int, default=300)
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--no-closure", action="store_true")
    ap.add_argument("--no-error", action="store_true")
    ap.add_argument("--closure-steps", type=int, default=50000)
    ap.add_argument("--closure-train-instances", type=int, default=512)
    ap.add_argument("--closure-train-steps", type=int, default=5000)
    ap.add_argument("--closure-eval-steps", type=int, default=5000)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only; the multi-GPU (particle-sharded) "
              "form is a design in DESIGN.md and is not part of this benchmark.", file=sys.stderr)
        raise SystemExit(2)

    from peridem_amd import capi

    assert capi.device_count() > 0, "bench.py needs a GPU"
    t_setup = time.perf_counter()
    spec = workload(args.workload)
    prec = capi.PREC_FP64 if args.precision == "fp64" else capi.PREC_MIXED
    eng = capi.GpuEngine(spec, prec)
    print(f"# workload {spec.name}: {spec.n_nodes} nodes, {spec.n_particles} particles, "
          f"setup {time.perf_counter() - t_setup:.1f} s", file=sys.stderr, flush=True)
    eng.integrate_begin()
    tms = None
    if not args.no_autotune:
        tms = round(eng.autotune(reps=5), 5)
    tuned = dict(zip(("pd_tpb", "mixed_tpb", "tile_nodes"), eng.launch_config()), force_pass_ms=tms,
                 autotuned=not args.no_autotune)
    if args.precision == "mixed":
        tuned["tiles_lds"], tuned["tiles_hbm"] = eng.tile_counts()
    print(f"# launch: {tuned}", file=sys.stderr, flush=True)
    eng.integrate_cd_graph(args.warmup)
    eng.synchronize()
    t0 = time.perf_counter()
    eng.integrate_cd_graph(args.steps)
    eng.synchronize()
    t1 = time.perf_counter()
    secs = t1 - t0
    nodes = spec.compute_nodes()
    value = nodes * args.steps / secs
    print(f"# graph region: {args.steps} steps in {secs:.3f} s", file=sys.stderr, flush=True)

    # eager region with HIP events around the bond-force kernel
    eng.profile_enable(True)
    eng.synchronize()
    t2 = time.perf_counter()
    eng.integrate_cd(args.prof_steps)
    eng.synchronize()
    t3 = time.perf_counter()
    kms, nlaunch = eng.profile_get()
    eng.profile_enable(False)
    bytes_launch, bytes_canon, mean_nbr = pd_bytes_per_launch(spec, eng, prec)
    achieved = bytes_launch / (kms * 1e-3) / 1e9 if kms > 0 else 0.0
    achieved_canon = bytes_canon / (kms * 1e-3) / 1e9 if kms > 0 else 0.0
    state = any(p["material_type"] == capi.MAT_PDSTATE for p in spec.particles)
    mixed = prec == capi.PREC_MIXED and state
    kname = "k_mixed_force" if mixed else ("k_pd_state" if state else
                                           ("k_pd_rnp_tile" if eng.launch_config()[0] == 512 else "k_pd_rnp"))
    ktag = kname[2:] + "_" + spec.name
    # achieved: the algorithm's own bytes (canonical FP64 state, 4-byte ids,
    # 1-bit bonds), independent of this build's layout and index arrays
    roof = {"bound": "hbm", "achieved": round(achieved_canon, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved_canon / HBM_PEAK_GBS, 4), "traffic": read_traffic(ktag),
            "kernel": kname,
            "kernel_ms_avg": round(kms, 5), "launches_timed": nlaunch,
            "bytes_per_launch": bytes_canon, "bytes_per_launch_layout": bytes_launch,
            "achieved_layout": round(achieved, 2),
            "frac_layout": round(achieved / HBM_PEAK_GBS, 4),
            "mean_neighbors": round(mean_nbr, 2),
            "timing": f"HIP events on the engine stream over {args.prof_steps} eager steps "
                      f"({t3 - t2:.3f} s)"}
    n, t = eng.time()
    st = eng.contact_state()
    n_cand = eng.neighbor_count(capi.NEIGH_CONTACT)
    broken, total = eng.broken_bonds()
    ferr = None
    if mixed and not args.no_error:
        ferr = force_error_vs_fp64(spec, eng)
        print(f"# force error vs FP64: {ferr}", file=sys.stderr, flush=True)
    eng.close()

    closure_res = None
    if not args.no_closure and args.workload == "compressive_2d":
        closure_res = closure_comparison(spec, args.closure_steps, 500, args.closure_train_instances,
                                         args.closure_train_steps, args.closure_eval_steps)

    cpu = None
    if not args.no_cpu:
        nthreads = args.cpu_threads or min(16, os.cpu_count() or 1)
        cpu = cpu_baseline(spec, nthreads, args.cpu_seconds, args.cpu_steps)
    out = {
        "metric": METRIC, "value": round(value, 1), "unit": "node-updates/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(secs / args.steps * 1e3, 5),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if prec == capi.PREC_FP64 else "f32/f64",
        "data": "synthetic (seeded generator, reference deck semantics)",
        "config": {"workload": spec.name, "nodes": spec.n_nodes, "nodes_integrated": nodes,
                   "particles": spec.n_particles, "dim": spec.dim, "dt": spec.dt,
                   "precision": args.precision, "step_mode": "hip graph replay",
                   "sim_time_end": t, "contact_rebuilds": st["n_rebuilds"],
                   "contact_interval": st["interval"], "contact_candidates": n_cand,
                   "broken_bond_fraction": broken / max(1, total), "launch": tuned},
        "roofline": roof, "cpu_baseline": cpu, "closure": closure_res,
    }
    if ferr is not None:
        out["force_error_vs_fp64"] = ferr
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
