"""Throughput of the particle hot path on one MI355X.

Default workload: LJ fluid, FCC 40^3 = 256,000 atoms at rho* = 0.85,
Nose-Hoover at T* = 0.85, dt = 0.005, rc = 2.5, neighbor cutoff 2.8
(skin 1.12 x 2 x 1.25), mixed precision (FP32 pair forces, FP64 state).
A step is one full scheduled step (integrate, rebuild when the skin is
spent, force, integrate + thermostat) of the whole system; particle-steps/s
= N x steps / wall seconds of the timed loop, with the state resident in HBM.

Prints one JSON line. Extra keys: the all-FP64 GPU run of the same workload,
the pair-force kernel's roofline (HIP-event timed inside the timed loop) and
the CPU baseline (the CPU checker on the host, 1 thread, a bounded sample).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "particle-steps/s (LJ 256k, DEM 4M) vs FP64 GPU & CPU; ensemble p50/p99 latency"
HBM_PEAK_GBS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def kernel_name(lg, mixed, fused=False):
    """Name of the pair-force kernel a layout runs, as the profiler shows it."""
    r = "float" if mixed else "double"
    if lg < 0:
        return f"k_lj_tile<{r},false,true>"
    return f"k_lj_post<{lg}>" if fused else f"k_lj<{r},false,{lg}>"


def post_fused(w, precision, lg):
    """Whether the library runs the thermostatted final half step inside the
    row pass (mdd_set_lj_post_fused, on by default): mixed Nose-Hoover LJ rows
    of at most 4 lanes, at most 262,144 particles."""
    from mddem_amd import hip
    return (w.model == "lj" and precision == hip.MIXED and w.nose_hoover is not None
            and 0 <= lg <= 2 and w.n <= 262144)


def layout_name(lg):
    return "cell tiles" if lg < 0 else f"rows, {1 << lg} lanes/particle"


def run_gpu(w, precision, steps, warmup, profile_every=10, tune=True, layout=None, graphs=True):
    """Warm up (the last four fifths of the warmup steps try the pair-force
    layouts: row kernels at 2, 4 and 8 lanes per particle and, for LJ, the
    cell-tile kernel; the one with the shortest wall time per step is kept),
    then time exactly `steps` steps."""
    from mddem_amd import hip
    s = hip.from_workload(w, precision, box_exit=False)
    s.set_launch_graphs(graphs)
    if layout is not None:
        s.set_force_lanes(layout)
        tune = False
    layout = 2 if layout is None else layout
    # set-up, untimed and outside the warmup count: run until the rebuild
    # graphs of both alternating buffer sets are captured and the rebuild
    # interval has been seen twice, so that a short warmup does not leave
    # graph capture or chunk-length learning inside the timed window
    setup = 0
    while setup < 600:
        th = s.thermo()
        if th["rebuilds"] >= 3 and s.launch_graph_stats()[1] >= (2 if graphs else 0):
            break
        s.run(20)
        setup += 20
    options = (-1, 1, 2, 3) if w.model == "lj" else (1, 2, 3)
    trial = warmup // (len(options) + 1) if tune else 0
    if trial >= 100:
        s.run(warmup - len(options) * trial - 1)  # equilibrate first
        best = None
        for lg in options:
            s.set_force_lanes(lg)
            s.run(trial // 4)  # relayout rebuild
            s.set_profiling(1)
            s.synchronize()
            t0 = time.perf_counter()
            s.run(trial - trial // 4)
            s.synchronize()
            per_step = (time.perf_counter() - t0) / (trial - trial // 4)
            ms, _ = s.force_kernel_stats()
            log(f"[bench]   tune: {layout_name(lg)}: {per_step * 1e6:.1f} us/step, "
                f"pair-force kernel {ms * 1e3:.1f} us")
            if best is None or per_step < best[0]:
                best = (per_step, lg)
        layout = best[1]
        s.set_force_lanes(layout)
        s.set_profiling(0)
        s.run(1)  # the last warmup step lays the list out for the chosen shape
    else:
        s.run(warmup)
    s.synchronize()
    th0 = s.thermo()
    s.set_profiling(profile_every)
    s.synchronize()
    t0 = time.perf_counter()
    s.run(steps)
    s.synchronize()
    t1 = time.perf_counter()
    th1 = s.thermo(energy=(w.model == "lj"))
    avg_ms, launches = s.force_kernel_stats()
    rb_ms, rb_gap_ms, rb_n = s.rebuild_stats()
    res = dict(seconds=t1 - t0, steps=steps, n=w.n, rebuilds=th1["rebuilds"] - th0["rebuilds"],
               npairs=th1["npairs"], force_ms=avg_ms, force_launches=launches,
               temperature=th1["temperature"], pe=th1["pe"], ke=th1["ke"],
               layout=layout_name(layout), tile=layout < 0, rebuild_ms=rb_ms,
               kernel_name=kernel_name(layout, precision == hip.MIXED,
                                       post_fused(w, precision, layout)),
               fused=post_fused(w, precision, layout), rebuild_gap_ms=rb_gap_ms,
               rebuilds_timed=rb_n, setup_steps=setup)
    s.close()
    return res


def run_dem(steps, warmup, n=4_000_000, log_every=0, every=500, tune_steps=200):
    """DEM line: the settling workload's 4M polydisperse glass spheres, walls
    and gravity, started as a bed at rest (every sphere touching the spheres
    it rests on), mixed and all-FP64 over the same window of the run: the
    FP64 leg first runs, untimed, the steps the mixed leg spends choosing
    the force pass's launch shape, then both take the same warmup and timed
    steps. Reports the mean half-list pairs k and contacts c per sphere over
    the timed window and the whole-step roofline with B_ps = S + 4 k + 56 c
    (S = 216 B mixed, 264 B FP64)."""
    from mddem_amd import hip
    from mddem_amd import workloads as W
    t0 = time.perf_counter()
    w = W.dem_packed_bed(n=n, every=every)
    log(f"[bench] dem bed of {w.n} spheres built in {time.perf_counter() - t0:.1f} s "
        f"(top at z = {w.pos[:, 2].max():.4f} m)")
    out = {}
    shape = None
    pre = 0
    for name, prec in (("mixed", hip.MIXED), ("fp64", hip.FP64)):
        s = hip.from_workload(w, prec, box_exit=False)
        if shape is None:
            # launch shape of the force pass, tuned during the first leg's
            # warmup: one lane per sphere (one kernel, or split by row
            # length) or 2 / 4 lanes per sphere; each trial after a relayout
            trials = []
            for lg, split in ((0, 0), (0, 1), (1, -1), (2, -1)):
                s.set_force_lanes(lg)
                s.set_dem_split(split)
                s.run(20)
                s.synchronize()
                t1 = time.perf_counter()
                s.run(tune_steps)
                s.synchronize()
                per = (time.perf_counter() - t1) / tune_steps
                trials.append((per, lg, split))
                log(f"[bench]   dem tune: lanes {1 << lg}, split {split}: {per * 1e3:.4f} ms/step")
            _, lg, split = min(trials)
            pre = s.thermo()["step"]
            shape = dict(lanes=1 << lg, split=split, trials=[
                dict(lanes=1 << t[1], split=t[2], ms_per_step=round(t[0] * 1e3, 4)) for t in trials])
        s.set_force_lanes({1: 0, 2: 1, 4: 2}[shape["lanes"]])
        s.set_dem_split(shape["split"])
        if s.thermo()["step"] < pre:
            s.run(pre - s.thermo()["step"])
        s.run(warmup)
        s.synchronize()
        th0 = s.thermo()
        s.set_profiling(5)
        # the timed window in segments, so that k and c are averaged over it
        seg = steps if log_every <= 0 else log_every
        pairs, contacts, done, wall = [], [], 0, 0.0
        while done < steps:
            k = min(seg, steps - done)
            s.synchronize()
            t0 = time.perf_counter()
            s.run(k)
            s.synchronize()
            wall += time.perf_counter() - t0
            done += k
            th = s.thermo()
            pairs.append(th["npairs"])
            contacts.append(th["ncontacts"])
        th1 = s.thermo()
        ms, _ = s.force_kernel_stats()
        rb_ms, rb_gap_ms, rb_n = s.rebuild_stats()
        rate = w.n * steps / wall
        kbar = float(np.mean(pairs)) / w.n
        cbar = float(np.mean(contacts)) / w.n
        state = 216.0 if prec == hip.MIXED else 264.0
        bps = state + 4.0 * kbar + 56.0 * cbar
        out[name] = dict(value=round(rate, 1), step_roofline_frac=round(rate * bps / 8e12, 4),
                         bytes_per_particle_step=round(bps, 1),
                         ms_per_step=round(1e3 * wall / steps, 4),
                         force_kernel_us=round(ms * 1e3, 2), k_half_pairs=round(kbar, 3),
                         c_contacts=round(cbar, 3), npairs=th1["npairs"],
                         ncontacts=th1["ncontacts"],
                         rebuilds=th1["rebuilds"] - th0["rebuilds"],
                         rebuild_us=round(rb_ms * 1e3, 1),
                         idle_before_rebuild_us=round(rb_gap_ms * 1e3, 1))
        s.close()
        log(f"[bench] dem {name}: {out[name]['value']:.4g} particle-steps/s, "
            f"{out[name]['ms_per_step']} ms/step, k {kbar:.3f} c {cbar:.3f}, "
            f"{out[name]['rebuilds']} rebuilds ({rb_ms * 1e3:.0f} us each), "
            f"force pass {ms * 1e3:.1f} us")
    out["mixed_over_fp64"] = round(out["mixed"]["value"] / out["fp64"]["value"], 3)
    out["force_pass_shape"] = shape
    out["workload"] = (f"DEM settling spheres as a bed at rest: {w.n} glass spheres "
                       "r~U[0.8,1.2] mm deposited on z=0 (each resting on the spheres below, "
                       "0.1 nm contact overlaps), box 0.4x0.4x1.2 m periodic x/y, plane walls "
                       "z=0 and z=1.2 m, gravity, dt=0.05 t_Rayleigh(0.8 mm), "
                       f"[neighbor] every={every} check=true")
    out["steps"] = steps
    out["warmup"] = warmup
    out["steps_before_warmup"] = pre
    return out, w


def force_kernel_bytes(n, npairs, force_bytes, pos_bytes=24.0):
    """Algorithmic HBM bytes of one pair-force launch: positions read once
    (24 B FP64 x, y, z; 16 B for the mixed pass's fixed-point int4), the
    half list streamed once (4 B per pair) and the force written once
    (3 x 4 B mixed, 3 x 8 B FP64)."""
    return n * pos_bytes + 4.0 * npairs + n * 3.0 * force_bytes


def _cpu_sample(args):
    """One replica of the CPU checker: `seconds` of scheduled steps after one
    untimed set-up step; returns (steps, wall seconds)."""
    w, seconds = args
    from tests.helpers import oracle_from_workload  # the checker; CPU leg only
    o = oracle_from_workload(w)
    o.step()
    steps = 0
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < seconds:
        o.run(5)
        steps += 5
    return steps, time.perf_counter() - t0


def cpu_baseline(w, seconds_budget=12.0, replicas=None):
    """The CPU checker (a C port of the reference path) on the host: 1 thread
    over a bounded number of steps of the same workload, then `replicas`
    independent copies at once (one process each), the embarrassingly
    parallel upper bound of the host. The checker keeps the reference's
    rebuild rule (decide_rebuild: also when any atom leaves the periodic
    box), which at this size rebuilds nearly every step."""
    import multiprocessing as mp
    nproc = os.cpu_count() or 1
    steps, dt = _cpu_sample((w, seconds_budget))
    out = dict(value=w.n * steps / dt, unit="particle-steps/s", cores=1, kind="port",
               sample=f"{steps} scheduled steps (after 1 untimed setup step) of the same "
                      f"{w.n}-atom workload on 1 host thread; rebuilds as the reference "
                      "decides them (skin spent or an atom outside the periodic box)",
               nproc=nproc)
    if replicas is None:
        replicas = min(nproc, 16)
    if replicas > 1:
        ctx = mp.get_context("fork")
        with ctx.Pool(replicas) as pool:
            res = pool.map(_cpu_sample, [(w, seconds_budget / 2)] * replicas)
        out["replicas"] = dict(value=sum(w.n * st / t for st, t in res), cores=replicas,
                               note="independent copies at once, one process each: an "
                                    "embarrassingly parallel upper bound")
    return out


def cpu_baseline_dem(w, seconds_budget=10.0):
    """The CPU checker on the DEM line's workload (the full bed, same
    neighbor settings) on 1 host thread: set-up and the first rebuild
    untimed, then whole steps until the budget is spent."""
    from tests.helpers import oracle_from_workload  # the checker; CPU leg only
    o = oracle_from_workload(w)
    o.step()
    steps = 0
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < seconds_budget:
        o.run(1)
        steps += 1
    dt = time.perf_counter() - t0
    return dict(value=w.n * steps / dt, unit="particle-steps/s", cores=1, kind="port",
                sample=f"{steps} scheduled steps (after 1 untimed setup step) of the same "
                       f"{w.n}-sphere bed on 1 host thread")


def traffic_from_profiles(name, kernel):
    """Calibrated HBM bytes per launch of `kernel` from the committed counter
    summary for this workload, or None when it covers another kernel."""
    path = os.path.join(ROOT, "profiles", f"traffic_{name}.json")
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        d = json.load(fh)
    return d.get("hbm_bytes_per_launch") if d.get("kernel") == kernel else None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20000)
    ap.add_argument("--warmup", type=int, default=1000)
    ap.add_argument("--ncell", type=int, default=40)
    ap.add_argument("--precision", choices=["mixed", "fp64"], default="mixed")
    ap.add_argument("--fp64-steps", type=int, default=-1,
                    help="timed steps of the all-FP64 leg (-1: the same as --steps)")
    ap.add_argument("--graphs", choices=["on", "off"], default="on",
                    help="issue neighbor rebuilds as captured HIP graphs")
    ap.add_argument("--layout", choices=["auto", "tile", "1", "2", "3"], default="2",
                    help="pair-force layout: tuned during warmup, cell tiles, or rows at "
                         "2^k lanes per particle (default 4)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--cpu-replicas", type=int, default=-1,
                    help="concurrent CPU checker copies (-1: min(nproc, 16); 0/1: none)")
    ap.add_argument("--precision-report", type=int, default=20000,
                    help="NVE steps of the mixed-vs-FP64 precision report (0 = skip)")
    ap.add_argument("--ensemble-systems", type=int, default=1024,
                    help="systems in the batched-ensemble measurement (0 = skip)")
    ap.add_argument("--ensemble-steps", type=int, default=5000)
    ap.add_argument("--serve-jobs", type=int, default=24,
                    help="jobs in the serving-stream measurement (0 = skip)")
    ap.add_argument("--dem-steps", type=int, default=3000,
                    help="timed steps of the DEM 4M line per precision (0 = skip)")
    ap.add_argument("--dem-warmup", type=int, default=200)
    ap.add_argument("--dem-tune-steps", type=int, default=200,
                    help="steps timed per launch shape of the DEM force pass")
    ap.add_argument("--dem-every", type=int, default=500,
                    help="[neighbor] every of the DEM line (rebuilds in the timed window)")
    ap.add_argument("--serve-steps", type=int, nargs=2, default=(1000, 50000),
                    help="step range of serving jobs (log-uniform)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); multi-GPU scaling is not "
              "part of this build", file=sys.stderr)
        sys.exit(2)

    from mddem_amd import hip
    from mddem_amd import workloads as W

    w = W.lj_fcc(ncell=args.ncell)
    prec = hip.MIXED if args.precision == "mixed" else hip.FP64
    log(f"[bench] workload {w.name} n={w.n} precision={args.precision} "
        f"steps={args.steps} warmup={args.warmup}")
    layout = None if args.layout == "auto" else (-1 if args.layout == "tile" else int(args.layout))
    graphs = args.graphs == "on"
    r = run_gpu(w, prec, args.steps, args.warmup, layout=layout, graphs=graphs)
    value = r["n"] * r["steps"] / r["seconds"]
    log(f"[bench] {value:.4g} particle-steps/s, {r['rebuilds']} rebuilds, "
        f"{r['layout']}, force kernel {r['force_ms'] * 1e3:.1f} us, "
        f"rebuild {r['rebuild_ms'] * 1e3:.1f} us + {r['rebuild_gap_ms'] * 1e3:.1f} us idle before "
        f"({r['rebuilds_timed']} timed)")

    fp64 = None
    r64 = None
    fp64_steps = args.steps if args.fp64_steps < 0 else args.fp64_steps
    if prec == hip.MIXED and fp64_steps > 0:
        r64 = run_gpu(w, hip.FP64, fp64_steps, args.warmup, layout=layout, graphs=graphs)
        fp64 = r64["n"] * r64["steps"] / r64["seconds"]
        log(f"[bench] fp64 {fp64:.4g} particle-steps/s, {r64['rebuilds']} rebuilds, force kernel "
            f"{r64['force_ms'] * 1e3:.1f} us")

    fbytes = 4 if prec == hip.MIXED else 8
    # the mixed row pass gathers fixed-point positions (mdd_set_lj_positions)
    pbytes = 16.0 if prec == hip.MIXED and not r["tile"] else 24.0
    alg = force_kernel_bytes(r["n"], r["npairs"], fbytes, pbytes)
    if r["fused"]:
        # the final half step in the same launch: velocity read and written,
        # mass read (the force it kicks with is the one the pass computed)
        alg += r["n"] * (32.0 + 32.0 + 8.0)
    achieved = alg / (r["force_ms"] * 1e-3) / 1e9 if r["force_ms"] > 0 else 0.0
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": traffic_from_profiles(w.name, r["kernel_name"]),
                "kernel": ("k_lj_tile (pair force + final half step)" if r["tile"]
                           else "k_lj_post (pair force + final half step)" if r["fused"]
                           else "k_lj (pair force)"), "bytes_per_launch": alg,
                "avg_launch_us": round(r["force_ms"] * 1e3, 2)}

    precision = None
    if args.precision_report > 0:
        from mddem_amd import precision as P
        precision = dict(force_error=P.force_errors(w),
                         nve=P.nve_drift(w, args.precision_report))
        log(f"[bench] precision: {json.dumps(precision)}")

    dem = None
    if args.dem_steps > 0:
        dem, wdem = run_dem(args.dem_steps, args.dem_warmup, log_every=500,
                            every=args.dem_every, tune_steps=args.dem_tune_steps)
        if args.precision_report > 0:
            from mddem_amd import precision as P
            dem["precision"] = dict(force_error=P.dem_force_errors(wdem))
            log(f"[bench] dem precision: {json.dumps(dem['precision'])}")
        if not args.no_cpu_baseline:
            log("[bench] dem cpu baseline ...")
            dem["cpu_baseline"] = cpu_baseline_dem(wdem, args.cpu_seconds)
            dem["gpu_over_cpu"] = round(dem["mixed"]["value"] / dem["cpu_baseline"]["value"], 1)
            log(f"[bench] dem cpu {dem['cpu_baseline']['value']:.4g} particle-steps/s")
        del wdem

    ens = None
    if args.ensemble_systems > 0:
        from mddem_amd import ensemble
        ens = ensemble.run(args.ensemble_systems, args.ensemble_steps, log=log)
        log(f"[bench] ensemble: {json.dumps(ens)}")

    serving = None
    if args.serve_jobs > 0:
        from mddem_amd import serve
        lo_s, hi_s = args.serve_steps
        jobs = serve.make_stream(args.serve_jobs, rate=30.0, steps=(lo_s, hi_s))
        for j in jobs:
            j.workload = serve.build_workload(j)
        serving = serve.Server(serve.default_budget(), slots=8).run(jobs)
        serving["stream"] = (f"{args.serve_jobs} jobs, Poisson 30/s, 70% granular gas 1k-200k, "
                             f"30% LJ FCC 4k-500k, steps log-uniform [{lo_s}, {hi_s}], mixed")
        log(f"[bench] serving: {json.dumps(serving)}")

    cpu = None
    if not args.no_cpu_baseline:
        log("[bench] cpu baseline ...")
        cpu = cpu_baseline(w, args.cpu_seconds,
                           None if args.cpu_replicas < 0 else args.cpu_replicas)
        log(f"[bench] cpu {cpu['value']:.4g} particle-steps/s")

    out = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * r["seconds"] / r["steps"], 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32-force/fp64-state" if prec == hip.MIXED else "fp64",
        "data": "synthetic (seeded FCC lattice, Maxwell-Boltzmann velocities)",
        "config": {"workload": f"LJ fluid FCC {args.ncell}^3 = {w.n} atoms, rho*=0.85, "
                               "T*=0.85 Nose-Hoover tau=1, dt=0.005, rc=2.5, skin 2.8",
                   "precision": args.precision, "rebuild": "displacement (lazy wrap)",
                   "rebuilds_in_timed_loop": r["rebuilds"],
                   "force_layout": r["layout"],
                   "rebuild_us": round(r["rebuild_ms"] * 1e3, 1),
                   "idle_before_rebuild_us": round(r["rebuild_gap_ms"] * 1e3, 1),
                   "setup_steps_untimed": r["setup_steps"],
                   "pairs_per_atom_half": round(r["npairs"] / r["n"], 3)},
        "fp64_gpu_value": None if fp64 is None else round(fp64, 1),
        "fp64_gpu": None if r64 is None else {
            "steps": r64["steps"], "warmup": args.warmup, "rebuilds_in_timed_loop": r64["rebuilds"],
            "force_kernel_us": round(r64["force_ms"] * 1e3, 2),
            "rebuild_us": round(r64["rebuild_ms"] * 1e3, 1)},
        "mixed_over_fp64": None if fp64 is None else round(value / fp64, 3),
        "roofline": roofline,
        "cpu_baseline": cpu,
        "gpu_over_cpu": None if cpu is None else round(value / cpu["value"], 1),
        # whole-step algorithmic bytes per particle-step: 120 B (mixed) or 144 B
        # (FP64) of state + 4 B per half-list pair, against the 8 TB/s HBM roofline
        "step_roofline_frac": round(
            value * ((120.0 if prec == hip.MIXED else 144.0) + 4.0 * r["npairs"] / r["n"]) / 8e12,
            4),
        "precision": precision,
        "dem_4m": dem,
        "ensemble": ens,
        "serving": serving,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
