"""Batched granular-gas ensemble: many independent periodic DEM systems of
uneven size, stepped together on one GPU.

Each size class of the ensemble is one batched context: its systems'
particles and bins are offset-indexed in shared arrays, and every phase of a
step is one launch over all of them. The autotuner picks the pair-force
launch shape (lanes per sphere) and the bin size of each class from short
timed trials on a copy of the class. Classes run concurrently, each on its
own stream and host thread.

Reported: particle-steps/s of the whole ensemble, the p50/p99 latency of a
step (over timed segments of the run) and of system completion, and the
granular-temperature decay of every system (Haff cooling).

    python -m mddem_amd.ensemble [--systems 1024] [--steps 5000]
"""
from __future__ import annotations

import argparse
import json
import sys
import threading
import time

import numpy as np

from mddem_amd import hip
from mddem_amd import workloads as W

CLASSES = ((200, 1000), (1000, 5000), (5000, 20001))
# force pass shapes: (log2 lanes per sphere, contacts queued per wavefront);
# the queue only applies at one lane per sphere
SHAPES = ((0, True), (0, False), (1, False), (2, False))
BIN_FACTORS = (1.0, 1.5)       # bin size / pair cutoff


def make_context(systems, precision=hip.MIXED, box_exit=False, bin_size=None,
                 lanes_log2=None, wave=None) -> hip.System:
    """One batched context holding `systems` (workloads.gas_system)."""
    w0 = systems[0]
    s = hip.System(hip.MODEL_DEM, precision)
    s.set_neighbor(w0.skin_fraction, w0.bin_size if bin_size is None else bin_size, 0, True,
                   1000, True)
    s.set_rebuild_on_box_exit(box_exit)
    s.set_dt(w0.dt)
    if lanes_log2 is not None:
        s.set_force_lanes(lanes_log2)
    if wave is not None:
        s.set_dem_wave(wave)
    for w in systems:
        m = w.materials
        s.add_system(w.lo, w.hi, m["youngs"][0], m["poisson"][0], m["restitution"][0],
                     m["friction"][0], w.pos, w.vel, w.mass, w.cutoff_radius, w.radius,
                     omega=w.omega, tag=w.tag)
    return s


def system_states(s, nsys):
    """Per system: (tag-ordered positions, velocities)."""
    st = s.state()
    out = []
    for k in range(nsys):
        f, c = s.system_range(k)
        tag = st["tag"][f:f + c]
        order = np.argsort(tag)
        out.append((st["pos"][f:f + c][order], st["vel"][f:f + c][order]))
    return out


def granular_temperature(vel, mass):
    vbar = (mass[:, None] * vel).sum(0) / mass.sum()
    u = vel - vbar
    return float((mass * (u * u).sum(1)).sum() / (3.0 * mass.sum()))


def tune_class(systems, precision, trial_steps=60):
    """Force pass shape and bin factor with the fastest trial step."""
    cut = 2.0 * float(systems[0].cutoff_radius[0]) * systems[0].skin_fraction
    best = None
    for bf in BIN_FACTORS:
        for lg, wave in SHAPES:
            s = make_context(systems, precision, bin_size=bf * cut, lanes_log2=lg, wave=wave)
            s.run(10)
            s.synchronize()
            t0 = time.perf_counter()
            s.run(trial_steps)
            s.synchronize()
            dt = time.perf_counter() - t0
            s.close()
            if best is None or dt < best[0]:
                best = (dt, lg, wave, bf)
    return dict(lanes_log2=best[1], wave=best[2], bin_factor=best[3], trial_s=best[0])


def run(nsys=1024, steps=5000, segment=100, precision=hip.MIXED, tune=True, seed=None, log=None):
    systems = W.ensemble(nsys, seed=seed)
    classes = []
    for lo, hi in CLASSES:
        idx = [k for k, w in enumerate(systems) if lo <= w.n < hi]
        if idx:
            classes.append(idx)
    plan = []
    for idx in classes:
        sub = [systems[k] for k in idx]
        cfg = tune_class(sub, precision) if tune else dict(lanes_log2=2, wave=False,
                                                            bin_factor=1.0)
        cut = 2.0 * float(sub[0].cutoff_radius[0]) * sub[0].skin_fraction
        plan.append((idx, sub, cfg, cut))
        if log:
            log(f"[ensemble] class n in [{min(w.n for w in sub)}, {max(w.n for w in sub)}]: "
                f"{len(sub)} systems, {sum(w.n for w in sub)} spheres, tuned {cfg}")
    ctxs = [make_context(sub, precision, bin_size=cfg["bin_factor"] * cut,
                         lanes_log2=cfg["lanes_log2"], wave=cfg["wave"])
            for idx, sub, cfg, cut in plan]
    for s in ctxs:
        s.run(1)  # setup and first rebuild, untimed
        s.synchronize()
    t_g0 = [granular_temperature(w.vel, w.mass) for w in systems]
    seg_times = [[] for _ in ctxs]
    done_at = [0.0 for _ in ctxs]
    errors = []
    barrier = threading.Barrier(len(ctxs) + 1)
    t0_box = [0.0]

    def drive(ci):
        s = ctxs[ci]
        try:
            barrier.wait()
            left = steps
            while left > 0:
                k = min(segment, left)
                a = time.perf_counter()
                s.run(k)
                s.synchronize()
                seg_times[ci].append((time.perf_counter() - a) / k)
                left -= k
            done_at[ci] = time.perf_counter() - t0_box[0]
        except Exception as e:  # surfaced below
            errors.append(f"{type(e).__name__}: {e}")

    threads = [threading.Thread(target=drive, args=(ci,)) for ci in range(len(ctxs))]
    for th in threads:
        th.start()
    t0_box[0] = time.perf_counter()
    barrier.wait()
    for th in threads:
        th.join()
    wall = time.perf_counter() - t0_box[0]
    if errors:
        raise RuntimeError("; ".join(errors))
    # per-system results
    t_g1 = [0.0] * len(systems)
    sys_done = [0.0] * len(systems)
    for ci, (idx, sub, cfg, cut) in enumerate(plan):
        states = system_states(ctxs[ci], len(sub))
        for j, k in enumerate(idx):
            t_g1[k] = granular_temperature(states[j][1], systems[k].mass)
            sys_done[k] = done_at[ci]
    for s in ctxs:
        s.close()
    total = sum(w.n for w in systems)
    step_lat = np.concatenate([np.asarray(t) for t in seg_times])
    ratio = np.asarray(t_g1) / np.asarray(t_g0)
    return dict(
        systems=nsys, spheres=total, steps=steps, precision="mixed" if precision == hip.MIXED
        else "fp64", value=round(total * steps / wall, 1), unit="particle-steps/s",
        wall_s=round(wall, 4),
        step_latency_p50_ms=round(1e3 * float(np.percentile(step_lat, 50)), 4),
        step_latency_p99_ms=round(1e3 * float(np.percentile(step_lat, 99)), 4),
        completion_p50_s=round(float(np.percentile(sys_done, 50)), 4),
        completion_p99_s=round(float(np.percentile(sys_done, 99)), 4),
        classes=[dict(n_range=[min(systems[k].n for k in idx), max(systems[k].n for k in idx)],
                      systems=len(idx), **cfg) for idx, sub, cfg, cut in plan],
        haff_cooled_fraction=round(float(np.mean(ratio < 1.0 - 1e-12)), 4),
        haff_tg_ratio_median=round(float(np.median(ratio)), 6),
        haff_tg_ratio_max=float(np.max(ratio)))


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--systems", type=int, default=1024)
    ap.add_argument("--steps", type=int, default=5000)
    ap.add_argument("--fp64", action="store_true")
    args = ap.parse_args(argv)
    rep = run(args.systems, args.steps, precision=hip.FP64 if args.fp64 else hip.MIXED,
              log=lambda m: print(m, file=sys.stderr, flush=True))
    print(json.dumps(rep), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
