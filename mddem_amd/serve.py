"""Serving loop: simulation jobs of mixed size on one GPU, continuously batched.

Jobs arrive over time (Poisson, seeded). A job is admitted as soon as a slot
is free and its estimated device footprint fits the remaining HBM budget, in
arrival order. An admitted job gets its own context (device-resident state,
its own HIP stream) and a host thread that drives its steps, so the device
work of all live jobs overlaps. A job's state never leaves the device until
it completes. Each job's result does not depend on what else runs: contexts
share nothing and every kernel is deterministic, so a job's final state is
bitwise the same alone and in the mix.

Reported: throughput (particle-steps/s over the makespan), and p50/p99 job
latency (completion minus arrival).

    python -m mddem_amd.serve [--jobs 32] [--rate 8] [--max-steps 5000]
"""
from __future__ import annotations

import argparse
import hashlib
import json
import math
import sys
import threading
import time
from dataclasses import dataclass, field

import numpy as np

from mddem_amd import workloads as W

# device bytes per particle (state, forces, alternates, neighbor rows,
# shear history, bins) used to admit jobs; measured allocations stay below
FOOTPRINT_PER_PARTICLE = {"dem": 3200, "lj": 1400}
FOOTPRINT_FIXED = 64 << 20


@dataclass
class Job:
    jid: int
    kind: str            # "dem" (granular gas) or "lj" (FCC fluid)
    n_target: int
    steps: int
    arrival: float       # seconds after the stream starts
    params: dict = field(default_factory=dict)
    workload: object = None
    admitted: float | None = None
    done: float | None = None
    digest: dict | None = None
    error: str | None = None
    final: dict | None = None

    @property
    def n(self) -> int:
        return self.workload.n if self.workload is not None else self.n_target

    @property
    def footprint(self) -> int:
        return FOOTPRINT_FIXED + FOOTPRINT_PER_PARTICLE[self.kind] * self.n


def make_stream(njobs=32, rate=8.0, seed=0x4D444445_4D04, steps=(1000, 50000),
                dem_n=(1000, 200_000), lj_n=(4000, 500_000), dem_share=0.7):
    """The mixed job stream: 70% granular gas of 1k-200k spheres, 30% LJ FCC
    of 4k-500k atoms at density 0.8-0.9 and T 0.7-1.2, step counts
    log-uniform, Poisson arrivals at `rate` jobs/s."""
    st = W.Stream(seed)
    u = st.uniform(6 * njobs).reshape(njobs, 6)
    t = 0.0
    jobs = []
    for k in range(njobs):
        t += -math.log(1.0 - u[k, 0]) / rate
        ns = int(round(math.exp(math.log(steps[0]) + u[k, 1] * math.log(steps[1] / steps[0]))))
        if u[k, 2] < dem_share:
            n = int(dem_n[0] + u[k, 3] * (dem_n[1] - dem_n[0]))
            jobs.append(Job(k, "dem", n, ns, t, dict(seed=W.SEED_BASE + 100 + k)))
        else:
            n = int(lj_n[0] + u[k, 3] * (lj_n[1] - lj_n[0]))
            jobs.append(Job(k, "lj", n, ns, t,
                            dict(rho=0.8 + 0.1 * u[k, 4], temperature=0.7 + 0.5 * u[k, 5],
                                 seed=W.SEED_BASE + 100 + k)))
    return jobs


def build_workload(job: Job):
    """The request payload: the job's seeded initial state."""
    if job.kind == "dem":
        box = 0.025 * (job.n_target / 500.0) ** (1.0 / 3.0)  # volume fraction 0.134
        return W.dem_gas(n=job.n_target, box=box, seed=job.params["seed"], steps=job.steps)
    ncell = max(2, int(round((job.n_target / 4.0) ** (1.0 / 3.0))))
    return W.lj_fcc(ncell=ncell, rho=job.params["rho"], temperature=job.params["temperature"],
                    seed=job.params["seed"], steps=job.steps)


def state_digest(s) -> dict:
    """Final state in tag order: SHA-256 of positions and velocities, and KE."""
    st = s.state()
    order = np.argsort(st["tag"])
    h = hashlib.sha256()
    h.update(np.ascontiguousarray(st["pos"][order]).tobytes())
    h.update(np.ascontiguousarray(st["vel"][order]).tobytes())
    th = s.thermo()
    replays, captures = s.launch_graph_stats()
    return dict(sha256=h.hexdigest(), ke=th["ke"], rebuilds=th["rebuilds"], step=th["step"],
                graph_replays=replays, graph_captures=captures)


def run_job(job: Job, precision, box_exit=False, keep_state=False) -> dict:
    """One job on its own context; returns its digest (and, with keep_state,
    its final tags, positions and velocities in the job's `final`)."""
    from mddem_amd import hip
    s = hip.from_workload(job.workload, precision, box_exit=box_exit)
    try:
        s.run(job.steps)
        if keep_state:
            st = s.state()
            job.final = dict(tag=st["tag"], pos=st["pos"], vel=st["vel"],
                             rebuilds=s.thermo()["rebuilds"])
        return state_digest(s)
    finally:
        s.close()


def admit(queue, live_bytes, live_count, budget, slots):
    """Jobs to start now: in arrival order, while a slot is free and the
    job's footprint fits the remaining budget (FIFO, no overtaking)."""
    out = []
    while queue and live_count + len(out) < slots:
        need = queue[0].footprint
        if live_bytes + need > budget:
            break
        live_bytes += need
        out.append(queue.pop(0))
    return out


class Server:
    """box_exit: rebuild when a particle leaves the periodic box (the
    reference's rule) instead of wrapping lazily; keep_state: keep each
    job's final state on the job (for comparisons)."""

    def __init__(self, hbm_budget: int, slots: int = 8, precision=None, box_exit=False,
                 keep_state=False):
        from mddem_amd import hip
        self.budget = int(hbm_budget)
        self.slots = int(slots)
        self.precision = hip.MIXED if precision is None else precision
        self.box_exit = box_exit
        self.keep_state = keep_state
        self.lock = threading.Condition()
        self.live_bytes = 0
        self.live = 0
        self.max_live = 0

    def _worker(self, job: Job, t0: float):
        try:
            job.digest = run_job(job, self.precision, self.box_exit, self.keep_state)
        except Exception as e:  # reported per job; the loop goes on
            job.error = f"{type(e).__name__}: {e}"
        job.done = time.perf_counter() - t0
        with self.lock:
            self.live_bytes -= job.footprint
            self.live -= 1
            self.lock.notify_all()

    def run(self, jobs):
        """Serve `jobs` (payloads already built) with their arrival times."""
        for j in jobs:
            if j.footprint > self.budget:
                raise ValueError(f"job {j.jid} needs {j.footprint} B, budget {self.budget} B")
        pending = sorted(jobs, key=lambda j: j.arrival)
        queue = []
        threads = []
        t0 = time.perf_counter()
        with self.lock:
            while pending or queue or self.live:
                now = time.perf_counter() - t0
                while pending and pending[0].arrival <= now:
                    queue.append(pending.pop(0))
                for j in admit(queue, self.live_bytes, self.live, self.budget, self.slots):
                    j.admitted = now
                    self.live_bytes += j.footprint
                    self.live += 1
                    self.max_live = max(self.max_live, self.live)
                    th = threading.Thread(target=self._worker, args=(j, t0), daemon=True)
                    th.start()
                    threads.append(th)
                wait = None
                if pending:
                    wait = max(0.0, pending[0].arrival - (time.perf_counter() - t0))
                if not pending and not queue and not self.live:
                    break
                self.lock.wait(timeout=wait if wait is not None else 1.0)
        for th in threads:
            th.join()
        makespan = max(j.done for j in jobs) - min(j.arrival for j in jobs)
        return report(jobs, makespan, self.max_live)


def report(jobs, makespan, max_live):
    lat = np.array([j.done - j.arrival for j in jobs])
    work = float(sum(j.n * j.steps for j in jobs))
    return dict(jobs=len(jobs), failed=sum(1 for j in jobs if j.error),
                particle_steps=work, makespan_s=round(makespan, 4),
                throughput=round(work / makespan, 1),
                latency_p50_s=round(float(np.percentile(lat, 50)), 4),
                latency_p99_s=round(float(np.percentile(lat, 99)), 4),
                latency_max_s=round(float(lat.max()), 4),
                queue_wait_p99_s=round(float(np.percentile(
                    [j.admitted - j.arrival for j in jobs], 99)), 4),
                max_concurrent=max_live)


def default_budget(fraction=0.5):
    from mddem_amd import hip
    free, total = hip.device_memory(0)
    return int(min(free, total) * fraction)


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=32)
    ap.add_argument("--rate", type=float, default=8.0, help="arrivals per second")
    ap.add_argument("--min-steps", type=int, default=1000)
    ap.add_argument("--max-steps", type=int, default=5000)
    ap.add_argument("--slots", type=int, default=8)
    ap.add_argument("--budget-gib", type=float, default=0.0, help="0: half of free HBM")
    args = ap.parse_args(argv)
    jobs = make_stream(args.jobs, args.rate, steps=(args.min_steps, args.max_steps))
    for j in jobs:
        j.workload = build_workload(j)
    budget = int(args.budget_gib * 2**30) if args.budget_gib > 0 else default_budget()
    rep = Server(budget, args.slots).run(jobs)
    print(json.dumps(rep), flush=True)
    return 0 if rep["failed"] == 0 else 1


if __name__ == "__main__":
    sys.exit(main())
