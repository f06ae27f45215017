"""Precision report of the mixed path against the all-FP64 GPU path.

Force error, from identical state: e_F = ||F_mixed - F_fp64||_2 / ||F_fp64||_2
over all particles, plus the largest per-particle error over the RMS force.
The states are those of the all-FP64 trajectory at the given steps after an
equilibration run: the seeded FCC start has zero net force on every atom by
symmetry, so a relative force error is only meaningful once the lattice has
melted.

NVE drift: velocity Verlet without the thermostat, E = KE + PE (LJ 12-6,
unshifted), from an equilibrated state (positions, velocities and forces of
the FP64 run after 1000 Nose-Hoover steps). The drift is |E(t) - E(0)| /
|E(0)| over the run, reported as the slope of a least-squares line through
the sampled points and as the largest deviation. The mixed run must stay
within 2x the FP64 run.

DEM (dem_force_errors): e_F and e_T of the contact pass alone; granular
contacts dissipate energy, so no NVE drift is defined there.

    python -m mddem_amd.precision [--ncell 40] [--nve-steps 2000]
"""
from __future__ import annotations

import argparse
import json
import sys

import numpy as np

from mddem_amd import hip
from mddem_amd import workloads as W


def _state_system(w, pos, vel, precision):
    s = hip.from_workload(w, precision, box_exit=False)
    s.upload(pos, vel, w.mass, w.cutoff_radius, radius=w.radius, omega=w.omega, tag=w.tag,
             atype=w.atype)
    return s


def _forces_by_tag(s):
    st = s.state()
    f, t = s.forces()
    order = np.argsort(st["tag"])
    return f[order], (None if t is None else t[order])


def force_errors(w, checkpoints=(0, 1, 10, 1000), equilibrate=1000):
    """e_F of the mixed pair forces at the FP64 trajectory's states, steps
    counted from the end of `equilibrate` thermostatted steps."""
    ref = hip.from_workload(w, hip.FP64, box_exit=False)
    ref.run(equilibrate)
    out = []
    done = 0
    for k in checkpoints:
        if k > done:
            ref.run(k - done)
            done = k
        st = ref.state()
        order = np.argsort(st["tag"])
        pos, vel = st["pos"][order], st["vel"][order]
        res = {}
        for name, prec in (("fp64", hip.FP64), ("mixed", hip.MIXED)):
            s = _state_system(w, pos, vel, prec)
            s.bin_neighbor_list()
            if w.model == "lj":
                s.lj_force()
            else:
                s.hertz_mindlin_contact_force()
            res[name] = _forces_by_tag(s)
            s.close()
        f64, fm = res["fp64"][0], res["mixed"][0]
        d = fm - f64
        frms = np.sqrt(np.mean(np.sum(f64 * f64, axis=1)))
        row = dict(step=equilibrate + k, e_F=float(np.linalg.norm(d) / np.linalg.norm(f64)),
                   max_over_rms=float(np.max(np.linalg.norm(d, axis=1)) / frms))
        if res["fp64"][1] is not None:
            t64, tm = res["fp64"][1], res["mixed"][1]
            if np.linalg.norm(t64) > 0:
                row["e_T"] = float(np.linalg.norm(tm - t64) / np.linalg.norm(t64))
        out.append(row)
    ref.close()
    return out


def nve_drift(w, steps=20000, samples=20, equilibrate=1000):
    """Energy drift of velocity Verlet (thermostat off) in both precisions,
    from one equilibrated state: `equilibrate` thermostatted FP64 steps melt
    the lattice first (from the lattice the energy moves by the melting and
    by pairs crossing the unshifted cutoff, which no precision changes).
    Both runs start from that state's positions, velocities and forces."""
    ref = hip.from_workload(w, hip.FP64, box_exit=False)
    ref.run(equilibrate)
    st = ref.state()
    f, _ = ref.forces()
    order = np.argsort(st["tag"])
    pos, vel, force = st["pos"][order], st["vel"][order], f[order]
    ref.close()
    res = {}
    for name, prec in (("fp64", hip.FP64), ("mixed", hip.MIXED)):
        s = _state_system(w, pos, vel, prec)
        s.set_nose_hoover(0.0, 1.0, enabled=False)
        s.set_integrators(True, False)
        s.set_forces(force)
        s.bin_neighbor_list()  # the energy pass needs the list (the forces move with the sort)
        t, e = [], []
        per = max(1, steps // samples)
        for k in range(samples + 1):
            if k:
                s.run(per)
            th = s.thermo(energy=True)
            t.append(th["step"])
            e.append(th["ke"] + th["pe"])
        s.close()
        e = np.asarray(e)
        rel = np.abs(e - e[0]) / abs(e[0])
        slope = float(np.polyfit(np.asarray(t, dtype=float), (e - e[0]) / abs(e[0]), 1)[0])
        res[name] = dict(E0=float(e[0]), max_rel_drift=float(rel.max()),
                         drift_per_step=slope, steps=int(t[-1]))
    res["mixed_over_fp64"] = (res["mixed"]["max_rel_drift"] /
                              max(res["fp64"]["max_rel_drift"], 1e-300))
    res["slope_mixed_over_fp64"] = (abs(res["mixed"]["drift_per_step"]) /
                                    max(abs(res["fp64"]["drift_per_step"]), 1e-300))
    res["start"] = f"after {equilibrate} Nose-Hoover FP64 steps from the lattice"
    return res


def dem_force_errors(w, checkpoints=(0, 1, 10, 1000)):
    """e_F and e_T of the mixed contact pass against the FP64 one at states of
    the FP64 trajectory (steps counted from the workload's start). Both
    passes start from the same positions, velocities and spins with no
    stored springs (an upload carries no contact history), so the
    tangential force is that of a contact's first step."""
    ref = hip.from_workload(w, hip.FP64, box_exit=False)
    out = []
    done = 0
    for k in checkpoints:
        if k > done:
            ref.run(k - done)
            done = k
        st = ref.state()
        order = np.argsort(st["tag"])
        pos, vel, om = st["pos"][order], st["vel"][order], st["omega"][order]
        res = {}
        for name, prec in (("fp64", hip.FP64), ("mixed", hip.MIXED)):
            s = hip.from_workload(w, prec, box_exit=False)
            s.upload(pos, vel, w.mass, w.cutoff_radius, radius=w.radius, omega=om, tag=w.tag,
                     atype=w.atype)
            s.bin_neighbor_list()
            s.hertz_mindlin_contact_force()
            res[name] = _forces_by_tag(s)
            s.close()
        (f64, t64), (fm, tm) = res["fp64"], res["mixed"]
        d = fm - f64
        frms = np.sqrt(np.mean(np.sum(f64 * f64, axis=1)))
        trms = np.sqrt(np.mean(np.sum(t64 * t64, axis=1)))
        row = dict(step=k, e_F=float(np.linalg.norm(d) / np.linalg.norm(f64)),
                   max_over_rms=float(np.max(np.linalg.norm(d, axis=1)) / frms))
        # no torque at all (no spin and no stored springs yet): e_T undefined
        if trms > 0.0:
            row["e_T"] = float(np.linalg.norm(tm - t64) / np.linalg.norm(t64))
            row["torque_max_over_rms"] = float(np.max(np.linalg.norm(tm - t64, axis=1)) / trms)
        else:
            row["e_T"] = None
        out.append(row)
    ref.close()
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--ncell", type=int, default=40)
    ap.add_argument("--nve-steps", type=int, default=20000)
    args = ap.parse_args(argv)
    w = W.lj_fcc(ncell=args.ncell)
    rep = dict(workload=w.name, force_error=force_errors(w), nve=nve_drift(w, args.nve_steps))
    print(json.dumps(rep), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
