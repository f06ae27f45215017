"""Writes the granular-gas fixture tests/golden/cfg0_state.csv and
tests/golden/cfg0_step10.npz.

The state is the 500-sphere periodic granular gas (glass, r = 1 mm, 25 mm box,
0.5 m/s) after 20,000 steps of the CPU checker, so that spheres are in
contact; the CSV holds the columns the reference's CSV ingest reads
(crates/dem_atom/src/insert.rs:652-776: x, y, z, r, vx, vy, vz) plus tag and
spin. The npz holds what 10 further FP64 steps from that state (no history
carried in) produce: forces and torques by tag, the half-list pairs as
canonical keys, and the shear history.

Run: python tests/golden/make_cfg0.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from mddem_amd import workloads as W  # noqa: E402
from tests.helpers import canonical_pairs, oracle_from_workload  # noqa: E402

COLUMNS = "tag,x,y,z,r,vx,vy,vz,wx,wy,wz"


def load_state():
    a = np.loadtxt(os.path.join(HERE, "cfg0_state.csv"), delimiter=",", skiprows=1)
    w = W.dem_gas(n=a.shape[0])
    w.pos = a[:, 1:4].copy()
    w.radius = a[:, 4].copy()
    w.cutoff_radius = a[:, 4].copy()
    w.vel = a[:, 5:8].copy()
    w.omega = a[:, 8:11].copy()
    w.mass = W.dem_mass(w.density, w.radius)
    tags = a[:, 0].astype(np.uint32)
    return w, tags


def main():
    w = W.dem_gas(n=500)
    o = oracle_from_workload(w)
    o.run(20000)
    st = o.atoms()
    order = np.argsort(st["tag"])
    rows = np.column_stack([st["tag"][order], st["pos"][order], w.radius[st["tag"][order]],
                            st["vel"][order], st["omega"][order]])
    np.savetxt(os.path.join(HERE, "cfg0_state.csv"), rows, delimiter=",", header=COLUMNS,
               comments="", fmt=["%d"] + ["%.17g"] * 10)
    w2, tags = load_state()
    o2 = oracle_from_workload(w2)
    o2.run(10)
    a = o2.atoms()
    order = np.argsort(a["tag"])
    go, gp, gc = o2.pairs()
    lo, hi, spring = o2.history()
    hk = np.lexsort((hi, lo))
    np.savez(os.path.join(HERE, "cfg0_step10.npz"), force=a["force"][order],
             torque=a["torque"][order], pos=a["pos"][order], vel=a["vel"][order],
             omega=a["omega"][order], pairs=canonical_pairs(go, gp, gc),
             hist_lo=lo[hk], hist_hi=hi[hk], hist_spring=spring[hk])
    print("pairs", len(go), "contacts with history", len(lo))


if __name__ == "__main__":
    main()
