"""Shared test helpers: configure the CPU checker from a Workload and compare
pair sets / forces between implementations."""
from __future__ import annotations

import json
import os

import numpy as np

from oracle.oracle import (HIST_REFERENCE, MODEL_DEM, MODEL_LJ, Oracle)

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def kats():
    with open(os.path.join(GOLDEN, "kats.json")) as fh:
        return json.load(fh)


def oracle_from_workload(w, history_mode=HIST_REFERENCE, sort_every=1000):
    o = Oracle(MODEL_DEM if w.model == "dem" else MODEL_LJ)
    o.set_box(w.lo, w.hi, w.periodic)
    o.set_neighbor(w.skin_fraction, w.bin_size, getattr(w, "every", 0), True, sort_every, True)
    # with every > 0 the reference permutes the atoms under the old list on
    # the step before a count rebuild; the HIP path keeps the list and the
    # order until the rebuild, and the checker is compared with that
    o.set_sort_only_on_rebuild(getattr(w, "every", 0) > 0)
    o.set_dt(w.dt)
    if w.model == "dem":
        m = w.materials
        o.set_materials(m["youngs"], m["poisson"], m["restitution"], m["friction"])
        o.set_history_mode(history_mode)
        if w.gravity is not None:
            o.set_gravity(w.gravity)
        for p, nrm, mat in w.walls:
            o.add_wall(p, nrm, mat)
    else:
        eps, sig, cut = w.lj
        o.set_lj(eps, sig, cut)
        if w.nose_hoover is not None:
            o.set_nose_hoover(*w.nose_hoover)
            o.set_integrators(False, False)
        else:
            o.set_integrators(True, False)
    o.set_atoms(w.pos, w.vel, w.mass, w.cutoff_radius, radius=w.radius, omega=w.omega,
                tag=w.tag, atype=w.atype)
    return o


def canonical_pairs(owner, partner, code):
    """(min tag, max tag, image code seen from the min tag) as a sorted array."""
    owner = np.asarray(owner, np.int64)
    partner = np.asarray(partner, np.int64)
    code = np.asarray(code, np.int64)
    swap = owner > partner
    lo = np.where(swap, partner, owner)
    hi = np.where(swap, owner, partner)
    c = np.where(swap, 26 - code, code)
    key = (lo << 37) | (hi << 5) | c
    return np.sort(key)


def brute_force_pairs(pos, cutoff, lo, hi, periodic):
    """All pairs (i < j, image) with |x_j + sL - x_i| < cutoff over the 27
    images of periodic axes. O(N^2 * 27): tiny inputs only."""
    pos = np.asarray(pos, np.float64)
    L = np.asarray(hi, np.float64) - np.asarray(lo, np.float64)
    n = len(pos)
    keys = []
    shifts = []
    for sx in (-1, 0, 1):
        for sy in (-1, 0, 1):
            for sz in (-1, 0, 1):
                s = np.array([sx, sy, sz])
                if np.any((s != 0) & ~np.asarray(periodic, bool)):
                    continue
                shifts.append(s)
    for i in range(n):
        for s in shifts:
            img = pos + s * L
            d = img - pos[i]
            r2 = np.sum(d * d, axis=1)
            code = (s[0] + 1) * 9 + (s[1] + 1) * 3 + (s[2] + 1)
            for j in np.nonzero(r2 < cutoff * cutoff)[0]:
                if j == i and code == 13:
                    continue
                if j > i or (j == i and code > 13):
                    keys.append((i << 37) | (int(j) << 5) | code)
    return np.sort(np.array(keys, dtype=np.int64))


def rel_err(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    den = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / den) if den > 0 else float(np.linalg.norm(a - b))


def canon_history(a, b, spring):
    """History entries as (lower tag, higher tag, value) sorted by the tag
    pair: the checker's per-owner store lists (owner, partner), the HIP path
    (lower, higher); the values are in the lower tag's frame in both."""
    a = np.asarray(a, np.int64)
    b = np.asarray(b, np.int64)
    lo, hi = np.minimum(a, b), np.maximum(a, b)
    k = np.lexsort((hi, lo))
    return lo[k], hi[k], np.asarray(spring)[k]
