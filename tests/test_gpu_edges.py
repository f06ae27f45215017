"""Edge cases of the HIP path against the CPU checker: bin cells holding more
particles than a wavefront has lanes (the cell build takes its members in
batches of 64), a single particle (no pairs at all, free flight across the
periodic faces), and a gas that contracts until its rows outgrow the row
capacity of the first build (the build is repeated with longer rows).
References: neighbor.rs:810-1080 (bin_neighbor_list), md_lj/src/lib.rs:488-633,
mddem_verlet/src/lib.rs:97-154."""
import sys

import numpy as np
import pytest

from mddem_amd import hip
from mddem_amd import workloads as W
from oracle.oracle import MODEL_LJ, Oracle
from tests.helpers import canonical_pairs, oracle_from_workload, rel_err

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _announce(request):
    print(f"\n[gpu case] {request.node.name}", flush=True)
    sys.stdout.flush()
    yield


def _oracle_rebuild(o):
    o.neighbor_setup()
    o.pbc()
    o.sort_atoms_by_bin()
    o.borders()
    o.bin_neighbor_list()


def _row_lengths(n, owner, partner):
    return np.bincount(owner, minlength=n) + np.bincount(partner, minlength=n)


@pytest.mark.parametrize("poly", [False, True])
def test_crowded_cells_pair_set_bit_exact(poly):
    # 1200 particles packed in a 3.5-wide cube of a 9 x 8 x 10 periodic box
    # (28 per unit volume) over a sparse background: bins of the 1.54 pair
    # cutoff hold over 100 particles, and rows of about 430 entries overflow
    # the first build's row capacity estimate
    st = W.Stream(31 + int(poly))
    L = np.array([9.0, 8.0, 10.0])
    bg = st.uniform(3 * 300).reshape(-1, 3) * L
    cl = np.array([3.0, 2.5, 4.0]) + st.uniform(3 * 1200).reshape(-1, 3) * 3.5
    pos = np.vstack([bg, cl])
    n = len(pos)
    cut_r = 0.55 + 0.15 * st.uniform(n) if poly else np.full(n, 0.7)
    per = (True, True, True)
    o = Oracle(MODEL_LJ)
    o.set_box((0, 0, 0), L, per)
    o.set_neighbor(1.1, 0.4, 0, True, 1000, True)
    o.set_atoms(pos, np.zeros_like(pos), np.ones(n), cut_r)
    s = hip.System(hip.MODEL_LJ)
    s.set_box((0, 0, 0), L, per)
    s.set_neighbor(1.1, 0.4, 0, True, 1000, True)
    s.upload(pos, np.zeros_like(pos), np.ones(n), cut_r)
    _oracle_rebuild(o)
    s.bin_neighbor_list()
    # a bin of at least the largest pair cutoff fits inside the cube along
    # every axis, so some cell holds more than 64 particles
    cut = 2.0 * cut_r.max() * 1.1
    nb = np.floor(L / cut).astype(int)
    cell = np.floor(pos / (L / nb)).astype(int)
    _, counts = np.unique(cell[:, 0] * 10000 + cell[:, 1] * 100 + cell[:, 2], return_counts=True)
    assert counts.max() > 64
    go, gp, gc = s.pairs()
    oo, op, oc = o.pairs()
    key = lambda a, b, c: np.sort((a.astype(np.int64) << 37) | (b.astype(np.int64) << 5) | c)
    assert len(go) == len(oo) > 100000
    assert np.array_equal(key(go, gp, gc), key(oo, op, oc))
    assert np.array_equal(canonical_pairs(go, gp, gc), canonical_pairs(oo, op, oc))
    assert _row_lengths(n, oo, op).max() > 300
    s.close()


def _single(model):
    if model == "lj":
        w = W.lj_fcc(ncell=4, thermostat=False)
        vel = np.array([[0.9, -0.7, 0.4]])
    else:
        w = W.dem_gas(n=20)
        vel = np.array([[61.0, -47.0, 33.0]])  # across the 25 mm box in 3000 steps
    w.pos = w.pos[:1].copy()
    w.vel = vel
    w.mass = w.mass[:1].copy()
    w.cutoff_radius = w.cutoff_radius[:1].copy()
    for f in ("radius", "omega", "atype"):
        v = getattr(w, f)
        if v is not None:
            setattr(w, f, v[:1].copy())
    return w


@pytest.mark.parametrize("model", ["lj", "dem"])
def test_single_particle_moves_freely(model):
    # no pair at any step; the particle crosses periodic faces, so the
    # wrap and the rebuilds it triggers are exercised with an empty list
    w = _single(model)
    o = oracle_from_workload(w)
    s = hip.from_workload(w, hip.FP64, box_exit=True)
    steps = 2000 if model == "lj" else 3000
    o.run(steps)
    s.run(steps)
    ref = o.atoms()
    st = s.state()
    L = w.hi - w.lo
    dp = st["pos"] - ref["pos"]
    dp -= L * np.round(dp / L)
    assert np.abs(dp).max() < 1e-12 * L.max()
    assert np.array_equal(st["vel"], ref["vel"])
    # the free flight itself: x(t) = x0 + v t modulo L
    x = np.mod(w.pos + w.vel * (steps * w.dt) - w.lo, L) + w.lo
    d = st["pos"] - x
    d -= L * np.round(d / L)
    assert np.abs(d).max() < 1e-9 * L.max()
    th = s.thermo()
    assert th["rebuilds"] == o.rebuilds and o.rebuilds >= 1
    assert s.npairs() == 0
    s.close()


def test_contracting_gas_outgrows_its_rows():
    # an LJ gas at rho = 0.3 whose velocities point at the box centre: over
    # 300 steps the centre's density more than doubles, so the rows outgrow
    # the capacity of the first build and the build is repeated with longer
    # rows; the trajectory follows the checker's
    w = W.lj_fcc(ncell=5, rho=0.3, thermostat=False)
    L = w.hi - w.lo
    c = w.lo + 0.5 * L
    d = w.pos - c
    d -= L * np.round(d / L)
    w.vel = -1.2 * d / (0.5 * L)
    o0 = oracle_from_workload(w)
    _oracle_rebuild(o0)
    a0, b0, _ = o0.pairs()
    row0 = _row_lengths(w.n, a0, b0).max()
    o = oracle_from_workload(w)
    s = hip.from_workload(w, hip.FP64, box_exit=True)
    o.run(300)
    s.run(300)
    ref = o.atoms()
    st = s.state()
    order_s = np.argsort(st["tag"])
    order_o = np.argsort(ref["tag"])
    assert rel_err(st["vel"][order_s], ref["vel"][order_o]) < 1e-9
    assert s.thermo()["rebuilds"] == o.rebuilds
    a1, b1, _ = s.pairs()
    assert _row_lengths(w.n, a1, b1).max() > 1.5 * row0 + 8
    s.close()
