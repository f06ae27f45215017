"""The LJ cell-tile layout (mdd_set_force_lanes(-1)) against the checker and
against the row layout.

Tile rows are 16-bit indices into each bin cell's candidate list; the tile
kernel stages the list's current positions in LDS and fuses the step's final
half step (Verlet kick, Nose-Hoover kick and scale) with the forces. Pair
sets expanded from tile rows must equal the checker's bit for bit; forces and
trajectories must match it to the same tolerances as the row kernels."""
import sys

import numpy as np
import pytest

from mddem_amd import hip
from mddem_amd import workloads as W
from tests.helpers import canonical_pairs, kats, oracle_from_workload, rel_err
from tests.test_gpu_parity import (FP64_TOL, MIXED_TOL, _build_case, _oracle_rebuild,
                                   _perturbed_lj, by_tag)

pytestmark = pytest.mark.gpu
K = kats()
TILE = -1


@pytest.fixture(autouse=True)
def _announce(request):
    print(f"\n[gpu case] {request.node.name}", flush=True)
    sys.stdout.flush()
    yield


@pytest.mark.parametrize("name", ["trig30", "rand_uniform", "rand_slab", "rand_poly", "lj_fcc_6"])
def test_tile_pair_set_bit_exact(name):
    o, s = _build_case(name)
    s.set_force_lanes(TILE)
    _oracle_rebuild(o)
    s.bin_neighbor_list()
    go, gp, gc = s.pairs()
    oo, op, oc = o.pairs()
    key = lambda a, b, c: np.sort((a.astype(np.int64) << 37) | (b.astype(np.int64) << 5) | c)
    assert len(oo) > 0
    assert len(go) == len(oo)
    assert np.array_equal(key(go, gp, gc), key(oo, op, oc))
    assert np.array_equal(canonical_pairs(go, gp, gc), canonical_pairs(oo, op, oc))
    assert s.npairs() == len(oo)


@pytest.mark.parametrize("prec", [hip.FP64, hip.MIXED])
def test_tile_forces_match_oracle(prec):
    w = _perturbed_lj()
    o = oracle_from_workload(w)
    s = hip.from_workload(w, prec)
    s.set_force_lanes(TILE)
    _oracle_rebuild(o)
    o.zero_all_forces()
    o.lj_force()
    o.reverse_send_force()
    s.bin_neighbor_list()
    s.lj_force()
    ref = o.atoms()
    st = s.state()
    f, _ = s.forces()
    tol = FP64_TOL if prec == hip.FP64 else MIXED_TOL
    assert rel_err(by_tag(st["tag"], f), by_tag(ref["tag"], ref["force"])) < tol


@pytest.mark.parametrize("thermostat", [True, False])
def test_tile_trajectory_matches_oracle(thermostat):
    # fused final half step: Nose-Hoover (kick, scale, per-cell kinetic
    # energy) or plain velocity Verlet
    w = W.lj_fcc(ncell=4, thermostat=thermostat)
    o = oracle_from_workload(w)
    s = hip.from_workload(w, hip.FP64, box_exit=True)
    s.set_force_lanes(TILE)
    o.run(300)
    s.run(300)
    ref = o.atoms()
    st = s.state()
    assert rel_err(by_tag(st["tag"], st["vel"]), by_tag(ref["tag"], ref["vel"])) < 1e-9
    th = s.thermo()
    if thermostat:
        assert abs(th["p_xi"] - o.p_xi) < 1e-9 * max(1.0, abs(o.p_xi))
    assert th["rebuilds"] == o.rebuilds


def test_tile_and_rows_agree_on_a_mixed_trajectory():
    w = W.lj_fcc(ncell=8)
    out = {}
    for layout in (2, TILE):
        s = hip.from_workload(w, hip.MIXED, box_exit=False)
        s.set_force_lanes(layout)
        s.run(200)
        st = s.state()
        th = s.thermo(energy=True)
        out[layout] = (by_tag(st["tag"], st["vel"]), th)
        s.close()
    # FP32 force sums in another order: trajectories stay close over 200 steps
    assert rel_err(out[TILE][0], out[2][0]) < 1e-4
    assert abs(out[TILE][1]["pe"] - out[2][1]["pe"]) < 1e-5 * abs(out[2][1]["pe"])
    assert out[TILE][1]["rebuilds"] == out[2][1]["rebuilds"]


def test_tile_runs_are_bitwise_deterministic():
    w = W.lj_fcc(ncell=5)
    outs = []
    for _ in range(2):
        s = hip.from_workload(w, hip.MIXED)
        s.set_force_lanes(TILE)
        s.run(150)
        st = s.state()
        outs.append((st["pos"], st["vel"], s.forces()[0]))
        s.close()
    for a, b in zip(outs[0], outs[1]):
        assert np.array_equal(a, b)


def test_tile_switches_back_to_rows():
    w = _perturbed_lj()
    s = hip.from_workload(w, hip.FP64)
    s.set_force_lanes(TILE)
    s.run(5)
    s.set_force_lanes(1)
    s.run(5)
    s.set_force_lanes(TILE)
    s.run(5)
    r = hip.from_workload(w, hip.FP64)
    r.run(15)
    a, b = s.state(), r.state()
    assert rel_err(by_tag(a["tag"], a["vel"]), by_tag(b["tag"], b["vel"])) < 1e-12


def test_tile_256k_fcc_full_size_properties():
    w = W.lj_fcc(ncell=40)
    s = hip.from_workload(w, hip.MIXED, box_exit=False)
    s.set_force_lanes(TILE)
    s.bin_neighbor_list()
    assert s.npairs() == w.n * K["fcc_neighbors_within_2p8"] // 2
    s.run(200)
    th = s.thermo(energy=True)
    assert np.all(np.abs(th["momentum"]) < 1e-9 * w.n)
    assert 0.3 < th["temperature"] < 1.5
    assert np.isfinite(th["pe"]) and th["pe"] < 0
