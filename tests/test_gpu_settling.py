"""The DEM settling physics (polydisperse Hertz-Mindlin with shear history,
two plane walls, gravity, periodic x/y) on the HIP path against the checker,
from a resting bed in which every sphere touches its neighbors: one force
pass, then a run through several rebuilds; and the full 4M-sphere bed's
size-independent properties. References: contact.rs:147-425,
dem_wall/src/lib.rs:862-1002, mddem_fixes/src/lib.rs:756-761."""
import math
import sys

import numpy as np
import pytest

from mddem_amd import hip
from mddem_amd import workloads as W
from tests.helpers import canon_history, oracle_from_workload, rel_err

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _announce(request):
    print(f"\n[gpu case] {request.node.name}", flush=True)
    sys.stdout.flush()
    yield


def bed(n=2500, vmax=0.5, spin=50.0, seed=None, every=0):
    w = W.dem_packed_bed(n=n, box=(0.02, 0.02, 1.2), vmax=vmax, seed=seed, every=every)
    w.omega = (W.Stream(77).uniform(3 * n).reshape(n, 3) * 2.0 - 1.0) * spin
    return w


def by_tag(tags, arr):
    return np.asarray(arr)[np.argsort(tags)]


def bed_stats(w):
    from scipy.spatial import cKDTree
    L = w.hi[:2]
    p = w.pos
    tr = cKDTree(np.column_stack([p[:, :2], p[:, 2] + 1.0]), boxsize=[L[0], L[1], 10.0])
    pr = tr.query_pairs(2.0 * w.radius.max() * w.skin_fraction, output_type="ndarray")
    d = p[pr[:, 0]] - p[pr[:, 1]]
    d[:, :2] -= L * np.round(d[:, :2] / L)
    dist = np.linalg.norm(d, axis=1)
    s = w.radius[pr[:, 0]] + w.radius[pr[:, 1]]
    return (np.sum(dist < w.skin_fraction * s), np.sum(dist < s), float((s - dist).max()))


def test_bed_is_a_resting_contact_network():
    # CPU-side facts the GPU cases rely on: most spheres touch, no overlap is
    # deeper than twice the contact overlap, the bottom layer rests on z = 0
    w = bed()
    pairs, contacts, deepest = bed_stats(w)
    assert contacts > 1.2 * w.n and pairs > contacts
    assert 0 < deepest <= 2.0000001e-10  # twice the 0.1 nm contact overlap
    assert np.sum(w.pos[:, 2] < w.radius) > 30


@pytest.mark.parametrize("prec", [hip.FP64, hip.MIXED])
def test_force_pass_matches_oracle(prec):
    w = bed()
    o = oracle_from_workload(w)
    o.set_virial(True)
    o.neighbor_setup()
    o.pbc()
    o.sort_atoms_by_bin()
    o.borders()
    o.bin_neighbor_list()
    o.zero_all_forces()
    o.hertz_mindlin_contact_force()
    o.reverse_send_force()
    o.wall_zero_force_accumulators()
    o.wall_contact_force()
    o.apply_gravity()
    s = hip.from_workload(w, prec)
    s.set_virial(1)
    s.bin_neighbor_list()
    s.hertz_mindlin_contact_force()
    s.wall_zero_force_accumulators()
    s.wall_contact_force()
    s.apply_gravity()
    ref = o.atoms()
    st = s.state()
    f, t = s.forces()
    tol = 1e-12 if prec == hip.FP64 else 1e-5
    assert s.thermo()["ncontacts"] > 1.2 * w.n
    assert rel_err(by_tag(st["tag"], f), by_tag(ref["tag"], ref["force"])) < tol
    assert rel_err(by_tag(st["tag"], t), by_tag(ref["tag"], ref["torque"])) < tol
    # contact virial (contact.rs:559-569) over the same pairs
    w_ref = o.contact_virial()
    w_got = np.asarray(s.thermo()["virial"])
    assert np.linalg.norm(w_got - w_ref) <= tol * np.linalg.norm(w_ref)
    assert w_ref[:3].sum() < 0  # repulsive contacts
    # wall force sums: the device adds 2^-32 N fixed-point values, one per
    # sphere touching the floor
    wf = s.walls()["force"]
    owf = o.walls()["force"]
    nfloor = int(np.sum(w.pos[:, 2] < w.radius))
    assert owf[0] > 0 and abs(wf[0] - owf[0]) <= nfloor * 2.0 ** -32 + 1e-12 * owf[0]
    assert wf[1] == 0.0 and owf[1] == 0.0


@pytest.mark.parametrize("prec", [hip.FP64, hip.MIXED])
def test_run_through_rebuilds_matches_oracle(prec):
    # the bed at rest with spinning spheres: gravity loads the contacts and
    # their springs grow under friction, while [neighbor] every = 100
    # rebuilds the list every 101 steps, so the shear history crosses several
    # rebuilds (both keep it per owning sphere, as the reference does). Over
    # 1000 steps (0.12 ms) the bed keeps most of its contacts: spheres that
    # rest on one sphere roll off it, and some lateral contacts open
    w = bed(vmax=0.0, every=100)
    o = oracle_from_workload(w)
    s = hip.from_workload(w, prec, box_exit=True)
    o.run(1000)
    s.run(1000)
    th = s.thermo()
    assert th["rebuilds"] == o.rebuilds and o.rebuilds >= 9
    assert th["ncontacts"] > 1.1 * w.n  # 1.32 n on the checker
    ref = o.atoms()
    st = s.state()
    L = w.hi - w.lo
    dp = by_tag(st["tag"], st["pos"]) - by_tag(ref["tag"], ref["pos"])
    dp[:, :2] -= L[:2] * np.round(dp[:, :2] / L[:2])
    if prec == hip.FP64:
        assert np.abs(dp).max() < 1e-12
        assert rel_err(by_tag(st["tag"], st["vel"]), by_tag(ref["tag"], ref["vel"])) < 1e-9
        assert rel_err(by_tag(st["tag"], st["omega"]), by_tag(ref["tag"], ref["omega"])) < 1e-8
    else:
        assert np.abs(dp).max() < 1e-9
        assert rel_err(by_tag(st["tag"], st["vel"]), by_tag(ref["tag"], ref["vel"])) < 1e-4
        assert rel_err(by_tag(st["tag"], st["omega"]), by_tag(ref["tag"], ref["omega"])) < 1e-3
    a, b, spring = canon_history(*s.history())
    oa, ob, ospring = canon_history(*o.history())
    assert np.array_equal(a, oa) and np.array_equal(b, ob)
    tol = 1e-8 if prec == hip.FP64 else 1e-3
    assert rel_err(spring, ospring) < tol


@pytest.mark.timeout(600)
def test_full_size_4m_bed_properties():
    # the settling config's 4,000,000 spheres as a bed at rest: the list
    # holds pairs and contacts from the first step and keeps most of them
    # (0.28 of them after 3000 steps, as the deep bed compresses), x/y
    # momentum is exchanged only between spheres (walls push along z,
    # gravity pulls along z) and every value stays finite through rebuilds
    # ([neighbor] every = 1000)
    w = W.dem_packed_bed(every=1000)
    print(f"[gpu case] bed built: n={w.n}", flush=True)
    p0 = np.sum(w.mass[:, None] * w.vel, axis=0)
    s = hip.from_workload(w, hip.MIXED, box_exit=False)
    s.run(1)
    th = s.thermo()
    c0 = th["ncontacts"]
    assert th["npairs"] > 2 * w.n and c0 > 1.5 * w.n
    s.run(3000)
    th = s.thermo()
    assert th["rebuilds"] >= 3
    assert th["npairs"] > 2 * w.n and th["ncontacts"] > 0.2 * c0
    p1 = np.asarray(th["momentum"])
    scale = np.sum(w.mass) * 0.1
    assert np.all(np.abs(p1[:2] - p0[:2]) < 1e-6 * scale)
    st = s.state()
    assert np.all(np.isfinite(st["pos"])) and np.all(np.isfinite(st["vel"]))
    assert np.all(np.isfinite(st["omega"]))
    assert st["pos"][:, 2].min() > 0.0 and st["pos"][:, 2].max() < 0.35


@pytest.mark.parametrize("lg", [0, 1])
def test_tile_order_gives_the_same_run(lg):
    # the pair pass visits the spheres band of bin columns by band instead of
    # in index order; every sphere's arithmetic is unchanged, so positions,
    # velocities, spins, forces and springs are the same bit for bit. The
    # 0.1 m box has 41 bin columns along y, three bands of 16.
    w = W.dem_packed_bed(n=20000, box=(0.1, 0.1, 1.2), every=100)
    out = []
    for tiles in (False, True):
        s = hip.from_workload(w, hip.MIXED, box_exit=True)
        s.set_dem_tiles(tiles)
        s.set_force_lanes(lg)
        s.set_dem_split(0)
        s.run(250)
        st = s.state()
        f, t = s.forces()
        a, b, spring = canon_history(*s.history())
        out.append((st["pos"], st["vel"], st["omega"], f, t, a, b, spring))
        assert s.thermo()["rebuilds"] >= 2 and s.thermo()["ncontacts"] > w.n
        s.close()
    for x, y in zip(out[0], out[1]):
        assert np.array_equal(x, y)


@pytest.mark.parametrize("prec,tiles", [(hip.MIXED, True), (hip.MIXED, False), (hip.FP64, True)])
def test_compacted_contacts_give_the_same_run(prec, tiles):
    # one lane per sphere: each wavefront queues its pairs in contact and runs
    # their arithmetic 64 at a time, then every sphere adds its own results
    # in row order; the pair arithmetic and each sphere's order of sums are
    # those of the per-lane loop, so the runs agree bit for bit
    w = W.dem_packed_bed(n=20000, box=(0.1, 0.1, 1.2), every=100)
    out = []
    for wave in (False, True):
        s = hip.from_workload(w, prec, box_exit=True)
        s.set_dem_tiles(tiles)
        s.set_force_lanes(0)
        s.set_dem_split(0)
        s.set_dem_wave(wave)
        s.run(250)
        st = s.state()
        f, t = s.forces()
        a, b, spring = canon_history(*s.history())
        out.append((st["pos"], st["vel"], st["omega"], f, t, a, b, spring))
        th = s.thermo()
        assert th["rebuilds"] >= 2 and th["ncontacts"] > w.n
        s.close()
    for x, y in zip(out[0], out[1]):
        assert np.array_equal(x, y)


def test_pairs_listed_twice_keep_their_springs_deterministically():
    # spheres of this bed that sit on the periodic boundary (np.mod rounds a
    # few to y = L) give pairs the reference lists twice, once through each
    # periodic image (the checker's half list shows them); at each rebuild
    # the history remap must give each side its own spring, the same way in
    # every run and by either remap method
    w = W.dem_packed_bed(n=20000, box=(0.1, 0.1, 1.2), every=100)
    o = oracle_from_workload(w)
    dup = 0
    for _ in range(3):
        o.run(1)
        a, b, _c = o.pairs()
        key = np.minimum(a, b).astype(np.int64) << 32 | np.maximum(a, b)
        dup += len(key) - len(np.unique(key))
    assert dup > 0
    runs = []
    for remap in (0, 0, 1):
        s = hip.from_workload(w, hip.MIXED, box_exit=True)
        s.set_dem_split(0)
        s.set_history_remap(remap)
        s.run(20)
        st = s.state()
        f, t = s.forces()
        runs.append((st["pos"], st["vel"], st["omega"], f, t))
        assert s.thermo()["rebuilds"] >= 10
        s.close()
    for r in runs[1:]:
        for x, y in zip(runs[0], r):
            assert np.array_equal(x, y)


def test_contact_queue_with_virial_steps_gives_the_same_run():
    # with the contact virial on every third step those steps run the
    # per-lane kernel (the queue has no virial), the others the queue: the
    # springs and contact bits pass between the two kernels, and the run is
    # the per-lane run bit for bit
    w = bed(n=2500, every=100)
    out = []
    for wave in (0, 1):
        s = hip.from_workload(w, hip.MIXED, box_exit=True)
        s.set_force_lanes(0)
        s.set_dem_split(0)
        s.set_dem_wave(wave)
        s.set_virial(3)
        s.run(301)
        st = s.state()
        f, t = s.forces()
        a, b, spring = canon_history(*s.history())
        th = s.thermo()
        out.append((st["pos"], st["vel"], st["omega"], f, t, a, b, spring,
                    np.asarray(th["virial"])))
        assert th["ncontacts"] > w.n // 5
        s.close()
    for x, y in zip(out[0], out[1]):
        assert np.array_equal(x, y)
