"""The LJ mixed pair pass on fixed-point positions (x * 2^32 / L per axis,
32-bit differences wrapping to the minimum image): pairs at the cutoff are
decided as in FP64, the periodic images need no shift, and the forces of a
melted state stay within the mixed budget of the FP64 pass. The displacement
check on fixed-point build positions decides as the double check does, also
at the threshold (neighbor.rs:564-601)."""
import sys

import numpy as np
import pytest

from mddem_amd import hip
from mddem_amd import workloads as W
from tests.helpers import rel_err

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _announce(request):
    print(f"\n[gpu case] {request.node.name}", flush=True)
    sys.stdout.flush()
    yield


def lj_at(r, mode, periodic=False, shift=0.0):
    s = hip.System(hip.MODEL_LJ, hip.MIXED)
    s.set_box((0.0, 0.0, 0.0), (20.0, 20.0, 20.0), (periodic,) * 3)
    s.set_lj(1.0, 1.0, 2.5)
    s.set_lj_positions(mode)
    x0 = 10.0 + shift
    s.upload([[x0, 10.0, 10.0], [x0 + r, 10.0, 10.0]], np.zeros((2, 3)), [1, 1], [1.5, 1.5])
    s.set_half_list([0, 1, 1], [1])
    s.lj_force()
    f, _ = s.forces()
    s.close()
    return f


@pytest.mark.parametrize("mode", [0, 1])
def test_pairs_at_the_cutoff_are_decided_as_in_fp64(mode):
    # r = 2.5 (1 -+ 1e-8): the FP32 r^2 cannot tell the two apart (its
    # rounding is ~1e-7), the FP64 decision can; the fixed-point unit here
    # is 20 / 2^32 = 4.7e-9, 1.9e-9 of r
    f_in = lj_at(2.5 * (1.0 - 1e-8), mode)
    f_out = lj_at(2.5 * (1.0 + 1e-8), mode)
    want = 24.0 * (2.0 * 2.5 ** -13 - 2.5 ** -7)  # force on the +x particle at rc
    assert abs(f_in[1, 0] - want) < 1e-5 * abs(want)
    assert f_in[0, 0] == -f_in[1, 0]
    assert f_out[0, 0] == 0.0 and f_out[1, 0] == 0.0


def test_a_pair_across_the_periodic_boundary_needs_no_shift():
    # the same pair, once inside the box and once straddling x = 0 / 20:
    # the wrapped 32-bit difference is the minimum image
    s = hip.System(hip.MODEL_LJ, hip.MIXED)
    s.set_box((0.0, 0.0, 0.0), (20.0, 20.0, 20.0), (True, True, True))
    s.set_lj(1.0, 1.0, 2.5)
    s.upload([[19.6, 10.0, 10.0], [0.7, 10.0, 10.0]], np.zeros((2, 3)), [1, 1], [1.4, 1.4])
    s.bin_neighbor_list()
    s.lj_force()
    f, _ = s.forces()
    st = s.state()
    s.close()
    r = 1.1
    want = 24.0 * (2.0 * r ** -13 - r ** -7)
    k = np.argsort(st["tag"])
    f = f[k]
    # particle 1 (x = 0.7) is +1.1 from particle 0 (x = 19.6) through the boundary
    assert abs(f[1, 0] - want) < 1e-5 * abs(want)
    assert f[0, 0] == -f[1, 0]


def test_melted_state_forces_within_the_mixed_budget():
    # a 4000-atom LJ fluid melted in FP64; both mixed pair passes against the
    # FP64 pass at the same state
    w = W.lj_fcc(ncell=10)
    ref = hip.from_workload(w, hip.FP64, box_exit=False)
    ref.run(300)
    st = ref.state()
    k = np.argsort(st["tag"])
    pos, vel = st["pos"][k], st["vel"][k]
    f64, _ = ref.forces()
    ref.close()
    out = {}
    for name, prec, mode in (("fp64", hip.FP64, 1), ("fixed", hip.MIXED, 1), ("geom64", hip.MIXED, 0)):
        s = hip.from_workload(w, prec, box_exit=False)
        s.set_lj_positions(mode)
        s.upload(pos, vel, w.mass, w.cutoff_radius, tag=w.tag)
        s.bin_neighbor_list()
        s.lj_force()
        f, _ = s.forces()
        out[name] = f[np.argsort(s.state()["tag"])]
        s.close()
    e_fixed = rel_err(out["fixed"], out["fp64"])
    e_geom = rel_err(out["geom64"], out["fp64"])
    assert e_fixed < 3e-6 and e_geom < 3e-6, (e_fixed, e_geom)


@pytest.mark.parametrize("lg", [0, 2])
def test_fused_final_half_gives_the_same_run(lg):
    # thermostatted mixed LJ: the final half step inside the row pass (kick,
    # rescale, kinetic-energy partials summed as the post kernel sums them)
    # against the post kernel, bit for bit through rebuilds
    w = W.lj_fcc(ncell=10)
    out = []
    for fused in (False, True):
        s = hip.from_workload(w, hip.MIXED, box_exit=True)
        s.set_force_lanes(lg)
        s.set_lj_post_fused(fused)
        s.run(500)
        st = s.state()
        f, _ = s.forces()
        th = s.thermo()
        out.append((st["pos"], st["vel"], f, np.array([th["p_xi"], th["ke"]])))
        assert th["rebuilds"] >= 5
        s.close()
    for x, y in zip(out[0], out[1]):
        assert np.array_equal(x, y)


def _threshold_walkers(mode):
    # 8 particles 10 apart (no pairs) moving along the axes at |v| = 0.5:
    # with dt = 0.005 each drifts 0.0025 per step, so its displacement
    # reaches the 0.15 threshold ((1.12 - 1) x 1.25) at every 60th step to
    # within a few ulp, inside the band where the fixed-point check reads
    # the double4 build position. The thermostat's target is the state's
    # own temperature (2 KE / ndof = 2 / 21 exactly times 21), so p_xi
    # stays 0, the scale is 1 and the velocities keep their exact values.
    w = W.lj_fcc(ncell=12)
    c = np.array([[x, y, z] for x in (5.0, 15.0) for y in (5.0, 15.0) for z in (5.0, 15.0)])
    v = np.zeros((8, 3))
    for k in range(8):
        v[k, k % 3] = 0.5 if k % 2 == 0 else -0.5
    w.pos, w.vel = c, v
    w.mass = np.ones(8)
    w.cutoff_radius = np.full(8, 1.25)
    ndof = 3 * 8 - 3
    T = 2.0 / ndof
    assert ndof * T == 2.0
    w.nose_hoover = (T, 1.0)
    s = hip.from_workload(w, hip.MIXED)
    s.set_lj_positions(mode)
    s.run(200)
    st = s.state()
    rb = s.thermo()["rebuilds"]
    s.close()
    return st, rb


def test_displacement_check_at_the_threshold_is_the_double_check():
    # the fixed-point build positions (mode 1) against the double4 check
    # (mode 0): with no pairs the two runs are the same arithmetic, so the
    # rebuild steps and the states agree bit for bit
    st1, rb1 = _threshold_walkers(1)
    st0, rb0 = _threshold_walkers(0)
    assert rb1 == rb0 >= 3
    assert np.array_equal(st1["pos"], st0["pos"])
    assert np.array_equal(st1["vel"], st0["vel"])
    assert np.all(np.abs(st1["vel"]) <= 0.5)
    assert np.array_equal(np.sort(np.abs(st1["vel"]).sum(axis=1)), np.full(8, 0.5))
