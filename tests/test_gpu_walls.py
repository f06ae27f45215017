"""The wall family on the GPU against the checker (dem_wall/src/lib.rs:
766-1241): plane walls with JKR / SJKR adhesion and twisting friction,
cylinder, sphere and region-surface walls (a union of a block and a cone cut
by a plane), and moving plane walls (constant velocity, oscillation, servo)
inside scheduled runs. A cloud of glass spheres is held
in a spherical container crossed by a cylinder, with an oscillating floor and
a servo lid, so that every wall kind carries contacts."""
import sys

import numpy as np
import pytest

from mddem_amd import hip
from mddem_amd import workloads as W
from tests.helpers import oracle_from_workload, rel_err

pytestmark = pytest.mark.gpu

CTR = np.array([0.015, 0.015, 0.015])
OBSTACLE = CTR + np.array([0.0, 0.0, -0.004])
R = 1e-3


@pytest.fixture(autouse=True)
def _announce(request):
    print(f"\n[gpu case] {request.node.name}", flush=True)
    sys.stdout.flush()
    yield


def by_tag(tags, arr):
    return np.asarray(arr)[np.argsort(tags)]


def _cloud(n=300, seed=21, speed=0.6):
    """n spheres with centres inside a ball of radius 0.011 about CTR (so the
    outer ones touch the container), separated by 2.1 r."""
    w = W.dem_gas(n=n, box=0.03, seed=seed, speed=speed)
    st = W.Stream(seed + 7)
    pts = []
    while len(pts) < n:
        p = CTR + (st.uniform(3) * 2 - 1) * 0.011
        if np.linalg.norm(p - CTR) > 0.011:
            continue
        if np.linalg.norm(p - OBSTACLE) < 0.002 + R + 1e-5:
            continue
        # clear of the region wall's cone (inside its bounding cylinder)
        if p[2] > 0.02 - R - 1e-4 and np.hypot(p[0] - 0.015, p[1] - 0.015) < 0.004 + R + 1e-4:
            continue
        if all(np.linalg.norm(p - q) > 2.1 * R for q in pts):
            pts.append(p)
    w.pos = np.array(pts)
    w.periodic = (False, False, False)
    w.gravity = (0.0, 0.0, -9.81)
    w.walls = []
    return w


def _walls(sys_, oracle):
    """The same walls on both sides (no sphere starts overlapping one)."""
    for s in (sys_, oracle):
        s.add_sphere_wall(CTR, 0.012, inside=True)                       # 0 container
        s.add_cylinder_wall(0, (0.015, 0.015), 0.012, lo=0.0, hi=0.03, inside=True)  # 1
        s.add_wall((0, 0, 0.0029), (0, 0, 1))                           # 2 floor
        s.set_wall_motion(2, 2, (2e-4, 2000.0))                         # oscillating
        s.add_wall((0, 0, 0.0271), (0, 0, -3.0))                        # 3 lid
        s.set_wall_motion(3, 3, (0.5, 0.2, 0.05))                       # servo
        s.add_sphere_wall(OBSTACLE, 0.002)                              # 4 obstacle
        s.add_wall((0.003, 0, 0), (1, 0, 0), bound_lo=(-1, -1, 0.012), bound_hi=(1, 1, 1))  # 5
        s.add_wall((0.026, 0, 0), (-1, 0, 0))                           # 6 (deactivated)
        s.set_wall_active(6, False)
        s.add_wall((0, 0, 0), (0, 1, 0))                                # 7 moves at 1 m/s
        s.set_wall_motion(7, 1, (0.0, 1.0, 0.0))
        _add_region(s)                                                  # 8 region
    return 9


# a slab under the container and a cone hanging from above, as one
# region-surface wall (spheres stay outside it)
REGION = {"type": "union", "regions": [
    {"type": "block", "min": [0.002, 0.002, 0.0], "max": [0.028, 0.028, 0.0031]},
    {"type": "intersect", "regions": [
        {"type": "cone", "center": [0.015, 0.015], "axis": "z", "rad_lo": 0.004,
         "rad_hi": 0.001, "lo": 0.02, "hi": 0.028},
        {"type": "plane", "point": [0.0, 0.0, 0.03], "normal": [0.0, 0.0, 1.0]}]},
]}


def _add_region(s):
    if isinstance(s, hip.System):
        s.add_region_wall(REGION)
    else:
        types, params = hip.flatten_region(REGION)
        s.add_region_wall(types, params)


EXTRAS = {
    "plain": {},
    "adhesive": dict(surface_energy=0.02, twisting_friction=0.1),
    "cohesive": dict(cohesion_energy=2e3, twisting_friction=0.05, rolling_friction=0.05),
}


def _pair(extras, prec=hip.FP64):
    w = _cloud()
    o = oracle_from_workload(w)
    s = hip.from_workload(w, prec, box_exit=True)
    if extras:
        for x in (o, s):
            x.set_material_extras(**{k: [v] for k, v in extras.items()})
    _walls(s, o)
    return w, o, s


def _near_walls():
    """Spheres placed at overlaps 8, 2 and 0.01 um and at a 2 nm gap from each
    wall (the gap is inside the JKR range at the adhesive settings)."""
    pts = []
    deltas = (8e-6, 2e-6, 1e-8, -2e-9)
    st = W.Stream(77)
    for k, d in enumerate(deltas):
        jx, jy = 0.006 + 0.004 * k, 0.008 + 0.003 * k
        pts.append((jx, jy, 0.0029 + R - d))                # floor
        pts.append((0.003 + R - d, jy, 0.013))              # bounded plane, in bounds
        pts.append((0.003 + R - d, jy, 0.011))              # bounded plane, out of bounds
        pts.append((jx, jy, 0.0271 - R + d))                # servo lid
        pts.append((0.026 - R + d, jy, 0.015))              # deactivated plane
        pts.append((jx, R - d, 0.02))                       # moving plane
        u = st.uniform(3) - 0.5
        u /= np.linalg.norm(u)
        pts.append(tuple(OBSTACLE + u * (0.002 + R - d)))   # obstacle sphere
        pts.append(tuple(CTR + u * (0.012 - R + d)))        # container sphere
        ang = 2 * np.pi * st.uniform(1)[0]
        rr = 0.012 - R + d
        pts.append((0.01 + 0.002 * k, 0.015 + rr * np.cos(ang), 0.015 + rr * np.sin(ang)))
        pts.append((jx, jy + 0.004, 0.0031 + R - d))        # region: slab top
        # region: the cone's side (radius 0.0025 at z = 0.024, slope 3 / 8)
        zc = 0.024
        nrm = np.array([np.cos(ang), np.sin(ang), 3 / 8])
        nrm /= np.linalg.norm(nrm)
        base = np.array([0.015 + 0.0025 * np.cos(ang), 0.015 + 0.0025 * np.sin(ang), zc])
        pts.append(tuple(base + nrm * (R - d)))
    return np.array(pts)


@pytest.mark.parametrize("extras", list(EXTRAS))
def test_wall_forces_match_oracle(extras):
    w, o, s = _pair(EXTRAS[extras])
    pos = _near_walls()
    n = len(pos)
    st0 = W.Stream(5)
    vel = (st0.uniform(3 * n).reshape(-1, 3) - 0.5) * 0.4
    om = (st0.uniform(3 * n).reshape(-1, 3) - 0.5) * 400
    m = np.full(n, w.mass[0])
    r = np.full(n, R)
    o.set_atoms(pos, vel, m, r, radius=r, omega=om)
    s.upload(pos, vel, m, r, radius=r, omega=om)
    o.zero_all_forces()
    o.wall_zero_force_accumulators()
    o.wall_contact_force()
    s.set_forces(np.zeros((n, 3)), np.zeros((n, 3)))
    s.wall_zero_force_accumulators()
    s.wall_contact_force()
    ref = o.atoms()
    f, t = s.forces()
    st = s.state()
    assert rel_err(by_tag(st["tag"], f), by_tag(ref["tag"], ref["force"])) < 1e-12
    if "twisting_friction" in EXTRAS[extras]:
        assert np.abs(t).max() > 0
    assert rel_err(by_tag(st["tag"], t), by_tag(ref["tag"], ref["torque"])) < 1e-12
    gw, ow = s.walls(), o.walls()
    assert np.abs(gw["force"] - ow["force"]).max() <= 1e-9 * np.abs(ow["force"]).max()
    # every wall but the deactivated one carries force
    assert np.all(ow["force"][[0, 1, 2, 3, 4, 5, 7, 8]] != 0.0) and ow["force"][6] == 0.0


@pytest.mark.parametrize("prec", [hip.FP64, hip.MIXED])
@pytest.mark.parametrize("extras", ["plain", "cohesive"])
def test_scheduled_runs_with_moving_walls_match_oracle(extras, prec):
    w, o, s = _pair(EXTRAS[extras], prec)
    touched = False
    for _ in range(4):
        o.run(100)
        s.run(100)
        fw = o.walls()["force"]
        touched = touched or fw[0] != 0.0 or fw[1] != 0.0
        assert np.abs(s.walls()["force"] - fw).max() <= 1e-6 * max(1.0, np.abs(fw).max())
    ref = o.atoms()
    st = s.state()
    tol = 1e-9 if prec == hip.FP64 else 1e-4
    dp = by_tag(st["tag"], st["pos"]) - by_tag(ref["tag"], ref["pos"])
    assert np.abs(dp).max() < tol * 0.03
    assert rel_err(by_tag(st["tag"], st["vel"]), by_tag(ref["tag"], ref["vel"])) < tol * 10
    gw, ow = s.walls(), o.walls()
    # the oscillating floor and the servo lid moved the same way
    assert np.abs(gw["point"] - ow["point"]).max() < 1e-12
    assert np.abs(gw["velocity"] - ow["velocity"]).max() < (1e-9 if prec == hip.FP64 else 1e-5)
    assert ow["velocity"][3, 2] < 0  # the lid descends toward its target force
    assert touched  # contacts happened on the curved walls


def test_wall_runs_are_bitwise_deterministic():
    outs = []
    for _ in range(2):
        w, o, s = _pair(EXTRAS["adhesive"], hip.MIXED)
        s.run(300)
        outs.append((s.state()["vel"], s.walls()["point"], s.walls()["force"]))
    for a, b in zip(*outs):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("prec", [hip.FP64, hip.MIXED])
def test_one_lane_split_pass_matches_oracle(prec):
    # at one lane per sphere the force pass runs the spheres with a nonempty
    # row (listed at each rebuild) and the others (walls and gravity only) in
    # two kernels; most of this cloud's spheres have no neighbor
    w, o, s = _pair(EXTRAS["plain"], prec)
    s.set_force_lanes(0)
    for _ in range(3):
        o.run(100)
        s.run(100)
        fw = o.walls()["force"]
        assert np.abs(s.walls()["force"] - fw).max() <= 1e-6 * max(1.0, np.abs(fw).max())
    ref = o.atoms()
    st = s.state()
    tol = 1e-9 if prec == hip.FP64 else 1e-4
    dp = by_tag(st["tag"], st["pos"]) - by_tag(ref["tag"], ref["pos"])
    assert np.abs(dp).max() < tol * 0.03
    assert rel_err(by_tag(st["tag"], st["vel"]), by_tag(ref["tag"], ref["vel"])) < tol * 10
    assert 0 < s.npairs() and s.thermo()["rebuilds"] == o.rebuilds
