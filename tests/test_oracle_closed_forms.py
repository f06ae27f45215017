"""Closed-form pins of the checker's Hertz-Mindlin damping and Mindlin
tangential force, worked out here from the reference's formulas with its
literal constant SQRT_5_3 = 0.9128709291752768 (dem_atom/src/lib.rs:63):

  S_n = 2 E* sqrt(R* d),  k_n = 4/3 E* sqrt(R* d),  k_t = 8 G* sqrt(R* d)
  F_n = max(k_n d - 2 beta SQRT_5_3 sqrt(S_n m_r) v_n, 0)   (contact.rs:252-320)
  s   = v_t dt on a contact's first step,  F_t = k_t s - gamma_t v_t,
  gamma_t = 2 SQRT_5_3 beta sqrt(k_t m_r)                    (contact.rs:349-408)
  plane wall: R* = r, m_r = m, F = max(k_n d - 2 beta SQRT_5_3 sqrt(S_n m) v_n, 0)
                                                            (dem_wall/src/lib.rs:925-968)

Each case is set up so that every damping term is a sizeable part of the
force and no Coulomb cap is active, and is held to 1e-12 relative. The last
test builds the checker with the constant scaled by sqrt(2) (sqrt(5/3)
literally, a plausible slip) and shows these cases fail on it. No GPU needed.
"""
import math
import os
import subprocess
import sys

import numpy as np
import pytest

from oracle.oracle import HIST_CANONICAL, HIST_REFERENCE, MODEL_DEM, Oracle, dem_mass

SQRT_5_3 = 0.9128709291752768
RHO = 2500.0
GLASS = dict(E=8.7e9, nu=0.3, e=0.95, mu=0.4)
STEEL = dict(E=200e9, nu=0.28, e=0.8, mu=0.3)
TOL = 1e-12
HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


def mix(a, b):
    """beta, mu, E*, G* of a material pair (dem_atom/src/lib.rs:430-466)."""
    le = math.log(math.sqrt(a["e"] * b["e"]))
    beta = -le / math.sqrt(math.pi ** 2 + le * le)
    mu = math.sqrt(a["mu"] * b["mu"])
    es = 1.0 / ((1 - a["nu"] ** 2) / a["E"] + (1 - b["nu"] ** 2) / b["E"])
    gs = 1.0 / (2 * (2 - a["nu"]) * (1 + a["nu"]) / a["E"]
                + 2 * (2 - b["nu"]) * (1 + b["nu"]) / b["E"])
    return beta, mu, es, gs


def mass(r):
    return RHO * 4.0 / 3.0 * math.pi * r ** 3


def first_contact(x0, x1, r0, r1, v0, v1, w0, w1, mat0=GLASS, mat1=GLASS, dt=1e-7, steps=1):
    """Force and torque on sphere 0 and 1 after `steps` force passes of a
    contact that starts with no stored spring, the geometry held fixed."""
    beta, mu, es, gs = mix(mat0, mat1)
    x0, x1, v0, v1, w0, w1 = (np.asarray(a, float) for a in (x0, x1, v0, v1, w0, w1))
    d = x1 - x0
    dist = np.linalg.norm(d)
    n = d / dist
    delta = min(r0 + r1 - dist, 0.5 * min(r0, r1))
    rs = r0 * r1 / (r0 + r1)
    mr = 1.0 / (1.0 / mass(r0) + 1.0 / mass(r1))
    sdr = math.sqrt(delta * rs)
    sn, kn, kt = 2 * es * sdr, 4.0 / 3.0 * es * sdr, 8 * gs * sdr
    vc0 = v0 + np.cross(w0, r0 * n)     # contact-point velocity of sphere 0
    vc1 = v1 + np.cross(w1, -r1 * n)    # and of sphere 1
    vr = vc1 - vc0
    vn = float(vr @ n)
    fn = max(kn * delta - 2 * beta * SQRT_5_3 * math.sqrt(sn * mr) * vn, 0.0)
    vt = vr - vn * n
    s = steps * vt * dt                 # v_t dt per pass, n and v_t fixed
    gt = 2 * SQRT_5_3 * beta * math.sqrt(kt * mr)
    ft = kt * s - gt * vt
    assert np.linalg.norm(kt * s) < mu * fn and np.linalg.norm(ft) < mu * fn, "cap active"
    f0 = -fn * n + ft
    t0 = np.cross(r0 * n, ft)
    t1 = np.cross(r1 * n, ft)
    return dict(f0=f0, f1=-f0, t0=t0, t1=t1, fn=fn, ft=ft, spring=s, kn_delta=kn * delta, n=n)


def check_pair(a, want, tol=TOL):
    """Normal and tangential parts of both forces and both torques, each held
    on its own (the normal force would hide a tangential error). tol bounds
    each part relative to its own size; a force rounded as a whole (FP32 sums)
    passes tol_sum, an error bound relative to the whole force."""
    n = want["n"]
    whole = want["fn"] + np.linalg.norm(want["ft"])
    for k, sgn in ((0, 1.0), (1, -1.0)):
        f = a["force"][k]
        assert abs(sgn * (f @ n) + want["fn"]) <= tol * want["fn"]
        ft = f - (f @ n) * n
        assert np.linalg.norm(ft - sgn * want["ft"]) <= max(10 * tol * np.linalg.norm(want["ft"]),
                                                            tol * whole if tol > TOL else 0.0)
        assert rel(a["torque"][k], want["t%d" % k]) < 10 * tol


def oracle_pair(x0, x1, r0, r1, v0, v1, w0, w1, mats=(GLASS,), types=(0, 0), dt=1e-7,
                owner=0, history=HIST_REFERENCE):
    o = Oracle(MODEL_DEM)
    o.set_box((-1, -1, -1), (1, 1, 1), (False, False, False))
    o.set_materials([m["E"] for m in mats], [m["nu"] for m in mats], [m["e"] for m in mats],
                    [m["mu"] for m in mats])
    o.set_dt(dt)
    o.set_history_mode(history)
    o.set_atoms([x0, x1], [v0, v1], [mass(r0), mass(r1)], [r0, r1], radius=[r0, r1],
                omega=[w0, w1], atype=list(types))
    # the pair sits in the row of `owner` (neighbor.rs half list)
    o.set_half_list([0, 1, 1] if owner == 0 else [0, 0, 1], [1] if owner == 0 else [0])
    return o


def rel(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def test_mass_and_material_closed_forms_agree_with_checker():
    assert dem_mass(RHO, 1e-3) == pytest.approx(mass(1e-3), rel=1e-15)
    o = Oracle(MODEL_DEM)
    o.set_materials([GLASS["E"], STEEL["E"]], [GLASS["nu"], STEEL["nu"]],
                    [GLASS["e"], STEEL["e"]], [GLASS["mu"], STEEL["mu"]])
    t = o.material_tables()
    beta, mu, es, gs = mix(GLASS, STEEL)
    assert t["beta"][0, 1] == pytest.approx(beta, rel=1e-14)
    assert t["friction"][0, 1] == pytest.approx(mu, rel=1e-14)
    assert t["e_eff"][0, 1] == pytest.approx(es, rel=1e-14)
    assert t["g_eff"][0, 1] == pytest.approx(gs, rel=1e-14)


@pytest.mark.parametrize("owner", [0, 1])
def test_sliding_pair_first_step_tangential_force(owner):
    # spring_history_accumulates_tangential_force (tangential.rs:353-408):
    # r = 1 mm glass, separation 1.9 mm, sphere 1 slides at 0.1 m/s along y,
    # dt = 1e-7. First step: F_t = k_t v_t dt - gamma_t v_t.
    args = ([0, 0, 0], [0.0019, 0, 0], 1e-3, 1e-3, [0, 0, 0], [0, 0.1, 0], [0] * 3, [0] * 3)
    want = first_contact(*args)
    # 0.017605 N of spring against 0.009049 N of damping
    assert want["ft"][1] == pytest.approx(0.0085560, rel=5e-5)
    o = oracle_pair(*args, owner=owner)
    o.hertz_mindlin_contact_force()
    a = o.atoms()
    check_pair(a, want)
    _, _, spring = o.history()
    # stored in the lower tag's frame whichever sphere owns the pair
    assert rel(spring[0], want["spring"]) < TOL


@pytest.mark.parametrize("history", [HIST_REFERENCE, HIST_CANONICAL])
def test_spring_accumulates_over_passes(history):
    # with the geometry held, pass k sees s = k v_t dt (contact.rs:366-373)
    args = ([0, 0, 0], [0.0019, 0, 0], 1e-3, 1e-3, [0, 0, 0], [0, 0.1, 0.05], [0] * 3, [0] * 3)
    o = oracle_pair(*args, history=history)
    for k in range(1, 4):
        o.zero_all_forces()
        o.hertz_mindlin_contact_force()
        want = first_contact(*args, steps=k)
        check_pair(o.atoms(), want)
        assert rel(o.history()[2][0], want["spring"]) < TOL


@pytest.mark.parametrize("vn", [-1.0, 0.05])
def test_normal_force_with_damping(vn):
    # d = 1e-6 m, e = 0.5: the damping term is 3x (closing) or 0.15x (opening)
    # of the elastic term k_n d
    mat = dict(GLASS, e=0.5)
    args = ([0, 0, 0], [0.002 - 1e-6, 0, 0], 1e-3, 1e-3, [0, 0, 0], [vn, 0, 0], [0] * 3, [0] * 3)
    want = first_contact(*args, mat0=mat, mat1=mat)
    damp = abs(want["fn"] - want["kn_delta"])
    assert damp > 0.1 * want["kn_delta"]
    o = oracle_pair(*args, mats=(mat,))
    o.hertz_mindlin_contact_force()
    a = o.atoms()
    assert rel(a["force"][0], want["f0"]) < TOL
    assert a["force"][0, 1] == 0.0 and a["force"][0, 2] == 0.0


def test_normal_force_clamped_at_zero_when_opening_fast():
    mat = dict(GLASS, e=0.5)
    o = oracle_pair([0, 0, 0], [0.002 - 1e-6, 0, 0], 1e-3, 1e-3, [0, 0, 0], [1.0, 0, 0],
                    [0] * 3, [0] * 3, mats=(mat,))
    o.hertz_mindlin_contact_force()
    a = o.atoms()
    assert np.all(a["force"] == 0.0)


@pytest.mark.parametrize("owner", [0, 1])
def test_unequal_radii_mixed_materials_with_spin(owner):
    # r0 = 0.8 mm glass against r1 = 1.2 mm steel along an oblique normal:
    # R* = 0.48 mm, m_r of unequal masses, contact-point velocities with
    # both spins (contact.rs:147-425)
    n = np.array([0.6, 0.8, 0.0])
    x1 = (1.999e-3 * n).tolist()
    args = ([0, 0, 0], x1, 0.8e-3, 1.2e-3, [0.05, 0.02, -0.01], [-0.8, -0.5, 0.2],
            [0, 0, 30.0], [10.0, -20.0, 5.0])
    want = first_contact(*args, mat0=GLASS, mat1=STEEL)
    # damping is a sizeable part of both the normal and the tangential force
    assert abs(want["fn"] - want["kn_delta"]) > 0.05 * want["fn"]
    o = oracle_pair(*args, mats=(GLASS, STEEL), types=(0, 1), owner=owner)
    o.hertz_mindlin_contact_force()
    a = o.atoms()
    check_pair(a, want)
    # the two torques differ by the lever ratio r0 / r1
    assert rel(a["torque"][0] * 1.2e-3, a["torque"][1] * 0.8e-3) < 1e-12


@pytest.mark.parametrize("vz", [-0.5, 0.05])
def test_plane_wall_damping(vz):
    # a sphere 0.999 mm above the plane z = 0 moving at vz, e = 0.6
    # (dem_wall/src/lib.rs:925-968): R* = r, m_r = m
    mat = dict(GLASS, e=0.6)
    r, z = 1e-3, 0.999e-3
    beta, _, _, _ = mix(mat, mat)
    es = 1.0 / (2 * (1 - mat["nu"] ** 2) / mat["E"])
    d = r - z
    sdr = math.sqrt(d * r)
    sn, kn = 2 * es * sdr, 4.0 / 3.0 * es * sdr
    want = max(kn * d - 2 * beta * SQRT_5_3 * math.sqrt(sn * mass(r)) * vz, 0.0)
    assert abs(want - kn * d) > 0.05 * kn * d
    o = Oracle(MODEL_DEM)
    o.set_box((-1, -1, -1), (1, 1, 1), (False, False, False))
    o.set_materials([mat["E"]], [mat["nu"]], [mat["e"]], [mat["mu"]])
    o.set_dt(1e-7)
    o.set_atoms([[0, 0, z]], [[0.3, 0, vz]], [mass(r)], [r], radius=[r], omega=[[0, 0, 0]])
    o.add_wall((0, 0, 0), (0, 0, 1))
    o.zero_all_forces()
    o.wall_zero_force_accumulators()
    o.wall_contact_force()
    f = o.atoms()["force"][0]
    # no tangential wall force (lib.rs:862-1002)
    assert f[0] == 0.0 and f[1] == 0.0
    assert f[2] == pytest.approx(want, rel=TOL)
    assert o.walls()["force"][0] == pytest.approx(want, rel=TOL)


def test_closed_forms_reject_a_checker_with_a_wrong_damping_constant(tmp_path):
    """Build oracle.c with SQRT_5_3 = sqrt(5/3) (sqrt(2) too large) and run
    this file against it: every case with a damping term must fail."""
    lib = tmp_path / "liboracle_mutant.so"
    src = os.path.join(ROOT, "oracle", "oracle.c")
    subprocess.run(["gcc", "-O1", "-ffp-contract=off", "-fPIC", "-shared", "-std=c11",
                    "-DSQRT_5_3=1.2909944487358056", "-o", str(lib), src, "-lm"], check=True)
    env = dict(os.environ, MDDEM_ORACLE_LIB=str(lib), PYTHONDONTWRITEBYTECODE="1")
    res = subprocess.run([sys.executable, "-m", "pytest", "-q", "-p", "no:cacheprovider",
                          "-k", "not reject", os.path.abspath(__file__), "-rf"],
                         cwd=ROOT, env=env, capture_output=True, text=True)
    out = res.stdout
    assert res.returncode != 0, out
    for case in ("test_sliding_pair_first_step_tangential_force[0]",
                 "test_sliding_pair_first_step_tangential_force[1]",
                 "test_spring_accumulates_over_passes",
                 "test_normal_force_with_damping[-1.0]", "test_normal_force_with_damping[0.05]",
                 "test_unequal_radii_mixed_materials_with_spin",
                 "test_plane_wall_damping[-0.5]", "test_plane_wall_damping[0.05]"):
        assert f"FAILED tests/test_oracle_closed_forms.py::{case}" in out, (case, out[-3000:])
    # the cases with no damping term still pass on the mutant
    assert "test_mass_and_material_closed_forms_agree_with_checker" not in out.split("short test summary")[-1]


@pytest.mark.parametrize("owner", [0, 1])
def test_contact_virial_of_one_pair(owner):
    # VirialStress::add_pair(d, F_i) with d = x_j - x_i and F_i = -F_n + F_t,
    # the force on the owner i (contact.rs:559-569, virial.rs:59-67): for one
    # pair the tensor is d (x) F_i, the same whichever sphere owns the pair
    n = np.array([0.6, 0.8, 0.0])
    x1 = 1.999e-3 * n
    args = ([0, 0, 0], x1.tolist(), 0.8e-3, 1.2e-3, [0.05, 0.02, -0.01], [-0.8, -0.5, 0.2],
            [0, 0, 30.0], [10.0, -20.0, 5.0])
    want = first_contact(*args, mat0=GLASS, mat1=STEEL)
    w = np.outer(x1, want["f0"])  # d from sphere 0 to 1, force on sphere 0
    want6 = np.array([w[0, 0], w[1, 1], w[2, 2], w[0, 1], w[0, 2], w[1, 2]])
    o = oracle_pair(*args, mats=(GLASS, STEEL), types=(0, 1), owner=owner)
    o.set_virial(True)
    o.hertz_mindlin_contact_force()
    got = o.contact_virial()
    assert np.linalg.norm(got - want6) <= 1e-11 * np.linalg.norm(want6)
    # the trace is -r F_n: repulsion gives a negative trace
    assert got[:3].sum() == pytest.approx(-want["fn"] * np.linalg.norm(x1), rel=1e-11)
    o.set_virial(False)
    o.zero_all_forces()
    o.hertz_mindlin_contact_force()
    assert np.array_equal(o.contact_virial(), got)  # inactive: left as it was
