"""Pins the checker's wall family against the reference's own known-answer
tests (dem_wall/src/lib.rs:1299-2450) and closed forms: plane walls with
adhesion, cohesion and twisting friction, cylinder and sphere walls, and
plane-wall motion (constant velocity, oscillation, servo). No GPU needed."""
import math

import numpy as np
import pytest

from oracle.oracle import MODEL_DEM, Oracle, dem_mass

R = 1e-3


def e_eff():
    return 1.0 / (2 * (1 - 0.3 ** 2) / 8.7e9)


def hertz(delta, r=R):
    """Static Hertz wall force (wall radius and mass infinite: R* = r)."""
    return 4.0 / 3.0 * e_eff() * math.sqrt(delta * r) * delta


def one(pos, vel=(0, 0, 0), omega=(0, 0, 0), dt=1e-7, extras=None, models=None):
    o = Oracle(MODEL_DEM)
    o.set_box((-1, -1, -1), (1, 1, 1), (False, False, False))
    o.set_materials([8.7e9], [0.3], [0.95], [0.4])
    if extras:
        o.set_material_extras(**{k: [v] for k, v in extras.items()})
    if models:
        o.set_contact_models(**models)
    o.set_dt(dt)
    m = dem_mass(2500.0, R)
    o.set_atoms([pos], [vel], [m], [R], radius=[R], omega=[omega])
    return o


def wall_force(o):
    o.zero_all_forces()
    o.wall_zero_force_accumulators()
    o.wall_contact_force()
    a = o.atoms()
    return a["force"][0], a["torque"][0], o.walls()["force"]


def test_plane_wall_hertz_magnitude_and_direction():
    # wall_repulsive_for_overlap / angled_wall_force_direction (lib.rs:1299, 1395)
    o = one((0, 0, 0.0008))
    assert o.add_wall((0, 0, 0), (0, 0, 2.0)) == 0  # normal normalised
    f, t, acc = wall_force(o)
    d = R - 0.0008
    assert f[2] == pytest.approx(hertz(d), rel=1e-12) and f[0] == 0.0 and f[1] == 0.0
    assert acc[0] == pytest.approx(f[2], rel=1e-15)
    o = one((0.0005, 0.0, 0.0005))
    o.add_wall((0, 0, 0), (1, 0, 1))
    f, _, _ = wall_force(o)
    n = np.array([1, 0, 1]) / math.sqrt(2)
    dist = 0.0005 * math.sqrt(2)
    assert np.allclose(f, hertz(R - dist) * n, rtol=1e-12)
    with pytest.raises(RuntimeError):
        o.add_wall((0, 0, 0), (0, 0, 0))


def test_plane_wall_bounds_inactive_and_far():
    # bounded_wall_ignores_out_of_bounds_atom, inactive_wall_applies_no_force,
    # wall_zero_for_no_overlap (lib.rs:1429, 1361, 1333)
    o = one((0, 0, 0.0008))
    o.add_wall((0, 0, 0), (0, 0, 1), bound_lo=(0.001, -1, -1), bound_hi=(1, 1, 1))
    assert np.all(wall_force(o)[0] == 0.0)
    o = one((0, 0, 0.0008))
    o.add_wall((0, 0, 0), (0, 0, 1))
    o.set_wall_active(0, False)
    assert np.all(wall_force(o)[0] == 0.0)
    o = one((0, 0, 0.002))
    o.add_wall((0, 0, 0), (0, 0, 1))
    assert np.all(wall_force(o)[0] == 0.0)


def test_plane_wall_adhesion_and_cohesion():
    # JKR gap regime: pure pull-off 3/2 pi gamma r; DMT: 2 pi gamma r in
    # contact only; SJKR: k_coh pi delta r (wall_cohesion_attractive_for_small_overlap)
    gap = 1e-9
    o = one((0, 0, R + gap), extras=dict(surface_energy=1.0))
    o.add_wall((0, 0, 0), (0, 0, 1))
    f, _, acc = wall_force(o)
    assert f[2] == pytest.approx(-1.5 * math.pi * 1.0 * R, rel=1e-12)
    assert acc[0] == pytest.approx(f[2], rel=1e-15)
    o = one((0, 0, R + gap), extras=dict(surface_energy=1.0), models=dict(dmt=True))
    o.add_wall((0, 0, 0), (0, 0, 1))
    assert np.all(wall_force(o)[0] == 0.0)
    z = R - 1e-9
    d = R - z  # the realised overlap
    o = one((0, 0, z), extras=dict(surface_energy=1.0), models=dict(dmt=True))
    o.add_wall((0, 0, 0), (0, 0, 1))
    f, _, _ = wall_force(o)
    assert f[2] == pytest.approx(hertz(d) - 2 * math.pi * R, rel=1e-12)
    o = one((0, 0, z), extras=dict(cohesion_energy=1e8))
    o.add_wall((0, 0, 0), (0, 0, 1))
    f, _, _ = wall_force(o)
    assert f[2] == pytest.approx(hertz(d) - 1e8 * math.pi * d * R, rel=1e-12) and f[2] < 0


def test_plane_wall_twisting_friction():
    # spin about the wall normal: tau = -sign(w.n) mu_tw |F| r n
    o = one((0, 0, 0.0009), omega=(0, 0, 50.0), extras=dict(twisting_friction=0.1))
    o.add_wall((0, 0, 0), (0, 0, 1))
    f, t, _ = wall_force(o)
    assert t[2] == pytest.approx(-0.1 * abs(f[2]) * R, rel=1e-14)
    assert t[0] == 0.0 and t[1] == 0.0


@pytest.mark.parametrize("inside", [True, False])
def test_cylinder_wall_radial_force(inside):
    # cylinder_inside_repels_toward_center, cylinder_force_always_points_
    # radially_inward (lib.rs:1694, 1871); outside walls push outward
    c = np.array([0.005, 0.005])
    for ang in np.linspace(0, 2 * np.pi, 7)[:-1]:
        rad = 0.0035 if inside else 0.0045
        p = (c[0] + rad * math.cos(ang), c[1] + rad * math.sin(ang), 0.005)
        o = one(p)
        o.add_cylinder_wall(2, c, 0.004, lo=0.0, hi=0.01, inside=inside)
        f, _, acc = wall_force(o)
        sgn = -1.0 if inside else 1.0
        n = sgn * np.array([math.cos(ang), math.sin(ang), 0.0])
        d = min(R - abs(0.004 - rad), 0.5 * R)
        assert np.allclose(f, hertz(d) * n, rtol=1e-9, atol=1e-12)
        assert acc[0] == pytest.approx(hertz(d), rel=1e-9)


def test_cylinder_axis_and_axial_bounds():
    # cylinder_axial_bounds_enforced (lib.rs:1950); an x-axis cylinder
    o = one((0.005, 0.005, 0.0125))
    o.add_cylinder_wall(2, (0.005, 0.005 - 0.0035), 0.004, lo=0.0, hi=0.01, inside=True)
    assert np.all(wall_force(o)[0] == 0.0)
    o = one((0.3, 0.0, 0.0035))
    o.add_cylinder_wall(0, (0.0, 0.0), 0.004, inside=True)
    f, _, _ = wall_force(o)
    assert f[0] == 0.0 and f[1] == 0.0 and f[2] == pytest.approx(-hertz(0.0005), rel=1e-9)


@pytest.mark.parametrize("inside", [True, False])
def test_sphere_wall_radial_force(inside):
    # sphere_inside_repels_toward_center, sphere_force_points_toward_center_
    # at_multiple_positions (lib.rs:1783, 2006)
    ctr = np.array([0.01, 0.01, 0.01])
    u = np.array([1.0, 2.0, -2.0]) / 3.0
    rad = 0.0045 if inside else 0.0055
    o = one(ctr + rad * u)
    o.add_sphere_wall(ctr, 0.005, inside=inside)
    f, _, _ = wall_force(o)
    sgn = -1.0 if inside else 1.0
    assert np.allclose(f, sgn * hertz(0.0005) * u, rtol=1e-9)
    o = one(ctr + 0.002 * u)
    o.add_sphere_wall(ctr, 0.005, inside=True)
    assert np.all(wall_force(o)[0] == 0.0)


def test_curved_walls_sjkr_cohesion():
    z = 0.004 - R + 1e-6
    d = R - (0.004 - z)  # the realised overlap
    o = one((0.0, 0.0, z), extras=dict(cohesion_energy=1e6))
    o.add_sphere_wall((0, 0, 0), 0.004, inside=True)
    f, _, _ = wall_force(o)
    assert f[2] == pytest.approx(-(hertz(d) - 1e6 * math.pi * d * R), rel=1e-12)


def test_constant_velocity_and_oscillating_walls_move():
    # constant_velocity_wall_moves, oscillating_wall_follows_sine (lib.rs:1502, 1531)
    o = one((0.5, 0.5, 0.5), dt=0.001)
    o.add_wall((0, 0, 0), (0, 0, 1))
    o.set_wall_motion(0, 1, (0.0, 0.0, -0.01))
    o.add_wall((0, 0, 0.1), (0, 0, 1))
    o.set_wall_motion(1, 2, (0.002, 50.0))
    o.wall_move()
    w = o.walls()
    assert abs(w["point"][0, 2] - (-0.00001)) < 1e-15
    phase = 2 * math.pi * 50.0 * 0.001
    assert abs(w["point"][1, 2] - (0.1 + 0.002 * math.sin(phase))) < 1e-12
    assert w["velocity"][1, 2] == pytest.approx(0.002 * 2 * math.pi * 50.0 * math.cos(phase),
                                                rel=1e-12)
    o.wall_move()
    w = o.walls()
    assert abs(w["point"][1, 2] - (0.1 + 0.002 * math.sin(2 * phase))) < 1e-12


def test_servo_wall_velocity_from_the_measured_force():
    # servo_wall_adjusts_velocity (lib.rs:1563): v = clamp(gain (target - F))
    # along the normal, F the contact force of the last pass
    o = one((0, 0, 0.0008), dt=0.001)
    o.add_wall((0, 0, 0), (0, 0, 1))
    o.set_wall_motion(0, 3, (100.0, 0.1, 0.001))
    wall_force(o)
    F = hertz(R - 0.0008)
    o.wall_move()
    w = o.walls()
    v = max(-0.1, min(0.1, 0.001 * (100.0 - F)))
    assert w["velocity"][0, 2] == pytest.approx(v, rel=1e-12)
    assert w["point"][0, 2] == pytest.approx(v * 0.001, rel=1e-12)
    # the clamp
    o.set_wall_motion(0, 3, (1e6, 0.1, 0.001))
    wall_force(o)
    o.wall_move()
    assert o.walls()["velocity"][0, 2] == 0.1


def test_moving_wall_relative_velocity_sets_the_damping():
    # moving_wall_relative_velocity_affects_damping (lib.rs:1603)
    fs = []
    for vz in (0.0, 0.05):
        o = one((0, 0, 0.0009))
        o.add_wall((0, 0, 0), (0, 0, 1))
        o.set_wall_motion(0, 1, (0.0, 0.0, vz))
        fs.append(wall_force(o)[0][2])
    assert fs[1] > fs[0] > 0
