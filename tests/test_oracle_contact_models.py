"""Pins the checker's contact-model family against the reference's own
known-answer tests (dem_granular/src/contact.rs:1149-2500) and closed forms:
SJKR cohesion, JKR and DMT adhesion, Hooke springs, constant and SDS rolling
resistance and twisting friction, and the optional mixing tables
(dem_atom/src/lib.rs:417-511). No GPU needed."""
import math

import numpy as np
import pytest

from oracle.oracle import HIST_REFERENCE, MODEL_DEM, Oracle, dem_mass

R = 1e-3  # the reference tests' sphere radius
R_EFF = R / 2


def pair(sep, extras=None, models=None, vel0=(0, 0, 0), vel1=(0, 0, 0), om0=(0, 0, 0),
         om1=(0, 0, 0), dt=1e-7, e=0.95, mu=0.4):
    """The reference's two-sphere fixture (push_test_atom_with_history,
    contact.rs:1000-1010): glass, r = 1 mm, half list [0 -> 1]."""
    o = Oracle(MODEL_DEM)
    o.set_box((-1, -1, -1), (1, 1, 1), (False, False, False))
    o.set_materials([8.7e9], [0.3], [e], [mu])
    if extras:
        o.set_material_extras(**{k: [v] for k, v in extras.items()})
    if models:
        o.set_contact_models(**models)
    o.set_dt(dt)
    o.set_history_mode(HIST_REFERENCE)
    m = dem_mass(2500.0, R)
    o.set_atoms([[0, 0, 0], [sep, 0, 0]], [vel0, vel1], [m, m], [R, R], radius=[R, R],
                omega=[om0, om1])
    o.set_half_list([0, 1, 1], [1])
    return o


def force(o, reps=1):
    for _ in range(reps):
        o.zero_all_forces()
        o.hertz_mindlin_contact_force()
    a = o.atoms()
    return a["force"], a["torque"]


def e_eff():
    return 1.0 / (2 * (1 - 0.3 ** 2) / 8.7e9)


def test_mixing_tables_follow_the_reference_rules():
    o = Oracle(MODEL_DEM)
    o.set_materials([8.7e9, 70e9, 3e9], [0.3, 0.25, 0.35], [0.9, 0.8, 0.7], [0.4, 0.3, 0.5])
    o.set_material_extras(rolling_friction=[0.04, 0.09, 0.0], twisting_friction=[-1.0, 4.0, 9.0],
                          cohesion_energy=[2.0, 8.0, 0.0], surface_energy=[1.0, 4.0, 0.0],
                          kn=[1e6, 3e6, 0.0], kt=[2e5, 2e5, 0.0],
                          rolling_stiffness=[2.0, 0.0, 6.0], rolling_damping=[4.0, 1.0, -1.0],
                          twisting_stiffness=[0.0, 0.0, 3.0], twisting_damping=[9.0, 4.0, 0.0])
    # geometric means
    assert o.extra_table("rolling_friction")[0, 1] == pytest.approx(0.06, rel=1e-15)
    assert o.extra_table("cohesion_energy")[0, 1] == 4.0
    assert o.extra_table("surface_energy")[0, 1] == 2.0
    assert o.extra_table("rolling_damping")[0, 1] == 2.0
    assert o.extra_table("twisting_damping")[0, 1] == 6.0
    # negative values count as zero where the reference clamps them
    assert o.extra_table("twisting_friction")[0, 1] == 0.0
    assert o.extra_table("twisting_friction")[1, 2] == 6.0
    assert o.extra_table("rolling_damping")[0, 2] == 0.0
    # harmonic means: Hooke stiffness needs both, SDS stiffness keeps the one set
    assert o.extra_table("kn")[0, 1] == pytest.approx(1.5e6, rel=1e-15)
    assert o.extra_table("kt")[0, 1] == 2e5
    assert o.extra_table("kn")[0, 2] == 0.0
    assert o.extra_table("rolling_stiffness")[0, 1] == 2.0
    assert o.extra_table("rolling_stiffness")[0, 2] == pytest.approx(3.0, rel=1e-15)
    assert o.extra_table("twisting_stiffness")[0, 1] == 0.0
    assert o.extra_table("twisting_stiffness")[1, 2] == 3.0
    for k in ("kn", "rolling_stiffness", "twisting_friction", "surface_energy"):
        t = o.extra_table(k)
        assert np.array_equal(t, t.T)


def test_zero_extras_reproduce_plain_hertz_mindlin_bitwise():
    # contact.rs:1212-1257: zero optional parameters change nothing
    kw = dict(vel1=(0.01, 0.2, -0.1), om0=(5, -3, 2), om1=(1, 4, -6))
    fa, ta = force(pair(0.0019, **kw), reps=3)
    zeros = {k: 0.0 for k in ("rolling_friction", "twisting_friction", "cohesion_energy",
                              "surface_energy", "rolling_stiffness", "twisting_stiffness")}
    fb, tb = force(pair(0.0019, extras=zeros, models=dict(rolling_sds=True, twisting_sds=True),
                        **kw), reps=3)
    assert np.array_equal(fa, fb) and np.array_equal(ta, tb)


def test_cohesion_attracts_with_area_proportional_force():
    # contact.rs:1164-1210 (cohesion 1e9 J/m^2, delta = 1e-8): attractive, Newton 3rd.
    # SJKR: F = k_n d - k_coh pi d R*; here the Hertz term is below 1 % of it
    o = pair(0.00199999, extras=dict(cohesion_energy=1e9))
    f, _ = force(o)
    d = 2 * R - 0.00199999
    f_coh = 1e9 * math.pi * d * R_EFF
    k_n = 4.0 / 3.0 * e_eff() * math.sqrt(R_EFF * d)
    assert f[0, 0] > 0
    assert f[0, 0] + f[1, 0] == 0.0
    assert abs(f[0, 0] - (f_coh - k_n * d)) / f_coh < 1e-9
    assert k_n * d < 0.01 * f_coh


def test_jkr_pulloff_force_matches_theory():
    # contact.rs:1268-1324: gap 1e-9 inside the JKR range, gamma = 1
    f, t = force(pair(2 * R + 1e-9, extras=dict(surface_energy=1.0)))
    expected = 1.5 * math.pi * 1.0 * R_EFF
    assert f[0, 0] > 0 and abs(f[0, 0] - expected) / expected < 1e-12
    assert f[0, 0] + f[1, 0] == 0.0
    assert np.all(t == 0.0)


def test_jkr_range_is_the_pulloff_distance():
    # interaction up to sum_r + (pi^2 gamma^2 R* / (4 E*^2))^(1/3) (contact.rs:163-175)
    gamma = 1.0
    d_po = (math.pi ** 2 * gamma ** 2 * R_EFF / (4 * e_eff() ** 2)) ** (1 / 3)
    inside, _ = force(pair(2 * R + 0.99 * d_po, extras=dict(surface_energy=gamma)))
    outside, _ = force(pair(2 * R + 1.01 * d_po, extras=dict(surface_energy=gamma)))
    assert inside[0, 0] > 0
    assert np.all(outside == 0.0)
    # contact.rs:1376-1411: far beyond it nothing either
    far, _ = force(pair(0.003, extras=dict(surface_energy=gamma)))
    assert np.all(far == 0.0)


def test_jkr_gap_regime_keeps_no_history():
    # the adhesion-only regime skips the tangential part and its history
    o = pair(2 * R + 1e-9, extras=dict(surface_energy=1.0), vel1=(0, 0.5, 0))
    f, _ = force(o)
    assert f[0, 1] == 0.0 and f[0, 2] == 0.0
    assert len(o.history7()[0]) == 0


def test_dmt_pulloff_force_and_range():
    # contact.rs:1997-2113: in contact F = k_n d - 2 pi gamma R*, none past contact
    d = 1e-10
    f, _ = force(pair(2 * R - d, extras=dict(surface_energy=1.0), models=dict(dmt=True)))
    f_dmt = 2 * math.pi * 1.0 * R_EFF
    k_n = 4.0 / 3.0 * e_eff() * math.sqrt(R_EFF * d)
    assert f[0, 0] > 0 and abs(f[0, 0] - (f_dmt - k_n * d)) / f_dmt < 1e-9
    gap, _ = force(pair(2 * R + 1e-9, extras=dict(surface_energy=1.0), models=dict(dmt=True)))
    assert np.all(gap == 0.0)
    # DMT pulls harder than JKR's 3/2 pi gamma R*
    jkr, _ = force(pair(2 * R - d, extras=dict(surface_energy=1.0)))
    assert f[0, 0] > jkr[0, 0] > 0


HOOKE = dict(kn=1e6, kt=5e5)


@pytest.mark.parametrize("d", [5e-5, 1e-4, 2e-4])
def test_hooke_normal_force_is_linear(d):
    # contact.rs:1429-1505 and 2268-2313: F = k_n d at rest, zero past contact
    f, _ = force(pair(2 * R - d, extras=HOOKE, models=dict(hooke=True)))
    d_real = 2 * R - (2 * R - d)
    assert abs(-f[0, 0] - 1e6 * d_real) / (1e6 * d_real) < 1e-12
    gap, _ = force(pair(0.0021, extras=HOOKE, models=dict(hooke=True)))
    assert np.all(gap == 0.0)


def test_hooke_ignores_surface_energy():
    # the Hooke kernel has no JKR / DMT branch (contact.rs:622-990)
    a, _ = force(pair(2 * R - 1e-4, extras=HOOKE, models=dict(hooke=True)))
    b, _ = force(pair(2 * R - 1e-4, extras=dict(HOOKE, surface_energy=1.0),
                      models=dict(hooke=True)))
    assert np.array_equal(a, b)


def test_hooke_head_on_restitution_matches_the_material():
    # linear spring-dashpot with gamma_n = 2 beta sqrt(k m_r) rebounds with
    # e_n = exp(-beta pi / sqrt(1 - beta^2)) = e (the clamp at F = 0 ends the
    # contact a little early, within 2 %)
    e = 0.9
    o = Oracle(MODEL_DEM)
    o.set_box((-0.01, -0.01, -0.01), (0.01, 0.01, 0.01), (False, False, False))
    o.set_materials([8.7e9], [0.3], [e], [0.4])
    o.set_material_extras(kn=[1e8], kt=[5e7])
    o.set_contact_models(hooke=True)
    o.set_neighbor(1.1, 0.005, 0, True, 1000, True)
    o.set_dt(2e-8)
    m = dem_mass(2500.0, R)
    v0 = 0.2
    o.set_atoms([[-R - 1e-7, 0, 0], [R + 1e-7, 0, 0]], [[v0 / 2, 0, 0], [-v0 / 2, 0, 0]],
                [m, m], [R, R], radius=[R, R])
    touched = False
    for _ in range(20000):
        o.step()
        a = o.atoms()
        gap = abs(a["pos"][1, 0] - a["pos"][0, 0]) - 2 * R
        if gap < 0:
            touched = True
        elif touched:
            break
    v = o.atoms()["vel"]
    cor = (v[1, 0] - v[0, 0]) / v0
    assert touched and abs(cor - e) / e < 0.02


def test_constant_twisting_torque_opposes_spin_with_the_cap_magnitude():
    # contact.rs:1508-1548: spin about the normal; |tau| = mu_tw |F_n| R*
    f, t = force(pair(0.0019, extras=dict(twisting_friction=0.05), om0=(100, 0, 0)))
    assert t[0, 0] < 0
    assert abs(t[0, 0] + 0.05 * abs(f[0, 0]) * R_EFF) < 1e-15 * abs(f[0, 0])
    assert t[0, 0] + t[1, 0] == 0.0
    # contact.rs:1550-1589: no spin, no twisting torque
    _, t0 = force(pair(0.0019, extras=dict(twisting_friction=0.05)))
    assert np.all(t0 == 0.0)


def test_constant_rolling_torque_opposes_relative_rolling():
    # contact.rs:1592-1633; the rolling part is the difference to mu_r = 0
    kw = dict(om0=(0, 0, 100))
    f0, t0 = force(pair(0.0019, **kw))
    f1, t1 = force(pair(0.0019, extras=dict(rolling_friction=0.3), **kw))
    assert np.array_equal(f0, f1)
    tr = t1[0] - t0[0]
    assert tr[2] < 0 and tr[0] == 0.0 and tr[1] == 0.0
    assert abs(-tr[2] - 0.3 * abs(f1[0, 0]) * R_EFF) / (0.3 * abs(f1[0, 0]) * R_EFF) < 1e-12


SDS_ROLL = dict(rolling_friction=0.3, rolling_stiffness=1e-3, rolling_damping=1e-8)
SDS_TWIST = dict(twisting_friction=0.3, twisting_stiffness=1e-3, twisting_damping=1e-8)


def test_sds_rolling_spring_accumulates_then_caps():
    # contact.rs:1670-1828: relative rolling w about z for N calls:
    # theta = N w dt, tau = -k theta - g w, capped at mu_r |F_n| R*
    w, n, dt = 10.0, 5, 1e-7
    o = pair(0.0019, extras=SDS_ROLL, models=dict(rolling_sds=True), om0=(0, 0, w), dt=dt)
    f_base, t_base = force(pair(0.0019, om0=(0, 0, w), dt=dt), reps=n)
    f, t = force(o, reps=n)
    theta = n * w * dt
    tau = -1e-3 * theta - 1e-8 * w
    assert abs((t[0, 2] - t_base[0, 2]) - tau) < 1e-12 * abs(tau)
    _, _, h = o.history7()
    assert abs(h[0, 5] - theta) < 1e-12 * theta          # lower tag's frame: +z
    # a stiff spring reaches the slider: |tau| = cap, theta = (tau + g w) / -k
    o = pair(0.0019, extras=dict(SDS_ROLL, rolling_stiffness=1e5),
             models=dict(rolling_sds=True), om0=(0, 0, w), dt=dt)
    f, t = force(o, reps=n)
    cap = 0.3 * abs(f[0, 0]) * R_EFF
    _, t_base = force(pair(0.0019, om0=(0, 0, w), dt=dt), reps=n)
    assert abs(abs(t[0, 2] - t_base[0, 2]) - cap) < 1e-12 * cap
    _, _, h = o.history7()
    assert abs(h[0, 5] - (-cap + 1e-8 * w) / -1e5) < 1e-12 * abs(h[0, 5])


def test_sds_twisting_spring_accumulates_then_caps():
    # contact.rs:1830-1923: spin w about the normal for N calls
    w, n, dt = 10.0, 5, 1e-7
    o = pair(0.0019, extras=SDS_TWIST, models=dict(twisting_sds=True), om0=(w, 0, 0), dt=dt)
    f, t = force(o, reps=n)
    phi = n * w * dt
    tau = -1e-3 * phi - 1e-8 * w
    assert abs(t[0, 0] - tau) < 1e-12 * abs(tau)
    assert t[0, 0] + t[1, 0] == 0.0
    _, _, h = o.history7()
    assert abs(h[0, 6] - phi) < 1e-12 * phi
    o = pair(0.0019, extras=dict(SDS_TWIST, twisting_stiffness=1e5),
             models=dict(twisting_sds=True), om0=(w, 0, 0), dt=dt)
    f, t = force(o, reps=n)
    cap = 0.3 * abs(f[0, 0]) * R_EFF
    assert abs(t[0, 0] + cap) < 1e-12 * cap


def test_constant_models_ignore_sds_parameters():
    # contact.rs:1925-1985
    kw = dict(om0=(30, 0, 20))
    a = force(pair(0.0019, extras=dict(rolling_friction=0.3, twisting_friction=0.2), **kw))
    b = force(pair(0.0019, extras=dict(rolling_friction=0.3, twisting_friction=0.2,
                                       rolling_stiffness=5.0, twisting_stiffness=5.0,
                                       rolling_damping=1.0, twisting_damping=1.0), **kw))
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


def test_history_slots_are_stored_in_the_lower_tag_frame():
    # sign = +1 when tag_i < tag_j (contact.rs:337-339, 598-602): seen from a
    # pair listed the other way round the stored spring and rolling values are
    # identical. The twist rate (w_i - w_j) . n is symmetric in the pair, so
    # the same sign convention stores the twist angle negated when the other
    # sphere owns the pair (the reference's convention, kept as it is)
    kw = dict(extras=dict(SDS_ROLL, **SDS_TWIST), models=dict(rolling_sds=True, twisting_sds=True),
              vel1=(0, 0.1, 0.05), om0=(4, 0, 7), om1=(0, -3, 1))
    a = pair(0.0019, **kw)
    force(a, reps=3)
    b = pair(0.0019, **kw)
    b.set_half_list([0, 0, 1], [0])           # the pair owned by tag 1
    force(b, reps=3)
    _, _, ha = a.history7()
    ob, pb, hb = b.history7()
    assert list(ob) == [1] and list(pb) == [0]
    assert np.allclose(ha[:, :6], hb[:, :6], rtol=1e-12, atol=1e-30)
    assert np.allclose(ha[:, 6], -hb[:, 6], rtol=1e-12, atol=1e-30) and ha[0, 6] != 0.0
