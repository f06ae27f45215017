"""Pins the checker's region surfaces and region walls against the
reference's own known answers (mddem_core/src/region.rs:1237-1343,
dem_wall/src/lib.rs:2168-2427) and against brute force: the signed distance
of every primitive equals the distance to a dense sampling of its surface,
with the sign given by containment. No GPU needed."""
import math

import numpy as np
import pytest

from oracle.oracle import MODEL_DEM, Oracle, dem_mass

BLOCK, SPHERE, CYL, CONE, PLANE = range(5)


def prim(kind, *vals):
    q = np.zeros(12)
    flat = np.concatenate([np.atleast_1d(np.asarray(v, dtype=float)) for v in vals])
    q[:len(flat)] = flat
    return kind, q


def region_oracle(prims, inside=False, pos=(0, 0, 0), r=1e-3):
    o = Oracle(MODEL_DEM)
    o.set_box((-1, -1, -1), (1, 1, 1), (False, False, False))
    o.set_materials([8.7e9], [0.3], [0.95], [0.4])
    o.set_dt(1e-7)
    m = dem_mass(2500.0, r)
    o.set_atoms([pos], [(0, 0, 0)], [m], [r], radius=[r], omega=[(0, 0, 0)])
    k = o.add_region_wall([p[0] for p in prims], [p[1] for p in prims], inside=inside)
    return o, k


def surf(prims, pos):
    o, k = region_oracle(prims)
    return o.region_surface(k, pos)


def test_reference_region_surface_answers():
    # region.rs:1237-1343
    d, p, n = surf([prim(SPHERE, (0, 0, 0), 1.0)], (2, 0, 0))
    assert abs(d - 1) < 1e-10 and abs(p[0] - 1) < 1e-10 and abs(n[0] - 1) < 1e-10
    d, p, n = surf([prim(SPHERE, (0, 0, 0), 1.0)], (0.5, 0, 0))
    assert abs(d + 0.5) < 1e-10 and abs(p[0] - 1) < 1e-10 and abs(n[0] - 1) < 1e-10
    d, _, n = surf([prim(BLOCK, (0, 0, 0), (2, 2, 2))], (1, 1, 1.8))
    assert abs(d + 0.2) < 1e-10 and abs(n[2] - 1) < 1e-10
    d, _, n = surf([prim(BLOCK, (0, 0, 0), (1, 1, 1))], (0.5, 0.5, 1.5))
    assert abs(d - 0.5) < 1e-10 and abs(n[2] - 1) < 1e-10
    d, p, n = surf([prim(PLANE, (0, 0, 0), (0, 0, 1))], (1, 2, 3))
    assert abs(d - 3) < 1e-10 and abs(p[2]) < 1e-10 and abs(n[2] - 1) < 1e-10
    d, _, _ = surf([prim(PLANE, (0, 0, 0), (0, 0, 1))], (1, 2, -1))
    assert abs(d + 1) < 1e-10
    d, _, n = surf([prim(CYL, (0, 0), 1.0, 2, 0.0, 10.0)], (0.8, 0, 5))
    assert abs(d + 0.2) < 1e-10 and abs(n[0] - 1) < 1e-10
    d, _, _ = surf([prim(CONE, (0, 0), 2, 2.0, 2.0, 0.0, 10.0)], (1.8, 0, 5))
    assert d < 0
    d, _, _ = surf([prim(SPHERE, (0, 0, 0), 1.0), prim(SPHERE, (5, 0, 0), 1.0)], (0.5, 0, 0))
    assert d < 0


def _contains(kind, q, x):
    if kind == BLOCK:
        return np.all((x >= q[0:3]) & (x <= q[3:6]), axis=1)
    if kind == SPHERE:
        return np.linalg.norm(x - q[0:3], axis=1) <= q[3]
    ax = int(q[3] if kind == CYL else q[2])
    oth = [a for a in range(3) if a != ax]
    rr = np.hypot(x[:, oth[0]] - q[0], x[:, oth[1]] - q[1])
    a = x[:, ax]
    if kind == CYL:
        return (a >= q[4]) & (a <= q[5]) & (rr <= q[2])
    lo, hi, rlo, rhi = q[5], q[6], q[3], q[4]
    t = np.clip((a - lo) / (hi - lo), 0, 1)
    return (a >= lo) & (a <= hi) & (rr <= rlo + t * (rhi - rlo))


def _surface_samples(kind, q, n=120):
    """Dense points on the primitive's surface."""
    u = np.linspace(0, 1, n)
    th = np.linspace(0, 2 * np.pi, n, endpoint=False)
    if kind == SPHERE:
        ph = np.linspace(0, np.pi, n)
        T, P = np.meshgrid(th, ph)
        d = np.stack([np.sin(P) * np.cos(T), np.sin(P) * np.sin(T), np.cos(P)], -1)
        return q[0:3] + q[3] * d.reshape(-1, 3)
    if kind == BLOCK:
        mn, mx = q[0:3], q[3:6]
        pts = []
        A, B = np.meshgrid(u, u)
        for ax in range(3):
            o = [a for a in range(3) if a != ax]
            for side in (mn[ax], mx[ax]):
                p = np.zeros((A.size, 3))
                p[:, ax] = side
                p[:, o[0]] = mn[o[0]] + A.ravel() * (mx[o[0]] - mn[o[0]])
                p[:, o[1]] = mn[o[1]] + B.ravel() * (mx[o[1]] - mn[o[1]])
                pts.append(p)
        return np.concatenate(pts)
    ax = int(q[3] if kind == CYL else q[2])
    lo, hi = (q[4], q[5]) if kind == CYL else (q[5], q[6])
    rad = (lambda a: np.full_like(a, q[2])) if kind == CYL else \
        (lambda a: q[3] + (a - lo) / (hi - lo) * (q[4] - q[3]))
    oth = [a for a in range(3) if a != ax]
    A, T = np.meshgrid(lo + u * (hi - lo), th)
    Rr = rad(A)
    side = np.zeros((A.size, 3))
    side[:, ax] = A.ravel()
    side[:, oth[0]] = q[0] + (Rr * np.cos(T)).ravel()
    side[:, oth[1]] = q[1] + (Rr * np.sin(T)).ravel()
    caps = []
    for a0 in (lo, hi):
        rmax = rad(np.array([a0]))[0]
        S, T2 = np.meshgrid(u * rmax, th)
        c = np.zeros((S.size, 3))
        c[:, ax] = a0
        c[:, oth[0]] = q[0] + (S * np.cos(T2)).ravel()
        c[:, oth[1]] = q[1] + (S * np.sin(T2)).ravel()
        caps.append(c)
    return np.concatenate([side] + caps)


@pytest.mark.parametrize("kind,q", [
    prim(BLOCK, (0.1, -0.2, 0.0), (0.9, 0.5, 0.7)),
    prim(SPHERE, (0.1, 0.2, -0.1), 0.6),
    prim(CYL, (0.1, -0.1), 0.5, 2, -0.4, 0.6),
    prim(CYL, (0.0, 0.2), 0.4, 0, -0.5, 0.5),
    prim(CONE, (0.0, 0.1), 2, 0.6, 0.25, -0.3, 0.7),
    prim(CONE, (0.1, 0.0), 1, 0.2, 0.5, -0.4, 0.4),
])
def test_signed_distance_matches_brute_force(kind, q):
    samples = _surface_samples(kind, q)
    rng = np.random.default_rng(5)
    pts = rng.uniform(-1.2, 1.2, size=(150, 3))
    inside = _contains(kind, q, pts)
    o, k = region_oracle([(kind, q)])
    for x, ins in zip(pts, inside):
        d, p, n = o.region_surface(k, x)
        brute = np.min(np.linalg.norm(samples - x, axis=1))
        # |d| is the distance to the surface (the sampling is within ~1.5 %
        # of the extent), and its sign says inside / outside
        assert abs(abs(d) - brute) < 0.03
        assert (d <= 0) == bool(ins) or abs(d) < 1e-9
        if not ins and d > 1e-6:
            # outside: the normal points from the surface point to x
            assert np.allclose(n, (x - p) / np.linalg.norm(x - p), atol=1e-12)
            assert abs(np.linalg.norm(x - p) - d) < 1e-12


def test_union_and_intersect_pick_the_nearest_child_surface():
    a, b = prim(SPHERE, (0, 0, 0), 1.0), prim(BLOCK, (0.5, -2, -2), (3, 2, 2))
    for x in [(0.2, 0.1, 0), (0.9, 0, 0), (2.0, 0.5, 0.3), (-1.5, 0, 0)]:
        da = surf([a], x)[0]
        db = surf([b], x)[0]
        dab = surf([a, b], x)[0]
        assert dab == (da if abs(da) <= abs(db) else db)


def test_region_walls_push_like_the_dedicated_walls():
    # region_wall_force_matches_dedicated_sphere (lib.rs:2348); region_sphere_
    # inside_repels, region_block_inside_repels, region_cone_inside_repels
    r = 1e-3
    ctr = np.array([0.005, 0.005, 0.005])
    for off in ([0.0035, 0, 0], [0, -0.0032, 0.001], [0.002, 0.002, 0.0025]):
        x = ctr + np.array(off)
        o, _ = region_oracle([prim(SPHERE, ctr, 0.004)], inside=True, pos=x)
        o.wall_contact_force()
        fr = o.atoms()["force"][0]
        o2 = Oracle(MODEL_DEM)
        o2.set_box((-1, -1, -1), (1, 1, 1), (False, False, False))
        o2.set_materials([8.7e9], [0.3], [0.95], [0.4])
        o2.set_dt(1e-7)
        o2.set_atoms([x], [(0, 0, 0)], [dem_mass(2500.0, r)], [r], radius=[r],
                     omega=[(0, 0, 0)])
        o2.add_sphere_wall(ctr, 0.004, inside=True)
        o2.wall_contact_force()
        fs = o2.atoms()["force"][0]
        assert np.allclose(fr, fs, rtol=1e-12, atol=1e-18) and np.linalg.norm(fs) > 0
        assert np.dot(fr, ctr - x) > 0
    o, _ = region_oracle([prim(SPHERE, ctr, 0.004)], inside=True, pos=ctr)
    o.wall_contact_force()
    assert np.all(o.atoms()["force"] == 0.0)
    o, _ = region_oracle([prim(BLOCK, (0, 0, 0), (0.01, 0.01, 0.01))], inside=True,
                         pos=(0.005, 0.005, 0.0095))
    o.wall_contact_force()
    f = o.atoms()["force"][0]
    assert f[2] < 0 and f[0] == 0.0 and f[1] == 0.0
    o, _ = region_oracle([prim(CONE, (0.005, 0.005), 2, 0.004, 0.002, 0.0, 0.01)], inside=True,
                         pos=(0.005 + 0.0028, 0.005, 0.005), r=0.0005)
    o.wall_contact_force()
    assert o.atoms()["force"][0][0] < 0
