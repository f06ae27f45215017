"""Seeded synthetic inputs for the particle hot path.

The reference draws every initial condition from an unseeded `rand::rng()`
(crates/dem_atom/src/insert.rs:446, crates/md_lattice/src/lib.rs:372), so the
same state can only reach it through files. These generators produce fixed
states from a documented counter-based PRNG (SplitMix64 over
`seed + index`), reproduce the reference's deterministic geometry (FCC
placement, mass / inertia / Rayleigh-time formulas) and hand the arrays to
both the HIP path and the CPU checker.

Seeds: workload k uses `SEED_BASE + k` unless the caller passes another.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

SEED_BASE = 0x4D44_4445_4D00

PI = 3.141592653589793
# glass (mddem_test_utils/src/lib.rs:170-175; granular_gas_benchmark)
GLASS = dict(youngs=8.7e9, poisson=0.3, restitution=0.95, friction=0.4, density=2500.0)

_MASK64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def splitmix64(x: np.ndarray) -> np.ndarray:
    """SplitMix64 finaliser applied elementwise to uint64 counters."""
    with np.errstate(over="ignore"):
        z = (x + np.uint64(0x9E3779B97F4A7C15)) & _MASK64
        z = ((z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)) & _MASK64
        z = ((z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)) & _MASK64
        return z ^ (z >> np.uint64(31))


class Stream:
    """Counter-based uniform / normal stream: value k = f(seed, k)."""

    def __init__(self, seed: int):
        self.seed = np.uint64(seed & 0xFFFFFFFFFFFFFFFF)
        self.counter = 0

    def uniform(self, n: int) -> np.ndarray:
        with np.errstate(over="ignore"):
            ctr = (np.arange(self.counter, self.counter + n, dtype=np.uint64)
                   * np.uint64(0xD1B54A32D192ED03) + self.seed) & _MASK64
        self.counter += n
        u = splitmix64(ctr) >> np.uint64(11)
        return (u.astype(np.float64) + 0.5) * (1.0 / 9007199254740992.0)

    def normal(self, n: int) -> np.ndarray:
        """Box-Muller, one deviate per pair of uniforms."""
        u1 = self.uniform(n)
        u2 = self.uniform(n)
        return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * PI * u2)


def dem_mass(density, radius):
    """m = rho * 4/3 * pi * r^3 (dem_atom/src/insert.rs:340), same evaluation order."""
    radius = np.asarray(radius, dtype=np.float64)
    return density * 4.0 / 3.0 * PI * ((radius * radius) * radius)


def rayleigh_time(youngs, poisson, density, radius):
    """Per-particle Rayleigh time pi r / alpha sqrt(rho / G) (insert.rs:1411-1427)."""
    g = youngs / (2.0 * (1.0 + poisson))
    alpha = 0.1631 * poisson + 0.876605
    return PI * np.asarray(radius, dtype=np.float64) / alpha * np.sqrt(density / g)


@dataclass
class Workload:
    """A complete single-system input: box, neighbor settings and atoms."""

    name: str
    model: str                      # "lj" or "dem"
    lo: np.ndarray
    hi: np.ndarray
    periodic: tuple
    pos: np.ndarray                 # (n, 3) float64
    vel: np.ndarray                 # (n, 3) float64
    mass: np.ndarray                # (n,)
    cutoff_radius: np.ndarray       # (n,) neighbor skin radius
    dt: float
    skin_fraction: float
    bin_size: float
    radius: np.ndarray | None = None
    omega: np.ndarray | None = None
    atype: np.ndarray | None = None
    materials: dict | None = None   # per-material arrays for DEM
    lj: tuple | None = None         # (epsilon, sigma, cutoff)
    nose_hoover: tuple | None = None  # (temperature, coupling)
    gravity: tuple | None = None
    walls: list = field(default_factory=list)  # (point, normal, material)
    steps: int = 0
    density: float | None = None
    every: int = 0                  # [neighbor] every (0: displacement rebuilds only)

    @property
    def n(self) -> int:
        return int(self.pos.shape[0])

    @property
    def tag(self) -> np.ndarray:
        return np.arange(self.n, dtype=np.uint32)


# ---------------------------------------------------------------------------
# LJ fluid on an FCC lattice (crates/md_lattice/src/lib.rs:232-454)
# ---------------------------------------------------------------------------

def fcc_lattice(ncell: int, rho: float, lo=0.0):
    """FCC sites for an `ncell`^3 block at number density rho.

    Geometry follows fcc_insert: a = (4/rho)^(1/3), L = ncell * a placed
    directly (so floor(L/a) cannot fall to ncell-1), a_x = L / n, loop order
    ix, iy, iz, basis (lib.rs:336-360)."""
    a_ideal = (4.0 / rho) ** (1.0 / 3.0)
    L = ncell * a_ideal
    ax = L / ncell
    basis = np.array([[0.0, 0.0, 0.0], [0.5, 0.5, 0.0], [0.5, 0.0, 0.5], [0.0, 0.5, 0.5]])
    ix, iy, iz = np.meshgrid(np.arange(ncell), np.arange(ncell), np.arange(ncell), indexing="ij")
    cells = np.stack([ix.ravel(), iy.ravel(), iz.ravel()], axis=1).astype(np.float64)
    pos = (cells[:, None, :] + basis[None, :, :]).reshape(-1, 3)
    pos = lo + pos * ax
    return pos, L


def maxwell_boltzmann(n: int, temperature: float, mass: float, stream: Stream) -> np.ndarray:
    """Gaussian velocities, COM removed, rescaled to T with ndof = 3N-3
    (md_lattice/src/lib.rs:369-431)."""
    sigma_v = math.sqrt(temperature / mass)
    v = stream.normal(3 * n).reshape(n, 3) * sigma_v
    v -= v.mean(axis=0)
    ndof = 3.0 * n - 3.0
    ke = 0.5 * float(np.sum(mass * np.sum(v * v, axis=1)))
    cur = 2.0 * ke / ndof
    if cur > 1e-20:
        v *= math.sqrt(temperature / cur)
    return v


def lj_fcc(ncell: int = 40, rho: float = 0.85, temperature: float = 0.85,
           seed: int | None = None, thermostat: bool = True, steps: int = 20000) -> Workload:
    """LJ fluid, FCC start, eps = sigma = m = 1, rc = 2.5.

    Neighbor cutoff 2.8 comes from cutoff_radius 1.25 with skin_fraction 1.12
    (examples/lj_argon/config.toml), which gives 2.8000000000000003 and a
    per-atom displacement budget of 0.15."""
    seed = SEED_BASE + 1 if seed is None else seed
    pos, L = fcc_lattice(ncell, rho)
    n = pos.shape[0]
    vel = maxwell_boltzmann(n, temperature, 1.0, Stream(seed))
    return Workload(
        name=f"lj_fcc_{n}", model="lj", lo=np.zeros(3), hi=np.full(3, L),
        periodic=(True, True, True), pos=pos, vel=vel, mass=np.ones(n),
        cutoff_radius=np.full(n, 1.25), dt=0.005, skin_fraction=1.12, bin_size=1.0,
        lj=(1.0, 1.0, 2.5), nose_hoover=(temperature, 1.0) if thermostat else None,
        steps=steps)


# ---------------------------------------------------------------------------
# DEM granular gas (examples/granular_gas_benchmark) and settling
# ---------------------------------------------------------------------------

def _min_image_d2(p, q, L):
    d = p - q
    d -= L * np.round(d / L)
    return np.sum(d * d, axis=-1)


def _rsa_periodic(n, radius, lo, hi, stream, max_tries=2_000_000):
    """Sequential addition with min-image separation > 1.1 (r_i + r_j) inside
    the box inset by the largest radius (insert.rs:530-560). O(N^2); small N."""
    L = hi - lo
    rmax = float(np.max(radius))
    pts = np.zeros((n, 3))
    k = 0
    tries = 0
    while k < n:
        u = stream.uniform(3)
        tries += 1
        if tries > max_tries:
            raise RuntimeError("random sequential addition did not converge")
        c = lo + rmax + u * (L - 2 * rmax)
        if k:
            d2 = _min_image_d2(pts[:k], c, L)
            md = (radius[k] + radius[:k]) * 1.1
            if np.any(d2 <= md * md):
                continue
        pts[k] = c
        k += 1
    return pts


def _isotropic(n, speed, stream):
    v = stream.normal(3 * n).reshape(n, 3)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    return v * speed


def glass_materials(n_mat=1, **over):
    m = dict(GLASS)
    m.update(over)
    return {k: np.full(n_mat, float(m[k])) for k in ("youngs", "poisson", "restitution", "friction")}


def dem_gas(n: int = 500, box: float = 0.025, radius: float = 1e-3, speed: float = 0.5,
            dt_factor: float = 0.05, seed: int | None = None, steps: int = 10000,
            restitution: float = 0.95, friction: float = 0.4) -> Workload:
    """Periodic granular gas of monodisperse glass spheres."""
    seed = SEED_BASE + 0 if seed is None else seed
    st = Stream(seed)
    r = np.full(n, radius)
    lo, hi = np.zeros(3), np.full(3, box)
    if n <= 4000:
        pos = _rsa_periodic(n, r, lo, hi, st)
    else:
        pos = _jittered_lattice(n, r, lo, hi, st)
    vel = _isotropic(n, speed, st)
    dens = GLASS["density"]
    t_r = float(np.min(rayleigh_time(GLASS["youngs"], GLASS["poisson"], dens, r)))
    return Workload(
        name=f"dem_gas_{n}", model="dem", lo=lo, hi=hi, periodic=(True, True, True),
        pos=pos, vel=vel, mass=dem_mass(dens, r), cutoff_radius=r.copy(), radius=r,
        omega=np.zeros((n, 3)), atype=np.zeros(n, np.uint32), dt=dt_factor * t_r,
        skin_fraction=1.1, bin_size=0.005,
        materials=glass_materials(restitution=restitution, friction=friction),
        steps=steps, density=dens)


def gas_system(n: int, phi: float, restitution: float, friction: float, seed: int,
               radius: float = 1e-3, speed: float = 0.5) -> Workload:
    """One system of the granular-gas ensemble: a periodic cube of edge
    L = (N 4/3 pi r^3 / phi)^(1/3), seeded lattice start, speed 0.5 m/s."""
    st = Stream(seed)
    L = (n * 4.0 / 3.0 * np.pi * radius ** 3 / phi) ** (1.0 / 3.0)
    r = np.full(n, radius)
    lo, hi = np.zeros(3), np.full(3, L)
    pos = _fcc_sites(n, L, radius, st)
    vel = _isotropic(n, speed, st)
    dens = GLASS["density"]
    t_r = float(np.min(rayleigh_time(GLASS["youngs"], GLASS["poisson"], dens, r)))
    return Workload(
        name=f"gas_{n}_{seed & 0xffff}", model="dem", lo=lo, hi=hi, periodic=(True, True, True),
        pos=pos, vel=vel, mass=dem_mass(dens, r), cutoff_radius=r.copy(), radius=r,
        omega=np.zeros((n, 3)), atype=np.zeros(n, np.uint32), dt=0.05 * t_r,
        skin_fraction=1.1, bin_size=0.005,
        materials=glass_materials(restitution=restitution, friction=friction), density=dens)


def _fcc_sites(n, L, radius, stream):
    """n non-overlapping spheres in a periodic cube: n sites drawn from the
    FCC lattice (m_x x m_y x m_z cells tiling the box) with the widest
    nearest-neighbor spacing, each jittered by less than half the slack."""
    m = int(math.ceil((n / 4.0) ** (1.0 / 3.0)))
    best = None
    for mx in (m - 1, m):
        for my in (m - 1, m):
            for mz in (m - 1, m):
                if min(mx, my, mz) < 1 or 4 * mx * my * mz < n:
                    continue
                ax, ay, az = L / mx, L / my, L / mz
                nn = min(ax, ay, az, math.hypot(ax, ay) / 2, math.hypot(ax, az) / 2,
                         math.hypot(ay, az) / 2)
                if best is None or nn > best[0]:
                    best = (nn, mx, my, mz)
    nn, mx, my, mz = best
    slack = nn - 2.002 * radius
    if slack <= 0:
        raise ValueError("volume fraction too high for lattice placement")
    basis = np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]])
    cells = np.stack(np.meshgrid(np.arange(mx), np.arange(my), np.arange(mz), indexing="ij"),
                     axis=-1).reshape(-1, 3)
    sites = (cells[:, None, :] + basis[None, :, :]).reshape(-1, 3) * np.array([L / mx, L / my, L / mz])
    keys = stream.uniform(sites.shape[0])
    pick = np.sort(np.argsort(keys, kind="stable")[:n])
    jitter = (stream.uniform(3 * n).reshape(n, 3) - 0.5) * (slack / math.sqrt(3.0))
    return np.mod(sites[pick] + jitter + 0.25 * np.array([L / mx, L / my, L / mz]), L)


def ensemble(nsys: int = 1024, seed: int | None = None, nmin: int = 200,
             nmax: int = 20000) -> list:
    """The batched granular-gas ensemble: N log-uniform in [nmin, nmax],
    restitution U[0.7, 0.99], friction U[0.1, 0.6], volume fraction
    U[0.05, 0.30], r = 1 mm; system k is seeded with base + k."""
    base = SEED_BASE + 3 if seed is None else seed
    u = Stream(base).uniform(4 * nsys).reshape(nsys, 4)
    out = []
    for k in range(nsys):
        n = int(round(math.exp(math.log(nmin) + u[k, 0] * math.log(nmax / nmin))))
        out.append(gas_system(n, 0.05 + 0.25 * u[k, 3], 0.7 + 0.29 * u[k, 1],
                              0.1 + 0.5 * u[k, 2], seed=(base << 12) + k))
    return out


def _jittered_lattice(n, radius, lo, hi, stream, gap=1.1):
    """n sites on the coarsest cubic-ish lattice that fits the box, each site
    jittered inside the slack left by gap * 2 * r_max; sites are drawn
    without replacement so the packing stays uniform."""
    L = hi - lo
    rmax = float(np.max(radius))
    vol_per = float(np.prod(L)) / n
    h = vol_per ** (1.0 / 3.0)
    while True:
        counts = np.maximum(np.floor(L / h).astype(int), 1)
        if int(np.prod(counts)) >= n:
            break
        h *= 0.995
    spacing = L / counts
    slack = float(np.min(spacing)) - gap * 2.0 * rmax
    if slack <= 0:
        raise ValueError("box too small for non-overlapping lattice placement")
    total = int(np.prod(counts))
    keys = stream.uniform(total)
    sites = np.sort(np.argsort(keys, kind="stable")[:n])
    ix = sites // (counts[1] * counts[2])
    iy = (sites // counts[2]) % counts[1]
    iz = sites % counts[2]
    centers = lo + (np.stack([ix, iy, iz], axis=1) + 0.5) * spacing
    jitter = (stream.uniform(3 * n).reshape(n, 3) - 0.5) * slack
    return centers + jitter


def dem_settling(n: int = 4_000_000, box=(0.4, 0.4, 1.2), rmin: float = 0.8e-3,
                 rmax: float = 1.2e-3, vmax: float = 0.1, dt_factor: float = 0.05,
                 seed: int | None = None, steps: int = 50000) -> Workload:
    """Polydisperse glass spheres settling under gravity between two plane
    walls (z = 0, n = +z and z = Lz, n = -z); periodic in x and y."""
    seed = SEED_BASE + 2 if seed is None else seed
    st = Stream(seed)
    lo, hi = np.zeros(3), np.asarray(box, dtype=np.float64)
    r = rmin + st.uniform(n) * (rmax - rmin)
    pos = _jittered_lattice(n, r, lo, hi, st)
    vel = (st.uniform(3 * n).reshape(n, 3) * 2.0 - 1.0) * vmax
    dens = GLASS["density"]
    t_r = float(np.min(rayleigh_time(GLASS["youngs"], GLASS["poisson"], dens, r)))
    walls = [((0.0, 0.0, 0.0), (0.0, 0.0, 1.0), 0), ((0.0, 0.0, hi[2]), (0.0, 0.0, -1.0), 0)]
    return Workload(
        name=f"dem_settling_{n}", model="dem", lo=lo, hi=hi, periodic=(True, True, False),
        pos=pos, vel=vel, mass=dem_mass(dens, r), cutoff_radius=r.copy(), radius=r,
        omega=np.zeros((n, 3)), atype=np.zeros(n, np.uint32), dt=dt_factor * t_r,
        skin_fraction=1.1, bin_size=0.0, materials=glass_materials(),
        gravity=(0.0, 0.0, -9.81), walls=walls, steps=steps, density=dens)


def _wrap2(d, L2):
    return d - L2 * np.round(d / L2)


def _deposited_bed(n, Lx, Ly, rmin, rmax, stream, overlap=5e-8, gap=1.2, look=4):
    """A resting polydisperse bed of n spheres on the plane z = 0, periodic in
    x and y, built layer by layer with no dynamics.

    Layer k sits on the sites of a triangular lattice of pitch gap * 2 r_max
    shifted by k hollows (A, B, C stacking). Each sphere is first placed
    touching its three supporting spheres of the layer below (the upper
    intersection of three spheres of radius r + r_j - overlap), or, where
    that point does not exist or lies outside their reach, on its lattice
    site resting on the highest supporter. Any sphere then closer than
    r_i + r_j - 2 overlap to an earlier sphere of its layer or to a sphere of
    the `look` layers below is raised until it rests on it, which can only
    lift it; the loop ends when nothing is raised. Layer 0 rests on the
    floor. Every contact therefore overlaps by `overlap` (to within
    rounding) and no pair overlaps by more than 2 * overlap."""
    a = 2.0 * rmax * gap
    nx = int(Lx // a)
    ny = int(Ly // (a * math.sqrt(3.0) / 2.0)) // 2 * 2
    if nx < 3 or ny < 4:
        raise ValueError("box too small for a deposited bed")
    ax, hy = Lx / nx, Ly / ny
    ii, rr = np.meshgrid(np.arange(nx), np.arange(ny), indexing="xy")
    base = np.stack([(ii * ax + (rr % 2) * ax / 2.0).ravel(), (rr * hy).ravel()], axis=1)
    m = base.shape[0]
    L2 = np.array([Lx, Ly])
    hollow = np.array([ax / 2.0, hy / 3.0])
    from scipy.spatial import cKDTree
    tree = cKDTree(base, boxsize=[Lx, Ly])
    _, sup = tree.query(np.mod(base + hollow, L2), k=3)
    smax = math.hypot(ax / 2.0, hy / 3.0)  # distance from a hollow to its supporters
    reach = 2.0 * rmax + 2.0 * smax + 1e-12
    # candidates of site s among the sites of the layer q below (q = 0: the
    # earlier sites of its own layer); the stacking makes them the same for
    # every layer
    cand = []
    for q in range(look + 1):
        lists = tree.query_ball_point(np.mod(base + q * hollow, L2), reach)
        width = max(len(x) for x in lists)
        C = np.full((m, width), -1, np.int64)
        for s, x in enumerate(lists):
            x = [t for t in x if q > 0 or t < s]
            C[s, :len(x)] = x
        cand.append(C)
    layers_p, layers_r = [], []
    placed = 0
    layer = 0
    while placed < n:
        xy = np.mod(base + layer * hollow, L2)
        r = rmin + stream.uniform(m) * (rmax - rmin)
        P = np.column_stack([xy, r - overlap])
        if layer:
            pp, pr = layers_p[-1], layers_r[-1]
            rel = pp[sup].copy()
            rel[:, :, :2] = _wrap2(rel[:, :, :2] - xy[:, None, :], L2)
            d = r[:, None] + pr[sup] - overlap
            p1, p2, p3 = rel[:, 0], rel[:, 1], rel[:, 2]
            ex = p2 - p1
            dd = np.linalg.norm(ex, axis=1)
            ex /= dd[:, None]
            t = p3 - p1
            ti = np.sum(ex * t, axis=1)
            ey = t - ti[:, None] * ex
            tj = np.linalg.norm(ey, axis=1)
            ey /= tj[:, None]
            ez = np.cross(ex, ey)
            ez *= np.sign(ez[:, 2:3])
            x = (d[:, 0] ** 2 - d[:, 1] ** 2 + dd ** 2) / (2.0 * dd)
            y = (d[:, 0] ** 2 - d[:, 2] ** 2 + ti ** 2 + tj ** 2) / (2.0 * tj) - (ti / tj) * x
            z2 = d[:, 0] ** 2 - x ** 2 - y ** 2
            sol = p1 + x[:, None] * ex + y[:, None] * ey + np.sqrt(np.maximum(z2, 0.0))[:, None] * ez
            tri = ((z2 > 0) & np.all(sol[:, None, 2] > rel[:, :, 2], axis=1)
                   & (np.hypot(sol[:, 0], sol[:, 1]) <= smax))
            rho2 = np.sum(rel[:, :, :2] ** 2, axis=2)
            P[:, 2] = np.max(rel[:, :, 2] + np.sqrt(np.maximum(d * d - rho2, 0.0)), axis=1)
            P[tri] = np.column_stack([np.mod(xy[tri] + sol[tri, :2], L2), sol[tri, 2]])
        rows = np.arange(m)
        for _ in range(1000):
            znew = P[rows, 2].copy()
            for q in range(min(look, layer) + 1):
                C = cand[q][rows]
                Q, R = (P, r) if q == 0 else (layers_p[-q], layers_r[-q])
                valid = C >= 0
                Cc = np.where(valid, C, 0)
                Qc = Q[Cc]
                dv = P[rows, None, :] - Qc
                dv[:, :, :2] = _wrap2(dv[:, :, :2], L2)
                dk = r[rows, None] + R[Cc] - overlap
                rho2 = dv[:, :, 0] ** 2 + dv[:, :, 1] ** 2
                bad = valid & (rho2 + dv[:, :, 2] ** 2 < (dk - overlap) ** 2)
                if bad.any():
                    zr = np.where(bad, Qc[:, :, 2] + np.sqrt(np.maximum(dk * dk - rho2, 0.0)),
                                  -np.inf)
                    znew = np.maximum(znew, zr.max(axis=1))
            moved = znew > P[rows, 2]
            if not moved.any():
                break
            P[rows, 2] = znew
            raised = np.zeros(m, bool)
            raised[rows[moved]] = True
            c0 = cand[0]
            rows = np.nonzero(raised | ((c0 >= 0) & raised[np.where(c0 >= 0, c0, 0)]).any(axis=1))[0]
        else:
            raise RuntimeError("deposited bed did not settle its conflicts")
        layers_p.append(P)
        layers_r.append(r)
        placed += m
        layer += 1
    return np.concatenate(layers_p)[:n], np.concatenate(layers_r)[:n]


def dem_packed_bed(n: int = 4_000_000, box=(0.4, 0.4, 1.2), rmin: float = 0.8e-3,
                   rmax: float = 1.2e-3, vmax: float = 0.0, overlap: float = 1e-10,
                   dt_factor: float = 0.05, seed: int | None = None, steps: int = 50000,
                   every: int = 0) -> Workload:
    """The settling workload's spheres, materials, walls and gravity, started
    from a bed resting on the floor (a deposited packing, `_deposited_bed`)
    instead of a dilute cloud: every sphere touches the spheres it rests on,
    each contact overlapping by `overlap`. The default, 0.1 nm, carries
    about 1/700 of a sphere's weight, so the bed starts touching and gravity
    loads its contacts (a 10 nm overlap carries about 1.4 weights and the
    bed springs apart). Velocities are uniform in [-vmax, vmax] per
    component (0: at rest). every: the [neighbor] every of the runs."""
    seed = SEED_BASE + 5 if seed is None else seed
    st = Stream(seed)
    lo, hi = np.zeros(3), np.asarray(box, dtype=np.float64)
    pos, r = _deposited_bed(n, hi[0], hi[1], rmin, rmax, st, overlap=overlap)
    if pos[:, 2].max() + rmax >= hi[2]:
        raise ValueError("bed higher than the box")
    vel = (st.uniform(3 * n).reshape(n, 3) * 2.0 - 1.0) * vmax
    dens = GLASS["density"]
    t_r = float(rayleigh_time(GLASS["youngs"], GLASS["poisson"], dens, rmin))
    walls = [((0.0, 0.0, 0.0), (0.0, 0.0, 1.0), 0), ((0.0, 0.0, hi[2]), (0.0, 0.0, -1.0), 0)]
    return Workload(
        name=f"dem_bed_{n}", model="dem", lo=lo, hi=hi, periodic=(True, True, False),
        pos=pos, vel=vel, mass=dem_mass(dens, r), cutoff_radius=r.copy(), radius=r,
        omega=np.zeros((n, 3)), atype=np.zeros(n, np.uint32), dt=dt_factor * t_r,
        skin_fraction=1.1, bin_size=0.0, materials=glass_materials(),
        gravity=(0.0, 0.0, -9.81), walls=walls, steps=steps, density=dens, every=every)
