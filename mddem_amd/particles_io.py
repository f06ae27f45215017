"""Particle state in the reference's own file formats, on both sides of the
GPU path: CSV particle ingest, restart files (JSON or bincode) and the
granular-temperature log.

- CSV ingest follows `read_csv_particles` (crates/dem_atom/src/insert.rs:652-776)
  with `insert_single_particle` (insert.rs:323-355) for the derived per-particle
  values: configurable columns, `#` comments and blank lines skipped, a first
  line starting with a letter taken as a header, velocities defaulting to 0,
  a fixed radius when there is no radius column, tags assigned in file order,
  materials from an optional atom-type column through a type map.
- Restart files hold `RestartData` (crates/mddem_print/src/lib.rs:298-375):
  serde JSON with the struct's field names, or bincode 1 with its default
  options (little-endian, fixed-width integers, u64 sequence lengths, fields
  in declaration order). The DEM extension buffer packs `DemAtom`'s fields per
  particle in declaration order (crates/dem_atom/src/lib.rs:531-557, packing
  by the AtomData derive, crates/mddem_derive/src/lib.rs:304-346): radius,
  density, 1/I, quaternion (w, x, y, z), omega, angular momentum, torque,
  body id: 17 values.
- `GranularTemp.txt` lines follow print_granular_temperature
  (crates/dem_granular/src/granular_temp.rs:128-133).
"""
from __future__ import annotations

import json
import math
import struct

import numpy as np

PI = 3.141592653589793
DEM_PACK = 17  # f64 per particle in DemAtom's restart buffer


class FormatError(ValueError):
    """A file the reference would refuse (it prints the message and exits)."""


# ---------------------------------------------------------------------------
# CSV particle ingest
# ---------------------------------------------------------------------------

def _parse_f64(s: str) -> float:
    # Rust's f64::from_str: decimal or exponent forms, inf / nan; no
    # underscores, no surrounding text (the fields are already trimmed)
    if "_" in s or s == "":
        raise ValueError(s)
    return float(s)


def _as_u32(x: float) -> int:
    """Rust's `f64 as u32`: truncate toward zero, saturate, NaN -> 0."""
    if math.isnan(x) or x <= 0.0:
        return 0
    if x >= 4294967295.0:
        return 4294967295
    return int(x)


def read_csv_particles(path, *, density, material=0, columns=None, radius=None,
                       type_map=None, first_tag=0):
    """Particles of a CSV file, as the reference's file insertion creates them.

    columns: dict of column indices for x, y, z (default 0, 1, 2) and, when
    present, radius, vx, vy, vz, atom_type. radius: the fixed radius used when
    there is no radius column. type_map: {file atom type: material index},
    consulted when an atom_type column is given; types not in it get
    `material`. Returns a dict of arrays: tag, type, pos, vel, radius,
    cutoff_radius, mass, inv_inertia, density, quat, omega.
    """
    cols = dict(x=0, y=1, z=2, radius=None, vx=None, vy=None, vz=None, atom_type=None)
    if columns:
        unknown = set(columns) - set(cols)
        if unknown:
            raise FormatError(f"unknown CSV column keys {sorted(unknown)}")
        cols.update(columns)
    rows = []
    with open(path, "r") as fh:
        for line_num, line in enumerate(fh):
            line = line.rstrip("\n").rstrip("\r")
            t = line.strip()
            if not t or t.startswith("#"):
                continue
            if line_num == 0 and t[0].isalpha():
                continue
            fields = [f.strip() for f in t.split(",")]

            def parse(idx, name):
                try:
                    return _parse_f64(fields[idx])
                except (IndexError, ValueError):
                    raise FormatError(
                        f"Failed to parse {name} (column {idx}) at line {line_num + 1} "
                        f"of '{path}'") from None

            x, y, z = parse(cols["x"], "x"), parse(cols["y"], "y"), parse(cols["z"], "z")
            if cols["radius"] is not None:
                r = parse(cols["radius"], "radius")
            elif radius is not None:
                r = float(radius)
            else:
                raise FormatError(f"No radius column or default radius at line "
                                  f"{line_num + 1} of '{path}'")
            v = [parse(cols[k], k) if cols[k] is not None else 0.0 for k in ("vx", "vy", "vz")]
            mat = material
            if cols["atom_type"] is not None:
                ft = _as_u32(parse(cols["atom_type"], "atom_type"))
                if type_map is not None and ft in type_map:
                    mat = type_map[ft]
            rows.append((x, y, z, r, v[0], v[1], v[2], mat))
    n = len(rows)
    a = np.array([r[:7] for r in rows], dtype=np.float64).reshape(n, 7)
    rad = a[:, 3].copy()
    rho = np.full(n, float(density))
    # insert_single_particle: m = rho * 4/3 * pi * r^3, 1/I = 1 / (0.4 m r^2)
    mass = rho * 4.0 / 3.0 * PI * ((rad * rad) * rad)
    quat = np.zeros((n, 4))
    quat[:, 0] = 1.0
    return dict(tag=np.arange(first_tag, first_tag + n, dtype=np.uint32),
                type=np.array([r[7] for r in rows], dtype=np.uint32),
                pos=a[:, 0:3].copy(), vel=a[:, 4:7].copy(), radius=rad,
                cutoff_radius=rad.copy(), mass=mass,
                inv_inertia=1.0 / (0.4 * mass * rad * rad), density=rho, quat=quat,
                omega=np.zeros((n, 3)))


def write_csv_particles(path, pos, radius, vel=None, header=True):
    """x,y,z,radius,vx,vy,vz rows that read_csv_particles (default columns plus
    radius=3, vx..vz=4..6) reads back exactly (17 significant digits)."""
    pos = np.asarray(pos, dtype=np.float64)
    n = pos.shape[0]
    vel = np.zeros((n, 3)) if vel is None else np.asarray(vel, dtype=np.float64)
    radius = np.broadcast_to(np.asarray(radius, dtype=np.float64), (n,))
    with open(path, "w") as fh:
        if header:
            fh.write("x,y,z,radius,vx,vy,vz\n")
        for i in range(n):
            vals = (*pos[i], radius[i], *vel[i])
            fh.write(",".join(repr(float(v)) for v in vals) + "\n")


CSV_WRITTEN_COLUMNS = dict(x=0, y=1, z=2, radius=3, vx=4, vy=5, vz=6)


# ---------------------------------------------------------------------------
# RestartData
# ---------------------------------------------------------------------------

# (field, kind) in declaration order; kinds: u64, f64, vu32 (Vec<u32>),
# vf64 (Vec<f64>), vvf64 (Vec<Vec<f64>>), vq (Vec<[f64; 4]>)
_RESTART_FIELDS = [
    ("natoms", "u64"), ("total_cycle", "u64"), ("dt", "f64"),
    ("tag", "vu32"), ("atom_type", "vu32"),
    ("pos_x", "vf64"), ("pos_y", "vf64"), ("pos_z", "vf64"),
    ("vel_x", "vf64"), ("vel_y", "vf64"), ("vel_z", "vf64"),
    ("force_x", "vf64"), ("force_y", "vf64"), ("force_z", "vf64"),
    ("mass", "vf64"), ("cutoff_radius", "vf64"), ("atom_data_buffers", "vvf64"),
    ("omega_x", "vf64"), ("omega_y", "vf64"), ("omega_z", "vf64"),
    ("torque_x", "vf64"), ("torque_y", "vf64"), ("torque_z", "vf64"),
    ("ang_mom_x", "vf64"), ("ang_mom_y", "vf64"), ("ang_mom_z", "vf64"),
    ("quaternion", "vq"),
]
_LEGACY = {"omega_x", "omega_y", "omega_z", "torque_x", "torque_y", "torque_z", "ang_mom_x",
           "ang_mom_y", "ang_mom_z", "quaternion"}


def _empty_restart():
    d = {}
    for name, kind in _RESTART_FIELDS:
        if kind == "u64":
            d[name] = 0
        elif kind == "f64":
            d[name] = 0.0
        elif kind == "vq":
            d[name] = np.zeros((0, 4))
        elif kind == "vvf64":
            d[name] = []
        else:
            d[name] = np.zeros(0, np.uint32 if kind == "vu32" else np.float64)
    return d


def encode_bincode(data) -> bytes:
    """RestartData as bincode 1 (default options)."""
    out = bytearray()
    for name, kind in _RESTART_FIELDS:
        v = data[name]
        if kind == "u64":
            out += struct.pack("<Q", int(v))
        elif kind == "f64":
            out += struct.pack("<d", float(v))
        elif kind == "vu32":
            a = np.ascontiguousarray(v, dtype="<u4")
            out += struct.pack("<Q", a.size) + a.tobytes()
        elif kind == "vf64":
            a = np.ascontiguousarray(v, dtype="<f8")
            out += struct.pack("<Q", a.size) + a.tobytes()
        elif kind == "vvf64":
            out += struct.pack("<Q", len(v))
            for buf in v:
                a = np.ascontiguousarray(buf, dtype="<f8")
                out += struct.pack("<Q", a.size) + a.tobytes()
        else:  # vq
            a = np.ascontiguousarray(v, dtype="<f8").reshape(-1, 4)
            out += struct.pack("<Q", a.shape[0]) + a.tobytes()
    return bytes(out)


def decode_bincode(raw: bytes):
    d = {}
    off = 0

    def take(nbytes):
        nonlocal off
        if off + nbytes > len(raw):
            raise FormatError("restart bincode ends early")
        b = raw[off:off + nbytes]
        off += nbytes
        return b

    def length():
        return struct.unpack("<Q", take(8))[0]

    for name, kind in _RESTART_FIELDS:
        if kind == "u64":
            d[name] = struct.unpack("<Q", take(8))[0]
        elif kind == "f64":
            d[name] = struct.unpack("<d", take(8))[0]
        elif kind == "vu32":
            n = length()
            d[name] = np.frombuffer(take(4 * n), dtype="<u4").astype(np.uint32)
        elif kind == "vf64":
            n = length()
            d[name] = np.frombuffer(take(8 * n), dtype="<f8").astype(np.float64)
        elif kind == "vvf64":
            m = length()
            bufs = []
            for _ in range(m):
                n = length()
                bufs.append(np.frombuffer(take(8 * n), dtype="<f8").astype(np.float64))
            d[name] = bufs
        else:
            n = length()
            d[name] = np.frombuffer(take(32 * n), dtype="<f8").astype(np.float64).reshape(n, 4)
    if off != len(raw):
        raise FormatError(f"restart bincode has {len(raw) - off} trailing bytes")
    return d


def _json_value(v):
    if isinstance(v, np.ndarray):
        return v.tolist()
    if isinstance(v, list):
        return [_json_value(x) for x in v]
    return v


def encode_json(data) -> str:
    """RestartData as serde JSON (field names as keys, arrays as lists)."""
    return json.dumps({name: _json_value(data[name]) for name, _ in _RESTART_FIELDS})


def decode_json(text: str):
    raw = json.loads(text)
    d = _empty_restart()
    for name, kind in _RESTART_FIELDS:
        if name not in raw:
            if name in _LEGACY:
                continue  # #[serde(default)]
            raise FormatError(f"restart JSON lacks field '{name}'")
        v = raw[name]
        if kind in ("u64",):
            d[name] = int(v)
        elif kind == "f64":
            d[name] = float(v)
        elif kind == "vu32":
            d[name] = np.asarray(v, dtype=np.uint32)
        elif kind == "vf64":
            d[name] = np.asarray(v, dtype=np.float64)
        elif kind == "vvf64":
            d[name] = [np.asarray(b, dtype=np.float64) for b in v]
        else:
            d[name] = np.asarray(v, dtype=np.float64).reshape(-1, 4)
    return d


def write_restart(path, data, fmt="bincode"):
    if fmt == "json":
        with open(path, "w") as fh:
            fh.write(encode_json(data))
    elif fmt == "bincode":
        with open(path, "wb") as fh:
            fh.write(encode_bincode(data))
    else:
        raise FormatError(f"unknown restart format '{fmt}'")


def read_restart(path, fmt=None):
    """fmt: 'json' or 'bincode'; by default from the extension as the
    reference names its files (restart_<step>_rank<r>.json / .bin)."""
    if fmt is None:
        fmt = "json" if str(path).endswith(".json") else "bincode"
    if fmt == "json":
        with open(path, "r") as fh:
            return decode_json(fh.read())
    with open(path, "rb") as fh:
        return decode_bincode(fh.read())


def restart_file_name(step, rank=0, fmt="bincode"):
    """restart_<step>_rank<rank>.<json|bin> (mddem_print/src/lib.rs:1177-1199)."""
    return f"restart_{step}_rank{rank}.{'json' if fmt == 'json' else 'bin'}"


def pack_dem(radius, density, inv_inertia, quat, omega, ang_mom=None, torque=None,
             body_id=None):
    """DemAtom's restart buffer: 17 values per particle in field order."""
    n = len(radius)
    z3 = np.zeros((n, 3))
    cols = [np.asarray(radius, np.float64)[:, None], np.asarray(density, np.float64)[:, None],
            np.asarray(inv_inertia, np.float64)[:, None], np.asarray(quat, np.float64),
            np.asarray(omega, np.float64), z3 if ang_mom is None else np.asarray(ang_mom),
            z3 if torque is None else np.asarray(torque),
            np.zeros((n, 1)) if body_id is None else np.asarray(body_id)[:, None]]
    return np.concatenate(cols, axis=1).reshape(-1)


def unpack_dem(buf):
    b = np.asarray(buf, dtype=np.float64)
    if b.size % DEM_PACK:
        raise FormatError(f"DemAtom restart buffer of {b.size} values is not a multiple of 17")
    a = b.reshape(-1, DEM_PACK)
    return dict(radius=a[:, 0].copy(), density=a[:, 1].copy(), inv_inertia=a[:, 2].copy(),
                quat=a[:, 3:7].copy(), omega=a[:, 7:10].copy(), ang_mom=a[:, 10:13].copy(),
                torque=a[:, 13:16].copy(), body_id=a[:, 16].copy())


def restart_from_system(s, step, dt, density=None):
    """A RestartData snapshot of a hip.System (particles in tag order).
    density: DEM particle density (scalar or per particle in tag order)."""
    st = s.state()
    order = np.argsort(st["tag"])
    f, t = s.forces()
    typ, mass, radii, cutr = s.properties()
    d = _empty_restart()
    n = len(order)
    d.update(natoms=n, total_cycle=int(step), dt=float(dt), tag=st["tag"][order],
             atom_type=typ[order], mass=mass[order], cutoff_radius=cutr[order])
    for k, name in enumerate(("x", "y", "z")):
        d[f"pos_{name}"] = st["pos"][order, k]
        d[f"vel_{name}"] = st["vel"][order, k]
        d[f"force_{name}"] = f[order, k]
    if "omega" in st:
        rad = radii[order]
        m = mass[order]
        rho = (m / (4.0 / 3.0 * PI * rad ** 3) if density is None
               else np.broadcast_to(np.asarray(density, np.float64), (n,)))
        d["atom_data_buffers"] = [pack_dem(rad, rho, 1.0 / (0.4 * m * rad * rad),
                                           st["quat"][order], st["omega"][order],
                                           torque=t[order])]
    return d


def upload_restart(s, data):
    """Load a RestartData snapshot into a configured hip.System (box,
    neighbor settings, materials and integrators set by the caller), as
    read_restart repopulates Atom and DemAtom (mddem_print/src/lib.rs:1222-1246):
    state, forces (the next step's first half kick uses them) and, for DEM,
    radius, omega, quaternion and torque from the DemAtom buffer."""
    n = len(data["tag"])
    pos = np.stack([data["pos_x"], data["pos_y"], data["pos_z"]], axis=1)
    vel = np.stack([data["vel_x"], data["vel_y"], data["vel_z"]], axis=1)
    force = np.stack([data["force_x"], data["force_y"], data["force_z"]], axis=1)
    dem = None
    if data["atom_data_buffers"]:
        dem = unpack_dem(data["atom_data_buffers"][0])
        if len(dem["radius"]) != n:
            raise FormatError("DemAtom restart buffer does not match the particle count")
    s.upload(pos, vel, np.asarray(data["mass"]), np.asarray(data["cutoff_radius"]),
             radius=None if dem is None else dem["radius"],
             omega=None if dem is None else dem["omega"], tag=np.asarray(data["tag"]),
             atype=np.asarray(data["atom_type"]))
    s.set_forces(force, None if dem is None else dem["torque"])
    if dem is not None:
        s.upload_quaternions(dem["quat"])
    return s


# ---------------------------------------------------------------------------
# GranularTemp.txt
# ---------------------------------------------------------------------------

def rust_exp(x: float, prec: int) -> str:
    """Rust's `{:.<prec>e}`: mantissa with `prec` decimals, exponent without a
    sign for non-negative values and without zero padding."""
    s = f"{float(x):.{prec}e}"
    if "e" not in s:  # inf / nan
        return {"inf": "inf", "-inf": "-inf", "nan": "NaN"}.get(s, s)
    mant, exp = s.split("e")
    return f"{mant}e{int(exp)}"


def granular_temp_line(step, dt, granular_temperature, ke, momentum):
    """One GranularTemp.txt line: step, physical time, T_g, KE, |p|."""
    px, py, pz = momentum
    mom = math.sqrt(px * px + py * py + pz * pz)
    return (f"{step} {rust_exp(step * dt, 6)} {rust_exp(granular_temperature, 10)} "
            f"{rust_exp(ke, 10)} {rust_exp(mom, 10)}")
