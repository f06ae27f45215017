"""ctypes bindings of libmddem_hip.so (the C ABI in include/mddem_hip.h).

`System` mirrors the reference's plugin systems one method each (names as in
the reference: neighbor_setup, bin_neighbor_list, lj_force,
hertz_mindlin_contact_force, initial_integration, nh_pre_initial, ...) and
`run(n)` is the scheduled step loop. Every call goes to the HIP library; if
the library is missing or the GPU is not a gfx950 the calls raise.
"""
from __future__ import annotations

import ctypes as C
import math
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmddem_hip.so")

MODEL_LJ, MODEL_DEM = 0, 1
FP64, MIXED = 0, 1
# optional per-material contact properties, in the order mdd_set_material_extras takes
MATERIAL_EXTRAS = ("rolling_friction", "twisting_friction", "cohesion_energy", "surface_energy",
                   "kn", "kt", "rolling_stiffness", "rolling_damping", "twisting_stiffness",
                   "twisting_damping")
ERRORS = {-1: "invalid", -2: "hip", -3: "overlap", -4: "capacity", -5: "unsupported",
          -6: "no gfx950 device", -7: "a step kept asking for a rebuild"}

_dp = C.POINTER(C.c_double)
_u32p = C.POINTER(C.c_uint32)
_i32p = C.POINTER(C.c_int32)
_u8p = C.POINTER(C.c_uint8)
_i64p = C.POINTER(C.c_int64)


_AXES = {"x": 0, "y": 1, "z": 2}


def flatten_region(region):
    """Leaf primitives of a region tree (Region, mddem_core/src/region.rs:
    40-78) in depth-first order, as (types, params[n, 12]) for
    mdd_add_region_wall. A region is a dict like the reference's TOML:
    {"type": "block", "min": [..], "max": [..]}, "sphere" (center, radius),
    "cylinder" (center[2], radius, axis, lo, hi), "cone" (center[2], axis,
    rad_lo, rad_hi, lo, hi), "plane" (point, normal), "union" / "intersect"
    (regions: [...]); both combinations take the child surface nearest the
    point, so a tree and its leaf list give the same surface."""
    types, rows = [], []

    def axis(r):
        a = r.get("axis", "z")
        return float(_AXES[a.lower()] if isinstance(a, str) else a)

    def walk(r):
        t = r["type"].lower()
        q = np.zeros(12)
        if t in ("union", "intersect"):
            if not r["regions"]:
                raise ValueError(f"{t} needs at least one region")
            for child in r["regions"]:
                walk(child)
            return
        if t == "block":
            q[0:3], q[3:6] = r["min"], r["max"]
            types.append(0)
        elif t == "sphere":
            q[0:3], q[3] = r["center"], r["radius"]
            types.append(1)
        elif t == "cylinder":
            q[0:2], q[2], q[3], q[4], q[5] = r["center"], r["radius"], axis(r), r["lo"], r["hi"]
            types.append(2)
        elif t == "cone":
            q[0:2], q[2] = r["center"], axis(r)
            q[3], q[4], q[5], q[6] = r["rad_lo"], r["rad_hi"], r["lo"], r["hi"]
            types.append(3)
        elif t == "plane":
            q[0:3], q[3:6] = r["point"], r["normal"]
            types.append(4)
        else:
            raise ValueError(f"unknown region type {t!r}")
        rows.append(q)

    walk(region)
    if len(types) > 8:
        raise ValueError("a region wall takes at most 8 primitives")
    return types, np.array(rows)


class MddError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class Params(C.Structure):
    _fields_ = [("device", C.c_int), ("model", C.c_int), ("precision", C.c_int),
                ("capacity", C.c_int64)]


class Thermo(C.Structure):
    _fields_ = [("step", C.c_int64), ("rebuilds", C.c_int64), ("npairs", C.c_int64),
                ("nghost", C.c_int64), ("ncontacts", C.c_int64), ("ke", C.c_double),
                ("temperature", C.c_double), ("pe", C.c_double), ("p_xi", C.c_double),
                ("momentum", C.c_double * 3), ("granular_temperature", C.c_double),
                ("virial", C.c_double * 6), ("pressure", C.c_double)]

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_}
        d["momentum"] = list(self.momentum)
        d["virial"] = list(self.virial)
        return d


_lib = None
_vp = C.c_void_p

SIGNATURES = {
    "mdd_create": (C.c_int, [C.POINTER(_vp), C.POINTER(Params)]),
    "mdd_destroy": (C.c_int, [_vp]),
    "mdd_last_error": (C.c_char_p, [_vp]),
    "mdd_version": (C.c_int, []),
    "mdd_set_box": (C.c_int, [_vp, _dp, _dp, C.POINTER(C.c_int)]),
    "mdd_set_neighbor": (C.c_int, [_vp, C.c_double, C.c_double, C.c_int, C.c_int, C.c_int,
                                   C.c_int]),
    "mdd_set_rebuild_on_box_exit": (C.c_int, [_vp, C.c_int]),
    "mdd_set_history_mode": (C.c_int, [_vp, C.c_int]),
    "mdd_set_history_remap": (C.c_int, [_vp, C.c_int]),
    "mdd_set_dem_split": (C.c_int, [_vp, C.c_int]),
    "mdd_set_dem_tiles": (C.c_int, [_vp, C.c_int]),
    "mdd_set_dem_wave": (C.c_int, [_vp, C.c_int]),
    "mdd_set_lj_post_fused": (C.c_int, [_vp, C.c_int]),
    "mdd_set_lj_positions": (C.c_int, [_vp, C.c_int]),
    "mdd_set_virial": (C.c_int, [_vp, C.c_int]),
    "mdd_set_dt": (C.c_int, [_vp, C.c_double]),
    "mdd_set_materials": (C.c_int, [_vp, C.c_int, _dp, _dp, _dp, _dp]),
    "mdd_set_material_extras": (C.c_int, [_vp, C.c_int, _dp]),
    "mdd_set_contact_models": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int, C.c_int]),
    "mdd_set_lj": (C.c_int, [_vp, C.c_double, C.c_double, C.c_double]),
    "mdd_set_nose_hoover": (C.c_int, [_vp, C.c_int, C.c_double, C.c_double]),
    "mdd_set_integrators": (C.c_int, [_vp, C.c_int, C.c_int]),
    "mdd_set_gravity": (C.c_int, [_vp, C.c_int, _dp]),
    "mdd_add_wall": (C.c_int, [_vp, _dp, _dp, C.c_int, _dp, _dp]),
    "mdd_clear_walls": (C.c_int, [_vp]),
    "mdd_add_cylinder_wall": (C.c_int, [_vp, C.c_int, C.c_double, C.c_double, C.c_double,
                                        C.c_double, C.c_double, C.c_int, C.c_int]),
    "mdd_add_sphere_wall": (C.c_int, [_vp, _dp, C.c_double, C.c_int, C.c_int]),
    "mdd_set_wall_motion": (C.c_int, [_vp, C.c_int, C.c_int, _dp]),
    "mdd_add_region_wall": (C.c_int, [_vp, C.c_int, C.POINTER(C.c_int), _dp, C.c_int, C.c_int]),
    "mdd_set_wall_active": (C.c_int, [_vp, C.c_int, C.c_int]),
    "mdd_get_walls": (C.c_int, [_vp, _dp, _dp, _dp, C.POINTER(C.c_int)]),
    "mdd_wall_move": (C.c_int, [_vp]),
    "mdd_wall_zero_force_accumulators": (C.c_int, [_vp]),
    "mdd_upload_state": (C.c_int, [_vp, C.c_int64, _u32p, _u32p, _dp, _dp, _dp, _dp, _dp, _dp]),
    "mdd_nlocal": (C.c_int64, [_vp]),
    "mdd_download_state": (C.c_int, [_vp, _u32p, _dp, _dp, _dp, _dp, _i32p]),
    "mdd_download_forces": (C.c_int, [_vp, _dp, _dp]),
    "mdd_upload_forces": (C.c_int, [_vp, _dp, _dp]),
    "mdd_neighbor_setup": (C.c_int, [_vp]),
    "mdd_bin_neighbor_list": (C.c_int, [_vp]),
    "mdd_set_half_list": (C.c_int, [_vp, C.c_int64, _u32p, _u32p]),
    "mdd_npairs": (C.c_int64, [_vp]),
    "mdd_download_pairs": (C.c_int, [_vp, C.c_int64, _u32p, _u32p, _u8p, _i64p]),
    "mdd_lj_force": (C.c_int, [_vp]),
    "mdd_hertz_mindlin_contact_force": (C.c_int, [_vp]),
    "mdd_wall_contact_force": (C.c_int, [_vp]),
    "mdd_apply_gravity": (C.c_int, [_vp]),
    "mdd_initial_integration": (C.c_int, [_vp]),
    "mdd_final_integration": (C.c_int, [_vp]),
    "mdd_initial_rotation": (C.c_int, [_vp]),
    "mdd_final_rotation": (C.c_int, [_vp]),
    "mdd_setup_nose_hoover": (C.c_int, [_vp]),
    "mdd_nh_pre_initial": (C.c_int, [_vp]),
    "mdd_nh_post_final": (C.c_int, [_vp]),
    "mdd_run": (C.c_int, [_vp, C.c_int64]),
    "mdd_download_history": (C.c_int, [_vp, C.c_int64, _u32p, _u32p, _dp, _i64p]),
    "mdd_download_history7": (C.c_int, [_vp, C.c_int64, _u32p, _u32p, _dp, _i64p]),
    "mdd_compute_thermo": (C.c_int, [_vp, C.c_int, C.POINTER(Thermo)]),
    "mdd_compute_thermo_noalloc": (C.c_int, [_vp, C.POINTER(Thermo)]),
    "mdd_synchronize": (C.c_int, [_vp]),
    "mdd_set_profiling": (C.c_int, [_vp, C.c_int]),
    "mdd_force_kernel_stats": (C.c_int, [_vp, _dp, _i64p]),
    "mdd_rebuild_stats": (C.c_int, [_vp, _dp, _dp, _i64p]),
    "mdd_set_force_lanes": (C.c_int, [_vp, C.c_int]),
    "mdd_set_launch_graphs": (C.c_int, [_vp, C.c_int]),
    "mdd_download_properties": (C.c_int, [_vp, _u32p, _dp, _dp, _dp]),
    "mdd_upload_quaternions": (C.c_int, [_vp, _dp]),
    "mdd_set_build_kernel": (C.c_int, [_vp, C.c_int]),
    "mdd_launch_graph_stats": (C.c_int, [_vp, _i64p, _i64p]),
    "mdd_device_memory": (C.c_int, [C.c_int, _i64p, _i64p]),
    "mdd_add_system": (C.c_int, [_vp, _dp, _dp, C.c_double, C.c_double, C.c_double, C.c_double,
                                 C.c_int64, _u32p, _dp, _dp, _dp, _dp, _dp, _dp]),
    "mdd_nsystems": (C.c_int, [_vp]),
    "mdd_system_range": (C.c_int, [_vp, C.c_int, _i64p, _i64p]),
}


def lib():
    """Load libmddem_hip.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: run `python -m mddem_amd.build`")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _d(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a if shape is None else a.reshape(shape)


def _p(a, t=_dp):
    return None if a is None else a.ctypes.data_as(t)


class System:
    """One particle system on one GPU (an App with SingleProcessComm)."""

    def __init__(self, model: int, precision: int = FP64, device: int = 0, capacity: int = 0):
        self.L = lib()
        self.model = model
        self.precision = precision
        self.nmat = 0
        h = _vp()
        rc = self.L.mdd_create(C.byref(h), C.byref(Params(device, model, precision, capacity)))
        if rc != 0:
            raise MddError(rc, "mdd_create failed")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.L.mdd_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _c(self, rc):
        if rc != 0:
            raise MddError(rc, self.L.mdd_last_error(self.h).decode())
        return rc

    # ---- configuration
    def set_box(self, lo, hi, periodic):
        self._c(self.L.mdd_set_box(self.h, _p(_d(lo)), _p(_d(hi)),
                                   (C.c_int * 3)(*[int(bool(p)) for p in periodic])))

    def set_neighbor(self, skin_fraction=1.0, bin_size=1.0, every=0, check=True, sort_every=1000,
                     newton=True):
        self._c(self.L.mdd_set_neighbor(self.h, skin_fraction, bin_size, every, int(check),
                                        sort_every, int(newton)))

    def set_rebuild_on_box_exit(self, enabled: bool):
        self._c(self.L.mdd_set_rebuild_on_box_exit(self.h, int(enabled)))

    def set_virial(self, every: int):
        """DEM contact virial on every `every`-th step's force pass (0: off)."""
        self._c(self.L.mdd_set_virial(self.h, int(every)))

    def set_dem_split(self, mode: int):
        """DEM pass at one lane per particle: -1 auto, 0 one kernel, 1 split."""
        self._c(self.L.mdd_set_dem_split(self.h, int(mode)))

    def set_lj_positions(self, mode: int):
        """LJ mixed pair pass: 1 fixed-point positions (default), 0 FP64 positions."""
        self._c(self.L.mdd_set_lj_positions(self.h, int(mode)))

    def set_dem_tiles(self, on: bool):
        """DEM pair pass in tile order (bands of bin columns) or index order."""
        self._c(self.L.mdd_set_dem_tiles(self.h, int(bool(on))))

    def set_dem_wave(self, on: bool):
        """DEM pair pass at one lane per particle: contacts compacted per wavefront (on) or
        each lane's own contacts (off)."""
        self._c(self.L.mdd_set_dem_wave(self.h, int(bool(on))))

    def set_lj_post_fused(self, on: bool):
        """Mixed thermostatted LJ: the final half step inside the row pass (on) or in its
        own kernel (off)."""
        self._c(self.L.mdd_set_lj_post_fused(self.h, int(bool(on))))

    def set_history_remap(self, method: int):
        """0: sorted 64-bit pair keys (default), 1: search of each old row."""
        self._c(self.L.mdd_set_history_remap(self.h, int(method)))

    def set_history_mode(self, mode: int):
        """0: history kept per owning particle (the reference), 1: per pair."""
        self._c(self.L.mdd_set_history_mode(self.h, int(mode)))

    def set_dt(self, dt):
        self._c(self.L.mdd_set_dt(self.h, dt))

    def set_materials(self, youngs, poisson, restitution, friction):
        self._c(self.L.mdd_set_materials(self.h, len(youngs), _p(_d(youngs)), _p(_d(poisson)),
                                         _p(_d(restitution)), _p(_d(friction))))
        self.nmat = len(youngs)

    def set_material_extras(self, **props):
        """Optional per-material contact properties by name (MATERIAL_EXTRAS),
        one value per material; those not given are zero."""
        n = self.nmat
        vals = np.zeros((len(MATERIAL_EXTRAS), n))
        for k, v in props.items():
            vals[MATERIAL_EXTRAS.index(k)] = v
        self._c(self.L.mdd_set_material_extras(self.h, n, _p(_d(vals))))

    def set_contact_models(self, hooke=False, dmt=False, rolling_sds=False, twisting_sds=False):
        self._c(self.L.mdd_set_contact_models(self.h, int(hooke), int(dmt), int(rolling_sds),
                                              int(twisting_sds)))

    def set_lj(self, epsilon=1.0, sigma=1.0, cutoff=2.5):
        self._c(self.L.mdd_set_lj(self.h, epsilon, sigma, cutoff))

    def set_nose_hoover(self, temperature, coupling, enabled=True):
        self._c(self.L.mdd_set_nose_hoover(self.h, int(enabled), temperature, coupling))

    def set_integrators(self, velocity_verlet, rotation):
        self._c(self.L.mdd_set_integrators(self.h, int(velocity_verlet), int(rotation)))

    def set_gravity(self, g, enabled=True):
        self._c(self.L.mdd_set_gravity(self.h, int(enabled), _p(_d(g))))

    def add_wall(self, point, normal, material=0, bound_lo=None, bound_hi=None):
        self._c(self.L.mdd_add_wall(self.h, _p(_d(point)), _p(_d(normal)), material,
                                    None if bound_lo is None else _p(_d(bound_lo)),
                                    None if bound_hi is None else _p(_d(bound_hi))))

    # ---- state
    def upload(self, pos, vel, mass, cutoff_radius, radius=None, omega=None, tag=None, atype=None):
        pos = _d(pos, (-1, 3))
        n = pos.shape[0]
        tag = np.ascontiguousarray(np.arange(n) if tag is None else tag, dtype=np.uint32)
        atype = None if atype is None else np.ascontiguousarray(atype, dtype=np.uint32)
        vel = _d(vel, (-1, 3))
        mass = _d(np.broadcast_to(mass, (n,)))
        cr = _d(np.broadcast_to(cutoff_radius, (n,)))
        rad = None if radius is None else _d(np.broadcast_to(radius, (n,)))
        om = None if omega is None else _d(omega, (-1, 3))
        self._c(self.L.mdd_upload_state(self.h, n, _p(tag, _u32p), _p(atype, _u32p), _p(pos),
                                        _p(vel), _p(om), _p(rad), _p(cr), _p(mass)))

    def add_system(self, lo, hi, youngs, poisson, restitution, friction, pos, vel, mass,
                   cutoff_radius, radius, omega=None, tag=None):
        """Append one periodic DEM system to a batched ensemble."""
        pos = _d(pos, (-1, 3))
        n = pos.shape[0]
        tag = np.ascontiguousarray(np.arange(n) if tag is None else tag, dtype=np.uint32)
        vel = _d(vel, (-1, 3))
        mass = _d(np.broadcast_to(mass, (n,)))
        cr = _d(np.broadcast_to(cutoff_radius, (n,)))
        rad = _d(np.broadcast_to(radius, (n,)))
        om = None if omega is None else _d(omega, (-1, 3))
        self._c(self.L.mdd_add_system(self.h, _p(_d(lo)), _p(_d(hi)), float(youngs),
                                      float(poisson), float(restitution), float(friction), n,
                                      _p(tag, _u32p), _p(pos), _p(vel), _p(om), _p(rad), _p(cr),
                                      _p(mass)))
        self._ensemble = True

    def nsystems(self) -> int:
        rc = self.L.mdd_nsystems(self.h)
        self._c(min(rc, 0))
        return rc

    def system_range(self, k: int):
        f, c = C.c_int64(0), C.c_int64(0)
        self._c(self.L.mdd_system_range(self.h, int(k), C.byref(f), C.byref(c)))
        return f.value, c.value

    @property
    def n(self):
        if getattr(self, "_ensemble", False):
            self.nsystems()  # commits staged systems
        return self.L.mdd_nlocal(self.h)

    def state(self):
        n = self.n
        dem = self.model == MODEL_DEM
        out = dict(tag=np.zeros(n, np.uint32), pos=np.zeros((n, 3)), vel=np.zeros((n, 3)),
                   image=np.zeros((n, 3), np.int32))
        if dem:
            out["omega"] = np.zeros((n, 3))
            out["quat"] = np.zeros((n, 4))
        self._c(self.L.mdd_download_state(self.h, _p(out["tag"], _u32p), _p(out["pos"]),
                                          _p(out["vel"]), _p(out.get("omega")),
                                          _p(out.get("quat")), _p(out["image"], _i32p)))
        return out

    def upload_quaternions(self, quat):
        q = np.ascontiguousarray(quat, dtype=np.float64).reshape(self.n, 4)
        self._c(self.L.mdd_upload_quaternions(self.h, _p(q)))

    def properties(self):
        """(type, mass, radius, cutoff_radius) in the current particle order."""
        n = self.n
        typ = np.zeros(n, np.uint32)
        mass, rad, cutr = np.zeros(n), np.zeros(n), np.zeros(n)
        self._c(self.L.mdd_download_properties(self.h, _p(typ, _u32p), _p(mass), _p(rad),
                                               _p(cutr)))
        return typ, mass, rad, cutr

    def forces(self):
        n = self.n
        f = np.zeros((n, 3))
        t = np.zeros((n, 3)) if self.model == MODEL_DEM else None
        self._c(self.L.mdd_download_forces(self.h, _p(f), _p(t)))
        return f, t

    def set_forces(self, force=None, torque=None):
        self._c(self.L.mdd_upload_forces(self.h, None if force is None else _p(_d(force)),
                                         None if torque is None else _p(_d(torque))))

    # ---- systems
    def neighbor_setup(self):
        self._c(self.L.mdd_neighbor_setup(self.h))

    def bin_neighbor_list(self):
        self._c(self.L.mdd_bin_neighbor_list(self.h))

    def set_half_list(self, offsets, indices):
        off = np.ascontiguousarray(offsets, dtype=np.uint32)
        idx = np.ascontiguousarray(indices, dtype=np.uint32)
        self._c(self.L.mdd_set_half_list(self.h, len(off) - 1, _p(off, _u32p), _p(idx, _u32p)))

    def npairs(self):
        v = self.L.mdd_npairs(self.h)
        if v < 0:
            self._c(v)
        return v

    def pairs(self):
        n = C.c_int64(0)
        self._c(self.L.mdd_download_pairs(self.h, 0, None, None, None, C.byref(n)))
        m = n.value
        o = np.zeros(m, np.uint32)
        p = np.zeros(m, np.uint32)
        c = np.zeros(m, np.uint8)
        self._c(self.L.mdd_download_pairs(self.h, m, _p(o, _u32p), _p(p, _u32p), _p(c, _u8p),
                                          C.byref(n)))
        return o, p, c

    def history(self):
        n = C.c_int64(0)
        self._c(self.L.mdd_download_history(self.h, 0, None, None, None, C.byref(n)))
        m = n.value
        a = np.zeros(m, np.uint32)
        b = np.zeros(m, np.uint32)
        s = np.zeros((m, 3))
        self._c(self.L.mdd_download_history(self.h, m, _p(a, _u32p), _p(b, _u32p), _p(s),
                                            C.byref(n)))
        return a, b, s

    def history7(self):
        """History with all 7 values per contact: spring, rolling, twist."""
        n = C.c_int64(0)
        self._c(self.L.mdd_download_history7(self.h, 0, None, None, None, C.byref(n)))
        m = n.value
        a = np.zeros(m, np.uint32)
        b = np.zeros(m, np.uint32)
        s = np.zeros((m, 7))
        self._c(self.L.mdd_download_history7(self.h, m, _p(a, _u32p), _p(b, _u32p), _p(s),
                                             C.byref(n)))
        return a, b, s

    def lj_force(self):
        self._c(self.L.mdd_lj_force(self.h))

    def hertz_mindlin_contact_force(self):
        self._c(self.L.mdd_hertz_mindlin_contact_force(self.h))

    def wall_contact_force(self):
        self._c(self.L.mdd_wall_contact_force(self.h))

    def add_cylinder_wall(self, axis, center, radius, lo=-math.inf, hi=math.inf, inside=False,
                          material=0):
        self._c(self.L.mdd_add_cylinder_wall(self.h, axis, center[0], center[1], radius, lo, hi,
                                             int(inside), material))

    def add_sphere_wall(self, center, radius, inside=False, material=0):
        self._c(self.L.mdd_add_sphere_wall(self.h, _p(_d(center)), radius, int(inside),
                                           material))

    def add_region_wall(self, region, inside=False, material=0):
        """A region-surface wall; `region` as flatten_region takes it."""
        types, params = flatten_region(region)
        t = (C.c_int * len(types))(*types)
        self._c(self.L.mdd_add_region_wall(self.h, len(types), t, _p(_d(params)), int(inside),
                                           material))

    def set_wall_motion(self, wall, motion, params=(0.0, 0.0, 0.0)):
        self._c(self.L.mdd_set_wall_motion(self.h, wall, motion, _p(_d(params))))

    def set_wall_active(self, wall, active):
        self._c(self.L.mdd_set_wall_active(self.h, wall, int(active)))

    def walls(self):
        """Per wall: last force pass' summed contact force, point, velocity."""
        n = C.c_int(0)
        self._c(self.L.mdd_get_walls(self.h, None, None, None, C.byref(n)))
        acc, pt, vel = np.zeros(n.value), np.zeros((n.value, 3)), np.zeros((n.value, 3))
        self._c(self.L.mdd_get_walls(self.h, _p(acc), _p(pt), _p(vel), C.byref(n)))
        return dict(force=acc, point=pt, velocity=vel)

    def wall_move(self):
        self._c(self.L.mdd_wall_move(self.h))

    def wall_zero_force_accumulators(self):
        self._c(self.L.mdd_wall_zero_force_accumulators(self.h))

    def apply_gravity(self):
        self._c(self.L.mdd_apply_gravity(self.h))

    def initial_integration(self):
        self._c(self.L.mdd_initial_integration(self.h))

    def final_integration(self):
        self._c(self.L.mdd_final_integration(self.h))

    def initial_rotation(self):
        self._c(self.L.mdd_initial_rotation(self.h))

    def final_rotation(self):
        self._c(self.L.mdd_final_rotation(self.h))

    def setup_nose_hoover(self):
        self._c(self.L.mdd_setup_nose_hoover(self.h))

    def nh_pre_initial(self):
        self._c(self.L.mdd_nh_pre_initial(self.h))

    def nh_post_final(self):
        self._c(self.L.mdd_nh_post_final(self.h))

    def run(self, nsteps: int):
        self._c(self.L.mdd_run(self.h, int(nsteps)))

    def synchronize(self):
        self._c(self.L.mdd_synchronize(self.h))

    def thermo(self, energy: bool = False) -> dict:
        t = Thermo()
        self._c(self.L.mdd_compute_thermo(self.h, int(energy), C.byref(t)))
        return t.as_dict()

    def thermo_noalloc(self) -> dict:
        """mdd_compute_thermo_noalloc: the thermo values without any
        allocation (static device scratch); no energy, virial or pressure."""
        t = Thermo()
        self._c(self.L.mdd_compute_thermo_noalloc(self.h, C.byref(t)))
        return t.as_dict()

    def set_profiling(self, every: int):
        self._c(self.L.mdd_set_profiling(self.h, int(every)))

    def force_kernel_stats(self):
        ms = C.c_double(0)
        n = C.c_int64(0)
        self._c(self.L.mdd_force_kernel_stats(self.h, C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def rebuild_stats(self):
        """(average device ms of a timed rebuild, average idle ms before one,
        rebuilds timed)."""
        ms, gap = C.c_double(0), C.c_double(0)
        n = C.c_int64(0)
        self._c(self.L.mdd_rebuild_stats(self.h, C.byref(ms), C.byref(gap), C.byref(n)))
        return ms.value, gap.value, n.value

    def set_force_lanes(self, log2_lanes: int):
        """Row kernels at 2^log2_lanes lanes per particle (0..3), or the LJ
        cell-tile kernel (-1)."""
        self._c(self.L.mdd_set_force_lanes(self.h, int(log2_lanes)))

    def set_build_kernel(self, mode: int):
        """0: by grid occupancy, 1: one wavefront per cell, 2: one lane per member."""
        self._c(self.L.mdd_set_build_kernel(self.h, int(mode)))

    def set_launch_graphs(self, enabled: bool):
        self._c(self.L.mdd_set_launch_graphs(self.h, int(bool(enabled))))

    def launch_graph_stats(self):
        """(rebuilds replayed from a captured graph, graphs captured)."""
        r, k = C.c_int64(0), C.c_int64(0)
        self._c(self.L.mdd_launch_graph_stats(self.h, C.byref(r), C.byref(k)))
        return r.value, k.value


def device_memory(device: int = 0):
    """(free, total) device memory in bytes."""
    f, t = C.c_int64(0), C.c_int64(0)
    rc = lib().mdd_device_memory(int(device), C.byref(f), C.byref(t))
    if rc != 0:
        raise MddError(rc, f"mdd_device_memory failed ({rc})")
    return f.value, t.value


def from_workload(w, precision=FP64, box_exit=True, sort_every=1000, device=0,
                  history_mode=0) -> System:
    """Configure a System the way the reference plugin groups would
    (GranularDefaultPlugins / LJDefaultPlugins) for a workloads.Workload."""
    s = System(MODEL_DEM if w.model == "dem" else MODEL_LJ, precision, device)
    s.set_box(w.lo, w.hi, w.periodic)
    s.set_neighbor(w.skin_fraction, w.bin_size, getattr(w, "every", 0), True, sort_every, True)
    s.set_rebuild_on_box_exit(box_exit)
    s.set_dt(w.dt)
    if w.model == "dem":
        m = w.materials
        s.set_materials(m["youngs"], m["poisson"], m["restitution"], m["friction"])
        s.set_history_mode(history_mode)
        if w.gravity is not None:
            s.set_gravity(w.gravity)
        for p, nrm, mat in w.walls:
            s.add_wall(p, nrm, mat)
    else:
        s.set_lj(*w.lj)
        if w.nose_hoover is not None:
            s.set_nose_hoover(*w.nose_hoover)
            s.set_integrators(False, False)
        else:
            s.set_integrators(True, False)
    s.upload(w.pos, w.vel, w.mass, w.cutoff_radius, radius=w.radius, omega=w.omega, tag=w.tag,
             atype=w.atype)
    return s
