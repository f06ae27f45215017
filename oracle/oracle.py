"""ctypes wrapper around the CPU oracle (oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the package smoke check and by
the CPU-baseline leg of bench.py. The product path (mddem_amd) never imports
it. Every method maps one-to-one onto a reference system; see oracle.h for the
file:line each one restates.
"""
from __future__ import annotations

import ctypes as C
import math
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_DEFAULT_LIB = os.path.join(_HERE, "liboracle.so")
# MDDEM_ORACLE_LIB names another build of oracle.c (the tests build variants
# with a planted mistake to show that their closed forms reject it)
_LIB_PATH = os.environ.get("MDDEM_ORACLE_LIB", _DEFAULT_LIB)

MODEL_LJ, MODEL_DEM = 0, 1
HIST_REFERENCE, HIST_CANONICAL = 0, 1
STATE_FULL_REBUILD, STATE_COMMUNICATE_ONLY = 0, 1

_dp = C.POINTER(C.c_double)
_u32p = C.POINTER(C.c_uint32)
_i32p = C.POINTER(C.c_int32)
_u8p = C.POINTER(C.c_uint8)
_ip = C.POINTER(C.c_int)


def build() -> str:
    """Compile liboracle.so with its Makefile (gcc, no GPU needed)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _DEFAULT_LIB


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        vp = C.c_void_p
        sig = {
            "orc_create": (vp, [C.c_int]),
            "orc_destroy": (None, [vp]),
            "orc_error": (C.c_char_p, [vp]),
            "orc_set_box": (None, [vp, _dp, _dp, _ip]),
            "orc_set_neighbor": (None, [vp, C.c_double, C.c_double, C.c_int, C.c_int, C.c_int, C.c_int]),
            "orc_set_dt": (None, [vp, C.c_double]),
            "orc_set_materials": (C.c_int, [vp, C.c_int, _dp, _dp, _dp, _dp]),
            "orc_get_material_tables": (C.c_int, [vp, _dp, _dp, _dp, _dp]),
            "orc_set_material_extras": (C.c_int, [vp, C.c_int, _dp]),
            "orc_get_extra_table": (C.c_int, [vp, C.c_int, _dp]),
            "orc_set_contact_models": (None, [vp, C.c_int, C.c_int, C.c_int, C.c_int]),
            "orc_set_lj": (None, [vp, C.c_double, C.c_double, C.c_double]),
            "orc_get_lj": (None, [vp, _dp]),
            "orc_set_nose_hoover": (None, [vp, C.c_int, C.c_double, C.c_double]),
            "orc_set_integrators": (None, [vp, C.c_int, C.c_int]),
            "orc_set_gravity": (None, [vp, C.c_int, _dp]),
            "orc_add_wall": (C.c_int, [vp, _dp, _dp, C.c_int, _dp, _dp]),
            "orc_add_cylinder_wall": (C.c_int, [vp, C.c_int, C.c_double, C.c_double, C.c_double,
                                                C.c_double, C.c_double, C.c_int, C.c_int]),
            "orc_add_sphere_wall": (C.c_int, [vp, _dp, C.c_double, C.c_int, C.c_int]),
            "orc_set_wall_motion": (C.c_int, [vp, C.c_int, C.c_int, _dp]),
            "orc_add_region_wall": (C.c_int, [vp, C.c_int, _ip, _dp, C.c_int, C.c_int]),
            "orc_region_surface": (C.c_double, [vp, C.c_int, _dp, _dp, _dp]),
            "orc_set_wall_active": (C.c_int, [vp, C.c_int, C.c_int]),
            "orc_get_walls": (C.c_int, [vp, _dp, _dp, _dp]),
            "orc_wall_move": (None, [vp]),
            "orc_wall_zero_force_accumulators": (None, [vp]),
            "orc_set_history_mode": (None, [vp, C.c_int]),
            "orc_set_atoms": (C.c_int, [vp, C.c_int, _u32p, _u32p, _dp, _dp, _dp, _dp, _dp, _dp]),
            "orc_nlocal": (C.c_int, [vp]),
            "orc_nghost": (C.c_int, [vp]),
            "orc_get_atoms": (None, [vp, _u32p, _dp, _dp, _dp, _dp, _dp, _dp, _i32p]),
            "orc_set_forces": (None, [vp, _dp, _dp]),
            "orc_neighbor_setup": (None, [vp]),
            "orc_bin_grid": (C.c_int, [vp, _dp]),
            "orc_decide_rebuild": (None, [vp]),
            "orc_zero_all_forces": (None, [vp]),
            "orc_remove_ghost_atoms": (None, [vp]),
            "orc_pbc": (None, [vp]),
            "orc_forward_comm": (None, [vp]),
            "orc_sort_atoms_by_bin": (None, [vp]),
            "orc_borders": (None, [vp]),
            "orc_bin_neighbor_list": (None, [vp]),
            "orc_hertz_mindlin_contact_force": (C.c_int, [vp]),
            "orc_wall_contact_force": (None, [vp]),
            "orc_apply_gravity": (None, [vp]),
            "orc_lj_force": (None, [vp]),
            "orc_lj_virial": (None, [vp, _dp]),
            "orc_set_virial": (None, [vp, C.c_int]),
            "orc_set_sort_only_on_rebuild": (None, [vp, C.c_int]),
            "orc_granular_temperature": (None, [vp, _dp]),
            "orc_contact_virial": (None, [vp, _dp]),
            "orc_lj_tails": (None, [vp, _dp, _dp]),
            "orc_pressure": (C.c_double, [vp, _dp, C.c_double]),
            "orc_reverse_send_force": (None, [vp]),
            "orc_initial_integration": (None, [vp]),
            "orc_final_integration": (None, [vp]),
            "orc_initial_rotation": (None, [vp]),
            "orc_final_rotation": (None, [vp]),
            "orc_setup_nose_hoover": (None, [vp]),
            "orc_nh_pre_initial": (None, [vp]),
            "orc_nh_post_final": (None, [vp]),
            "orc_compute_ke": (C.c_double, [vp]),
            "orc_step": (C.c_int, [vp]),
            "orc_run": (C.c_int, [vp, C.c_int64]),
            "orc_state": (C.c_int, [vp]),
            "orc_p_xi": (C.c_double, [vp]),
            "orc_rebuild_count": (C.c_int64, [vp]),
            "orc_overlap_warnings": (C.c_int64, [vp]),
            "orc_set_half_list": (C.c_int, [vp, C.c_int, _u32p, _u32p]),
            "orc_npairs": (C.c_int64, [vp]),
            "orc_get_pairs": (C.c_int64, [vp, _u32p, _u32p, _u8p]),
            "orc_history_count": (C.c_int64, [vp]),
            "orc_get_history": (C.c_int64, [vp, _u32p, _u32p, _dp]),
            "orc_get_history7": (C.c_int64, [vp, _u32p, _u32p, _dp]),
            "orc_set_history": (C.c_int, [vp, C.c_int64, _u32p, _u32p, _dp]),
            "orc_dem_mass": (C.c_double, [C.c_double, C.c_double]),
            "orc_dem_inv_inertia": (C.c_double, [C.c_double, C.c_double]),
            "orc_rayleigh_dt": (C.c_double, [C.c_double, C.c_double, C.c_double, C.c_double]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _d(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _ptr(a, t=_dp):
    return None if a is None else a.ctypes.data_as(t)


# optional per-material properties, in the order the C API takes them
MATERIAL_EXTRAS = ("rolling_friction", "twisting_friction", "cohesion_energy", "surface_energy",
                   "kn", "kt", "rolling_stiffness", "rolling_damping", "twisting_stiffness",
                   "twisting_damping")


class Oracle:
    """One reference-path system on the CPU (a SingleProcessComm App)."""

    def __init__(self, model: int):
        self.L = lib()
        self.h = self.L.orc_create(model)
        self.model = model

    def __del__(self):
        h = getattr(self, "h", None)
        if h:
            self.L.orc_destroy(h)
            self.h = None

    def _check(self, rc):
        if rc != 0:
            raise RuntimeError(self.L.orc_error(self.h).decode())
        return rc

    # configuration
    def set_box(self, lo, hi, periodic):
        self.L.orc_set_box(self.h, _ptr(_d(lo)), _ptr(_d(hi)),
                           (C.c_int * 3)(*[int(p) for p in periodic]))

    def set_neighbor(self, skin_fraction=1.0, bin_size=1.0, every=0, check=True,
                     sort_every=1000, newton=True):
        self.L.orc_set_neighbor(self.h, skin_fraction, bin_size, every, int(check),
                                sort_every, int(newton))

    def set_dt(self, dt):
        self.L.orc_set_dt(self.h, dt)

    def set_materials(self, youngs, poisson, restitution, friction):
        n = len(youngs)
        self._check(self.L.orc_set_materials(self.h, n, _ptr(_d(youngs)), _ptr(_d(poisson)),
                                             _ptr(_d(restitution)), _ptr(_d(friction))))

    def material_tables(self):
        n = self.L.orc_get_material_tables(self.h, None, None, None, None)
        out = [np.zeros(n * n) for _ in range(4)]
        self.L.orc_get_material_tables(self.h, *[_ptr(o) for o in out])
        return {k: v.reshape(n, n) for k, v in zip(("beta", "friction", "e_eff", "g_eff"), out)}

    def set_material_extras(self, **props):
        """Optional per-material properties by name (MATERIAL_EXTRAS), one
        value per material; properties not given are zero."""
        n = self.L.orc_get_material_tables(self.h, None, None, None, None)
        vals = np.zeros((len(MATERIAL_EXTRAS), n))
        for k, v in props.items():
            vals[MATERIAL_EXTRAS.index(k)] = v
        self._check(self.L.orc_set_material_extras(self.h, n, _ptr(_d(vals))))

    def extra_table(self, name):
        n = self.L.orc_get_material_tables(self.h, None, None, None, None)
        out = np.zeros(n * n)
        self.L.orc_get_extra_table(self.h, MATERIAL_EXTRAS.index(name), _ptr(out))
        return out.reshape(n, n)

    def set_contact_models(self, hooke=False, dmt=False, rolling_sds=False, twisting_sds=False):
        self.L.orc_set_contact_models(self.h, int(hooke), int(dmt), int(rolling_sds),
                                      int(twisting_sds))

    def set_lj(self, epsilon=1.0, sigma=1.0, cutoff=2.5):
        self.L.orc_set_lj(self.h, epsilon, sigma, cutoff)

    def granular_temperature(self):
        """print_granular_temperature's values: T_g, KE, |p| and p."""
        w = np.zeros(6)
        self.L.orc_granular_temperature(self.h, _ptr(w))
        return dict(granular_temperature=w[0], ke=w[1], momentum_mag=w[2],
                    momentum=w[3:6].copy())

    def set_virial(self, on=True):
        self.L.orc_set_virial(self.h, int(bool(on)))

    def contact_virial(self):
        """xx yy zz xy xz yz of the last contact force pass."""
        w = np.zeros(6)
        self.L.orc_contact_virial(self.h, _ptr(w))
        return w

    def lj_virial(self):
        """W (xx, yy, zz, xy, xz, yz) over the current half list."""
        w = np.zeros(6)
        self.L.orc_lj_virial(self.h, _ptr(w))
        return w

    def lj_tails(self):
        e, p = C.c_double(0), C.c_double(0)
        self.L.orc_lj_tails(self.h, C.byref(e), C.byref(p))
        return e.value, p.value

    def pressure(self, w, p_tail):
        return self.L.orc_pressure(self.h, _ptr(_d(w)), float(p_tail))

    def lj_coeffs(self):
        out = np.zeros(3)
        self.L.orc_get_lj(self.h, _ptr(out))
        return out

    def set_nose_hoover(self, temperature, coupling, enabled=True):
        self.L.orc_set_nose_hoover(self.h, int(enabled), temperature, coupling)

    def set_integrators(self, velocity_verlet, rotation):
        self.L.orc_set_integrators(self.h, int(velocity_verlet), int(rotation))

    def set_gravity(self, g, enabled=True):
        self.L.orc_set_gravity(self.h, int(enabled), _ptr(_d(g)))

    def _index(self, k):
        if k < 0:
            raise RuntimeError(self.L.orc_error(self.h).decode() or "wall rejected")
        return k

    def add_wall(self, point, normal, material=0, bound_lo=(-1e300,) * 3, bound_hi=(1e300,) * 3):
        return self._index(self.L.orc_add_wall(self.h, _ptr(_d(point)), _ptr(_d(normal)),
                                               material, _ptr(_d(bound_lo)), _ptr(_d(bound_hi))))

    def add_cylinder_wall(self, axis, center, radius, lo=-math.inf, hi=math.inf, inside=False,
                          material=0):
        return self._index(self.L.orc_add_cylinder_wall(self.h, axis, center[0], center[1],
                                                        radius, lo, hi, int(inside), material))

    def add_sphere_wall(self, center, radius, inside=False, material=0):
        return self._index(self.L.orc_add_sphere_wall(self.h, _ptr(_d(center)), radius,
                                                      int(inside), material))

    def add_region_wall(self, types, params, inside=False, material=0):
        """Leaf primitives (type codes and 12 parameters each, oracle.h)."""
        t = (C.c_int * len(types))(*[int(x) for x in types])
        q = _d(np.asarray(params, dtype=np.float64).reshape(len(types), 12))
        return self._index(self.L.orc_add_region_wall(self.h, len(types), t, _ptr(q),
                                                      int(inside), material))

    def region_surface(self, wall, pos):
        """(signed distance, closest point, outward normal) of a region wall."""
        pt, nrm = np.zeros(3), np.zeros(3)
        d = self.L.orc_region_surface(self.h, wall, _ptr(_d(pos)), _ptr(pt), _ptr(nrm))
        return d, pt, nrm

    def set_wall_motion(self, wall, motion, params=(0.0, 0.0, 0.0)):
        self._check(self.L.orc_set_wall_motion(self.h, wall, motion, _ptr(_d(params))))

    def set_wall_active(self, wall, active):
        self._check(self.L.orc_set_wall_active(self.h, wall, int(active)))

    def walls(self):
        n = self.L.orc_get_walls(self.h, None, None, None)
        acc, pt, vel = np.zeros(n), np.zeros((n, 3)), np.zeros((n, 3))
        self.L.orc_get_walls(self.h, _ptr(acc), _ptr(pt), _ptr(vel))
        return dict(force=acc, point=pt, velocity=vel)

    def set_sort_only_on_rebuild(self, on=True):
        self.L.orc_set_sort_only_on_rebuild(self.h, int(bool(on)))

    def set_history_mode(self, mode):
        self.L.orc_set_history_mode(self.h, mode)

    # state
    def set_atoms(self, pos, vel, mass, cutoff_radius, radius=None, omega=None, tag=None,
                  atype=None):
        pos = _d(pos).reshape(-1, 3)
        n = pos.shape[0]
        tag = np.ascontiguousarray(np.arange(n) if tag is None else tag, dtype=np.uint32)
        atype = None if atype is None else np.ascontiguousarray(atype, dtype=np.uint32)
        vel = _d(vel).reshape(-1, 3)
        mass = _d(np.broadcast_to(mass, (n,)))
        cr = _d(np.broadcast_to(cutoff_radius, (n,)))
        rad = None if radius is None else _d(np.broadcast_to(radius, (n,)))
        om = None if omega is None else _d(omega).reshape(-1, 3)
        self._keep = (pos, vel, mass, cr, rad, om, tag, atype)
        self._check(self.L.orc_set_atoms(self.h, n, _ptr(tag, _u32p), _ptr(atype, _u32p),
                                         _ptr(pos), _ptr(vel), _ptr(om), _ptr(rad), _ptr(cr),
                                         _ptr(mass)))

    @property
    def nlocal(self):
        return self.L.orc_nlocal(self.h)

    @property
    def nghost(self):
        return self.L.orc_nghost(self.h)

    def atoms(self):
        n = self.nlocal
        out = dict(tag=np.zeros(n, np.uint32), pos=np.zeros((n, 3)), vel=np.zeros((n, 3)),
                   force=np.zeros((n, 3)), omega=np.zeros((n, 3)), torque=np.zeros((n, 3)),
                   quat=np.zeros((n, 4)), image=np.zeros((n, 3), np.int32))
        self.L.orc_get_atoms(self.h, _ptr(out["tag"], _u32p), _ptr(out["pos"]), _ptr(out["vel"]),
                             _ptr(out["force"]), _ptr(out["omega"]), _ptr(out["torque"]),
                             _ptr(out["quat"]), _ptr(out["image"], _i32p))
        return out

    def set_forces(self, force=None, torque=None):
        f = None if force is None else _d(force)
        t = None if torque is None else _d(torque)
        self.L.orc_set_forces(self.h, _ptr(f), _ptr(t))

    # systems
    def neighbor_setup(self):
        self.L.orc_neighbor_setup(self.h)

    def bin_grid(self):
        out = np.zeros(13)
        self.L.orc_bin_grid(self.h, _ptr(out))
        return dict(max_cutoff=out[0], ghost_cutoff=out[1], bin_size=out[2:5],
                    bin_origin=out[5:8], bin_count=out[8:11].astype(int),
                    n_forward=int(out[11]), has_self=bool(out[12]))

    def __getattr__(self, name):
        # the remaining systems take only the handle: orc_<name>(h)
        if name.startswith("_"):
            raise AttributeError(name)
        fn = getattr(self.L, "orc_" + name, None)
        if fn is None:
            raise AttributeError(name)
        return lambda: fn(self.h)

    def hertz_mindlin_contact_force(self):
        return self._check(self.L.orc_hertz_mindlin_contact_force(self.h))

    def step(self):
        return self._check(self.L.orc_step(self.h))

    def run(self, nsteps):
        return self._check(self.L.orc_run(self.h, nsteps))

    def set_half_list(self, offsets, indices):
        off = np.ascontiguousarray(offsets, dtype=np.uint32)
        idx = np.ascontiguousarray(indices, dtype=np.uint32)
        self._keep_list = (off, idx)
        self._check(self.L.orc_set_half_list(self.h, len(off) - 1, _ptr(off, _u32p),
                                             _ptr(idx, _u32p)))

    def pairs(self):
        n = self.L.orc_npairs(self.h)
        o = np.zeros(n, np.uint32)
        p = np.zeros(n, np.uint32)
        c = np.zeros(n, np.uint8)
        self.L.orc_get_pairs(self.h, _ptr(o, _u32p), _ptr(p, _u32p), _ptr(c, _u8p))
        return o, p, c

    def history(self):
        n = self.L.orc_history_count(self.h)
        o = np.zeros(n, np.uint32)
        p = np.zeros(n, np.uint32)
        s = np.zeros((n, 3))
        self.L.orc_get_history(self.h, _ptr(o, _u32p), _ptr(p, _u32p), _ptr(s))
        return o, p, s

    def history7(self):
        """History entries with all 7 values: spring (3), rolling (3), twist."""
        n = self.L.orc_history_count(self.h)
        o = np.zeros(n, np.uint32)
        p = np.zeros(n, np.uint32)
        s = np.zeros((n, 7))
        self.L.orc_get_history7(self.h, _ptr(o, _u32p), _ptr(p, _u32p), _ptr(s))
        return o, p, s

    def set_history(self, owner, partner, spring):
        o = np.ascontiguousarray(owner, dtype=np.uint32)
        p = np.ascontiguousarray(partner, dtype=np.uint32)
        s = _d(spring).reshape(-1, 3)
        self._check(self.L.orc_set_history(self.h, len(o), _ptr(o, _u32p), _ptr(p, _u32p),
                                           _ptr(s)))

    @property
    def p_xi(self):
        return self.L.orc_p_xi(self.h)

    @property
    def rebuilds(self):
        return self.L.orc_rebuild_count(self.h)

    @property
    def state(self):
        return self.L.orc_state(self.h)


def dem_mass(density, radius):
    return lib().orc_dem_mass(density, radius)


def dem_inv_inertia(mass, radius):
    return lib().orc_dem_inv_inertia(mass, radius)


def rayleigh_time(youngs, poisson, density, radius):
    return lib().orc_rayleigh_dt(youngs, poisson, density, radius)
