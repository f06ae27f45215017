"""The HIP Hertz-Mindlin and plane-wall kernels against the same closed forms
that pin the checker (worked out from contact.rs:252-408 and
dem_wall/src/lib.rs:925-968 with the literal SQRT_5_3 = 0.9128709291752768),
through the C ABI. FP64 holds 1e-12 per force part; mixed (FP32 force
algebra and sums) 1e-5."""
import sys

import numpy as np
import pytest

from mddem_amd import hip
from tests.test_oracle_closed_forms import (GLASS, SQRT_5_3, STEEL, check_pair, first_contact,
                                            mass, mix)

pytestmark = pytest.mark.gpu
TOLS = {hip.FP64: 1e-12, hip.MIXED: 1e-5}


@pytest.fixture(autouse=True)
def _announce(request):
    print(f"\n[gpu case] {request.node.name}", flush=True)
    sys.stdout.flush()
    yield


def gpu_pair(prec, x0, x1, r0, r1, v0, v1, w0, w1, mats=(GLASS,), types=(0, 0), dt=1e-7,
             owner=0):
    s = hip.System(hip.MODEL_DEM, prec)
    s.set_box((-1, -1, -1), (1, 1, 1), (False, False, False))
    s.set_materials([m["E"] for m in mats], [m["nu"] for m in mats], [m["e"] for m in mats],
                    [m["mu"] for m in mats])
    s.set_dt(dt)
    s.upload([x0, x1], [v0, v1], [mass(r0), mass(r1)], [r0, r1], radius=[r0, r1],
             omega=[w0, w1], atype=list(types))
    s.set_half_list([0, 1, 1] if owner == 0 else [0, 0, 1], [1] if owner == 0 else [0])
    return s


def as_atoms(s):
    f, t = s.forces()
    order = np.argsort(s.state()["tag"])
    return dict(force=f[order], torque=t[order])


@pytest.mark.parametrize("prec", [hip.FP64, hip.MIXED])
@pytest.mark.parametrize("owner", [0, 1])
def test_sliding_pair_passes(prec, owner):
    # tangential.rs:353-408's sliding pair, three passes with the spring kept
    args = ([0, 0, 0], [0.0019, 0, 0], 1e-3, 1e-3, [0, 0, 0], [0, 0.1, 0.05], [0] * 3, [0] * 3)
    s = gpu_pair(prec, *args, owner=owner)
    for k in range(1, 4):
        s.hertz_mindlin_contact_force()
        want = first_contact(*args, steps=k)
        check_pair(as_atoms(s), want, TOLS[prec])
        _, _, spring = s.history()
        assert np.linalg.norm(spring[0] - want["spring"]) <= 1e-12 * np.linalg.norm(want["spring"])


@pytest.mark.parametrize("prec", [hip.FP64, hip.MIXED])
@pytest.mark.parametrize("vn", [-1.0, 0.05, 1.0])
def test_damped_normal_force(prec, vn):
    mat = dict(GLASS, e=0.5)
    args = ([0, 0, 0], [0.002 - 1e-6, 0, 0], 1e-3, 1e-3, [0, 0, 0], [vn, 0, 0], [0] * 3, [0] * 3)
    s = gpu_pair(prec, *args, mats=(mat,))
    s.hertz_mindlin_contact_force()
    f = as_atoms(s)["force"]
    if vn == 1.0:  # opening fast: clamped at zero
        assert np.all(f == 0.0)
        return
    want = first_contact(*args, mat0=mat, mat1=mat)
    assert abs(-f[0, 0] - want["fn"]) <= TOLS[prec] * want["fn"]
    assert f[0, 1] == 0.0 and f[0, 2] == 0.0


@pytest.mark.parametrize("prec", [hip.FP64, hip.MIXED])
@pytest.mark.parametrize("owner", [0, 1])
def test_unequal_radii_mixed_materials_with_spin(prec, owner):
    n = np.array([0.6, 0.8, 0.0])
    args = ([0, 0, 0], (1.999e-3 * n).tolist(), 0.8e-3, 1.2e-3, [0.05, 0.02, -0.01],
            [-0.8, -0.5, 0.2], [0, 0, 30.0], [10.0, -20.0, 5.0])
    want = first_contact(*args, mat0=GLASS, mat1=STEEL)
    s = gpu_pair(prec, *args, mats=(GLASS, STEEL), types=(0, 1), owner=owner)
    s.hertz_mindlin_contact_force()
    check_pair(as_atoms(s), want, TOLS[prec])


@pytest.mark.parametrize("prec", [hip.FP64, hip.MIXED])
@pytest.mark.parametrize("vz", [-0.5, 0.05])
def test_plane_wall_damping(prec, vz):
    mat = dict(GLASS, e=0.6)
    r, z = 1e-3, 0.999e-3
    beta, _, _, _ = mix(mat, mat)
    es = 1.0 / (2 * (1 - mat["nu"] ** 2) / mat["E"])
    d = r - z
    sdr = (d * r) ** 0.5
    sn, kn = 2 * es * sdr, 4.0 / 3.0 * es * sdr
    want = max(kn * d - 2 * beta * SQRT_5_3 * (sn * mass(r)) ** 0.5 * vz, 0.0)
    s = hip.System(hip.MODEL_DEM, prec)
    s.set_box((-1, -1, -1), (1, 1, 1), (False, False, False))
    s.set_materials([mat["E"]], [mat["nu"]], [mat["e"]], [mat["mu"]])
    s.set_dt(1e-7)
    s.upload([[0, 0, z]], [[0.3, 0, vz]], [mass(r)], [r], radius=[r], omega=[[0, 0, 0]])
    s.add_wall((0, 0, 0), (0, 0, 1))
    s.wall_zero_force_accumulators()
    s.wall_contact_force()
    f, _ = s.forces()
    assert f[0, 0] == 0.0 and f[0, 1] == 0.0
    assert abs(f[0, 2] - want) <= TOLS[prec] * want
    # the wall's force sum is fixed point at 2^-32 N
    assert abs(s.walls()["force"][0] - want) <= 2.0 ** -31 + 1e-12 * want
