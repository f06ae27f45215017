"""Step observables computed on the device against the checker: the LJ pair
virial W = sum d (x) F (md_lj/src/lib.rs:510-585) and the virial pressure
P = rho T - tr(W) / (3 V) + P_tail (md_measure/src/lib.rs:605-651,
tails md_lj/src/lib.rs:397-443), for the row and cell-tile layouts. The full
rows give each pair's d (x) F half to each side; only summation order
differs from the checker's half list."""
import sys

import numpy as np
import pytest

from mddem_amd import hip
from tests.helpers import oracle_from_workload
from tests.test_gpu_parity import _oracle_rebuild, _perturbed_lj

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _announce(request):
    print(f"\n[gpu case] {request.node.name}", flush=True)
    sys.stdout.flush()
    yield


@pytest.mark.parametrize("layout", [2, -1])
@pytest.mark.parametrize("prec", [hip.FP64, hip.MIXED])
def test_virial_and_pressure_match_oracle(layout, prec):
    w = _perturbed_lj()
    o = oracle_from_workload(w)
    _oracle_rebuild(o)
    w_ref = o.lj_virial()
    _, p_tail = o.lj_tails()
    p_ref = o.pressure(w_ref, p_tail)
    s = hip.from_workload(w, prec)
    s.set_force_lanes(layout)
    s.bin_neighbor_list()
    th = s.thermo(energy=True)
    got = np.asarray(th["virial"])
    scale = np.abs(w_ref[:3]).max()
    tol = 1e-12 if prec == hip.FP64 else 1e-5
    assert np.abs(got - w_ref).max() < tol * scale
    assert abs(th["pressure"] - p_ref) < tol * max(1.0, abs(p_ref)) * 10
    # FCC at rho 0.85: nearest neighbours at 1.185 sigma, past the force
    # zero 2^(1/6): attraction dominates and d.F > 0 (tr W > 0)
    assert got[0] + got[1] + got[2] > 0


def test_thermo_three_particles_hand_computed():
    # print_granular_temperature / compute_ke on the device (granular_temp.rs:
    # 60-99, atom.rs:489-502): T_g = 11/54, KE = 3, p = (1, 2, -3)
    r = [1e-3, 1.2e-3, 1.4e-3]
    s = hip.System(hip.MODEL_DEM, hip.FP64)
    s.set_box((-1, -1, -1), (1, 1, 1), (False, False, False))
    s.set_materials([8.7e9], [0.3], [0.95], [0.4])
    s.set_dt(1e-7)
    s.upload([[0, 0, 0], [0.1, 0, 0], [0.2, 0, 0]], [[1, 0, 0], [0, 1, 0], [0, 0, -1]],
             [1.0, 2.0, 3.0], r, radius=r)
    th = s.thermo()
    assert th["granular_temperature"] == pytest.approx(11.0 / 54.0, rel=1e-15)
    assert th["ke"] == pytest.approx(3.0, rel=1e-15)
    assert th["momentum"] == [1.0, 2.0, -3.0]
    assert th["temperature"] == pytest.approx(2.0 * 3.0 / 6.0, rel=1e-15)


@pytest.mark.parametrize("prec", [hip.FP64, hip.MIXED])
def test_thermo_matches_oracle_after_a_run(prec):
    # a drifting granular gas after 300 steps: KE, T, T_g and p of the device
    # reductions against print_granular_temperature / compute_ke on the same
    # state (the velocities themselves come from the device run)
    from mddem_amd import workloads as W
    w = W.dem_gas(n=2000, box=0.04)
    w.vel = w.vel + np.array([0.2, -0.1, 0.05])
    s = hip.from_workload(w, prec)
    s.run(300)
    st = s.state()
    o = oracle_from_workload(w)
    o.set_atoms(st["pos"], st["vel"], w.mass[st["tag"]],
                w.cutoff_radius[st["tag"]], radius=w.radius[st["tag"]], tag=st["tag"])
    g = o.granular_temperature()
    th = s.thermo()
    assert th["granular_temperature"] == pytest.approx(g["granular_temperature"], rel=1e-12)
    assert th["ke"] == pytest.approx(g["ke"], rel=1e-12)
    assert th["ke"] == pytest.approx(o.compute_ke(), rel=1e-12)
    assert np.allclose(th["momentum"], g["momentum"], rtol=1e-12, atol=1e-15)
    assert np.linalg.norm(th["momentum"]) == pytest.approx(g["momentum_mag"], rel=1e-12)
    assert th["temperature"] == pytest.approx(2.0 * g["ke"] / (3.0 * w.n - 3.0), rel=1e-12)


@pytest.mark.parametrize("prec", [hip.FP64, hip.MIXED])
def test_dem_contact_virial_of_one_pair(prec):
    # VirialStress::add_pair(d, -F_n + F_t) for the unequal-radius pair of the
    # closed-form cases (contact.rs:559-569): d (x) F_0, either owner
    from tests.test_gpu_closed_forms import gpu_pair
    from tests.test_oracle_closed_forms import GLASS, STEEL, first_contact
    n = np.array([0.6, 0.8, 0.0])
    x1 = 1.999e-3 * n
    args = ([0, 0, 0], x1.tolist(), 0.8e-3, 1.2e-3, [0.05, 0.02, -0.01], [-0.8, -0.5, 0.2],
            [0, 0, 30.0], [10.0, -20.0, 5.0])
    want = first_contact(*args, mat0=GLASS, mat1=STEEL)
    wm = np.outer(x1, want["f0"])
    want6 = np.array([wm[0, 0], wm[1, 1], wm[2, 2], wm[0, 1], wm[0, 2], wm[1, 2]])
    tol = 1e-11 if prec == hip.FP64 else 1e-5
    for owner in (0, 1):
        s = gpu_pair(prec, *args, mats=(GLASS, STEEL), types=(0, 1), owner=owner)
        s.set_virial(1)
        s.hertz_mindlin_contact_force()
        got = np.asarray(s.thermo()["virial"])
        assert np.linalg.norm(got - want6) <= tol * np.linalg.norm(want6)
