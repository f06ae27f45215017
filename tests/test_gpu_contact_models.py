"""The contact-model family on the GPU against the checker: SJKR cohesion,
JKR and DMT adhesion, Hooke springs, constant and SDS rolling resistance and
twisting friction (dem_granular/src/contact.rs:103-615, 622-990), on a
compressed, spinning granular gas where most spheres touch. Force passes
with history carried between them, full scheduled runs, and the 7 history
values per contact. Pair arithmetic is FP64 in both precisions (the optional
models keep only the per-particle sums in FP32 when mixed)."""
import sys

import numpy as np
import pytest

from mddem_amd import hip
from mddem_amd import workloads as W
from tests.helpers import canon_history, oracle_from_workload, rel_err

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12
MIXED_TOL = 1e-5

SDS_ROLL = dict(rolling_friction=0.2, rolling_stiffness=2e-2, rolling_damping=1e-8)
SDS_TWIST = dict(twisting_friction=0.2, twisting_stiffness=2e-2, twisting_damping=1e-8)
CONFIGS = {
    "sjkr": (dict(cohesion_energy=5e3), {}),
    "jkr": (dict(surface_energy=0.05), {}),
    "dmt": (dict(surface_energy=0.05), dict(dmt=True)),
    "hooke": (dict(kn=2e4, kt=1e4), dict(hooke=True)),
    "hooke_sjkr": (dict(kn=2e4, kt=1e4, cohesion_energy=5e3), dict(hooke=True)),
    "roll_const": (dict(rolling_friction=0.2), {}),
    "roll_sds": (SDS_ROLL, dict(rolling_sds=True)),
    "twist_const": (dict(twisting_friction=0.2), {}),
    "twist_sds": (SDS_TWIST, dict(twisting_sds=True)),
    "all_sds": (dict(SDS_ROLL, **SDS_TWIST, cohesion_energy=5e3),
                dict(rolling_sds=True, twisting_sds=True)),
}


@pytest.fixture(autouse=True)
def _announce(request):
    print(f"\n[gpu case] {request.node.name}", flush=True)
    sys.stdout.flush()
    yield


def by_tag(tags, arr):
    return np.asarray(arr)[np.argsort(tags)]


def _dense_dem(n=1500, seed=5):
    # a gas compressed so that many spheres overlap slightly, spinning
    w = W.dem_gas(n=n, box=0.025 * (n / 500) ** (1 / 3), seed=seed)
    w.pos = w.pos * 0.83
    w.hi = w.hi * 0.83
    w.vel = (W.Stream(seed + 1).uniform(3 * n).reshape(-1, 3) - 0.5) * 0.4
    w.omega = (W.Stream(seed + 2).uniform(3 * n).reshape(-1, 3) - 0.5) * 400.0
    return w


def _pair(w, config, prec):
    extras, models = CONFIGS[config]
    o = oracle_from_workload(w)
    o.set_material_extras(**{k: [v] for k, v in extras.items()})
    o.set_contact_models(**models)
    s = hip.from_workload(w, prec, box_exit=True)
    s.set_material_extras(**{k: [v] for k, v in extras.items()})
    s.set_contact_models(**models)
    return o, s


def _canonical_history(a, b, h):
    # (lower tag, higher tag) order: the checker lists (owner, partner)
    return canon_history(a, b, h)


@pytest.mark.parametrize("prec", [hip.FP64, hip.MIXED])
@pytest.mark.parametrize("config", list(CONFIGS))
def test_force_passes_match_oracle(config, prec):
    w = _dense_dem()
    o, s = _pair(w, config, prec)
    o.neighbor_setup()
    o.pbc()
    o.sort_atoms_by_bin()
    o.borders()
    o.bin_neighbor_list()
    s.bin_neighbor_list()
    for _ in range(4):
        o.zero_all_forces()
        o.hertz_mindlin_contact_force()
        o.reverse_send_force()
        s.hertz_mindlin_contact_force()
    ref = o.atoms()
    st = s.state()
    f, t = s.forces()
    tol = FP64_TOL if prec == hip.FP64 else MIXED_TOL
    assert rel_err(by_tag(st["tag"], f), by_tag(ref["tag"], ref["force"])) < tol
    assert rel_err(by_tag(st["tag"], t), by_tag(ref["tag"], ref["torque"])) < tol
    ga, gb, gh = _canonical_history(*s.history7())
    oa, ob, oh = _canonical_history(*o.history7())
    assert len(ga) > 100
    assert np.array_equal(ga, oa) and np.array_equal(gb, ob)
    # spring and rolling / twist values: FP64 pair arithmetic in both precisions
    assert rel_err(gh, oh) < 1e-10
    if config in ("roll_sds", "all_sds"):
        assert np.abs(gh[:, 3:6]).max() > 0
    if config in ("twist_sds", "all_sds"):
        assert np.abs(gh[:, 6]).max() > 0


@pytest.mark.parametrize("config", ["jkr", "hooke", "all_sds", "twist_const"])
def test_scheduled_runs_match_oracle(config):
    w = _dense_dem(n=800, seed=11)
    o, s = _pair(w, config, hip.FP64)
    o.run(150)
    s.run(150)
    ref = o.atoms()
    st = s.state()
    L = w.hi - w.lo
    dp = by_tag(st["tag"], st["pos"]) - by_tag(ref["tag"], ref["pos"])
    dp -= L * np.round(dp / L)
    assert np.abs(dp).max() < 1e-9 * L.max()
    assert rel_err(by_tag(st["tag"], st["vel"]), by_tag(ref["tag"], ref["vel"])) < 1e-8
    assert rel_err(by_tag(st["tag"], st["omega"]), by_tag(ref["tag"], ref["omega"])) < 1e-7
    assert s.thermo()["rebuilds"] == o.rebuilds


def test_history_slots_survive_a_rebuild():
    w = _dense_dem(n=1200, seed=9)
    extras, models = CONFIGS["all_sds"]
    s = hip.from_workload(w, hip.FP64, box_exit=False)
    s.set_material_extras(**{k: [v] for k, v in extras.items()})
    s.set_contact_models(**models)
    s.run(20)
    a0, b0, h0 = _canonical_history(*s.history7())
    s.bin_neighbor_list()  # rebuild at the same positions: new rows, remapped history
    a1, b1, h1 = _canonical_history(*s.history7())
    assert len(a0) > 100
    assert np.array_equal(a0, a1) and np.array_equal(b0, b1) and np.array_equal(h0, h1)


def test_plain_model_is_restored_and_rejects_bad_tables():
    w = _dense_dem(n=600, seed=4)
    s = hip.from_workload(w, hip.FP64)
    with pytest.raises(ValueError):
        s.set_material_extras(kn=[1e4, 2e4])  # two values for one material
    with pytest.raises(hip.MddError):
        s._c(s.L.mdd_set_material_extras(s.h, 2, None))  # wrong material count
    s.set_material_extras(rolling_friction=[0.2])
    # new materials clear the optional tables: plain Hertz-Mindlin again,
    # bit for bit the run of a fresh system
    m = w.materials
    s.set_materials(m["youngs"], m["poisson"], m["restitution"], m["friction"])
    s2 = hip.from_workload(w, hip.FP64)
    s.run(30)
    s2.run(30)
    assert np.array_equal(s.state()["vel"], s2.state()["vel"])
    h = s.history7()[2]
    assert h.shape[1] == 7 and len(h) > 0 and np.all(h[:, 3:] == 0.0)
