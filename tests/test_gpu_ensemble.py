"""Batched ensemble: every system of a batched context follows the same
trajectory as the same system run alone (single-system GPU path) and as the
CPU checker, within the FP64 tolerance; rebuild timing is shared across the
batch, which only changes summation order."""
import numpy as np
import pytest

from mddem_amd import ensemble, hip
from mddem_amd import workloads as W
from tests.helpers import oracle_from_workload, rel_err

pytestmark = pytest.mark.gpu


def _systems():
    spec = [(200, 0.05, 0.95, 0.4), (350, 0.25, 0.7, 0.6), (900, 0.12, 0.85, 0.1),
            (1500, 0.3, 0.99, 0.3), (2600, 0.18, 0.75, 0.5)]
    return [W.gas_system(n, phi, e, mu, seed=1000 + k) for k, (n, phi, e, mu) in enumerate(spec)]


def _by_tag(st):
    order = np.argsort(st["tag"])
    return st["pos"][order], st["vel"][order]


def _periodic_diff(a, b, L):
    d = a - b
    return d - L * np.round(d / L)


def _collided(w, v):
    # spheres whose velocity changed: only contacts change one in these gases
    return int(np.sum(np.linalg.norm(v - w.vel, axis=1) > 1e-12 * np.abs(w.vel).max()))


def test_batched_systems_match_alone_runs():
    # 4000 steps: the denser gases have had collisions by then (a 100-step
    # window has none)
    systems = _systems()
    s = ensemble.make_context(systems, hip.FP64, box_exit=True)
    assert s.nsystems() == len(systems)
    s.run(4000)
    got = ensemble.system_states(s, len(systems))
    hits = 0
    for w, (p, v) in zip(systems, got):
        a = hip.from_workload(w, hip.FP64, box_exit=True)
        a.run(4000)
        pa, va = _by_tag(a.state())
        L = w.hi[0] - w.lo[0]
        assert np.abs(_periodic_diff(p, pa, L)).max() < 1e-12 * L
        assert rel_err(v, va) < 1e-9
        hits += _collided(w, v)
        a.close()
    assert hits > 20


def test_batched_system_matches_cpu_checker():
    systems = [_systems()[k] for k in (1, 3)]
    s = ensemble.make_context(systems, hip.FP64, box_exit=True)
    s.run(4000)
    got = ensemble.system_states(s, len(systems))
    hits = 0
    for w, (p, v) in zip(systems, got):
        o = oracle_from_workload(w)
        o.run(4000)
        ref = o.atoms()
        order = np.argsort(ref["tag"])
        L = w.hi[0] - w.lo[0]
        assert np.abs(_periodic_diff(p, ref["pos"][order], L)).max() < 1e-12 * L
        assert rel_err(v, ref["vel"][order]) < 1e-9
        hits += _collided(w, v)
    assert hits > 10


def test_small_ensemble_run_reports():
    rep = ensemble.run(nsys=24, steps=300, segment=100, tune=False)
    assert rep["systems"] == 24 and rep["value"] > 0
    # inelastic collisions only remove energy (Haff cooling); over 300 steps
    # many systems have had no collision yet
    assert rep["haff_tg_ratio_max"] <= 1.0 + 1e-9


def test_batched_systems_at_one_lane_match_alone_runs():
    # one lane per sphere: a system run alone splits its force pass into the
    # spheres with neighbors and the rest; the batch runs one pass
    systems = _systems()
    s = ensemble.make_context(systems, hip.FP64, box_exit=True)
    s.set_force_lanes(0)
    s.run(150)
    got = ensemble.system_states(s, len(systems))
    for w, (p, v) in zip(systems, got):
        a = hip.from_workload(w, hip.FP64, box_exit=True)
        a.set_force_lanes(0)
        a.run(150)
        pa, va = _by_tag(a.state())
        L = w.hi[0] - w.lo[0]
        assert np.abs(_periodic_diff(p, pa, L)).max() < 1e-12 * L
        assert rel_err(v, va) < 1e-9
        a.close()


def test_batched_contact_queue_gives_the_same_run():
    # one lane per sphere in a batch: the wavefront's pairs in contact queued
    # and computed 64 at a time (each sphere's system materials looked up by
    # the queued pair's owner) against each lane's own loop, bit for bit
    systems = _systems()
    speed0 = np.sort(np.concatenate([np.linalg.norm(w.vel, axis=1) for w in systems]))
    out = []
    for wave in (False, True):
        s = ensemble.make_context(systems, hip.MIXED, box_exit=True, lanes_log2=0, wave=wave)
        s.run(4000)
        st = s.state()
        f, t = s.forces()
        out.append((st["pos"], st["vel"], st["omega"], f, t))
        s.close()
    # collisions happened (a few dozen spheres by 4000 steps): only contacts
    # change a speed in these gases
    speed = np.sort(np.linalg.norm(out[0][1], axis=1))
    assert np.sum(np.abs(speed - speed0) > 1e-9 * speed0.max()) > 10
    for x, y in zip(out[0], out[1]):
        assert np.array_equal(x, y)
