"""Runs without a thermostat defer each step's final half step (Verlet kick,
final rotation) into the next step's first pass. Every mdd_run ends with the
deferred half issued on its own, so run(N) and N x run(1) integrate the same
steps with the passes fused and unfused: the states must be bitwise equal."""
import sys

import numpy as np
import pytest

from mddem_amd import hip
from mddem_amd import workloads as W
from tests.test_gpu_parity import _dense_dem

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _announce(request):
    print(f"\n[gpu case] {request.node.name}", flush=True)
    sys.stdout.flush()
    yield


def _state(s):
    st = s.state()
    order = np.argsort(st["tag"])
    f, t = s.forces()
    out = [st["pos"][order], st["vel"][order], f[order]]
    if st.get("omega") is not None:
        out += [st["omega"][order], st["quat"][order], t[order]]
    return out


@pytest.mark.parametrize("case", ["dem_mixed", "dem_fp64", "lj_nve_mixed"])
def test_deferred_final_half_matches_separate_passes(case):
    if case.startswith("dem"):
        w = _dense_dem(n=1500, seed=31)
        w.vel = w.vel * 5.0
        w.skin_fraction = 1.05
        steps = 240
        prec = hip.MIXED if case.endswith("mixed") else hip.FP64
    else:
        w = W.lj_fcc(ncell=6, thermostat=False)
        steps = 120
        prec = hip.MIXED
    a = hip.from_workload(w, prec, box_exit=False)
    a.run(steps)
    b = hip.from_workload(w, prec, box_exit=False)
    for _ in range(steps):
        b.run(1)
    ta, tb = a.thermo(), b.thermo()
    assert ta["step"] == tb["step"] == steps
    assert ta["rebuilds"] == tb["rebuilds"] >= 2
    for x, y in zip(_state(a), _state(b)):
        assert np.array_equal(x, y)
