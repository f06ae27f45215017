"""mdd_compute_thermo_noalloc (static device scratch, no allocation) against
mdd_compute_thermo through the C ABI: equal bit for bit on two calls in a
row, the second on a smaller system, and a clear error past the scratch's
particle limit."""
import sys

import numpy as np
import pytest

from mddem_amd import hip
from mddem_amd import workloads as W

pytestmark = pytest.mark.gpu

FIELDS = ("step", "rebuilds", "npairs", "nghost", "ncontacts", "ke", "temperature", "p_xi",
          "momentum", "granular_temperature")


@pytest.fixture(autouse=True)
def _announce(request):
    print(f"\n[gpu case] {request.node.name}", flush=True)
    sys.stdout.flush()
    yield


def same(a, b):
    for k in FIELDS:
        assert a[k] == b[k], (k, a[k], b[k])
    for k in ("pe", "pressure"):
        assert b[k] == 0.0
    assert b["virial"] == [0.0] * 6


def test_noalloc_thermo_equals_thermo_on_two_calls():
    w1 = W.dem_packed_bed(n=6000, box=(0.03, 0.03, 1.2), vmax=0.5)
    s1 = hip.from_workload(w1, hip.MIXED)
    s1.run(120)
    a = s1.thermo()
    b = s1.thermo_noalloc()
    same(a, b)
    assert b["npairs"] > 0 and b["ncontacts"] > 0
    # the second call: a smaller system, another context on the same GPU
    w2 = W.lj_fcc(ncell=4)
    s2 = hip.from_workload(w2, hip.FP64)
    s2.run(60)
    a2 = s2.thermo()
    b2 = s2.thermo_noalloc()
    same(a2, b2)
    assert b2["npairs"] == a2["npairs"] > 0 and b2["p_xi"] != 0.0
    # and the larger one again: nothing of the smaller call leaks into it
    same(s1.thermo(), s1.thermo_noalloc())


def test_noalloc_thermo_rejects_beyond_the_static_limit():
    n = 8388608 + 1
    L = 500.0
    s = hip.System(hip.MODEL_LJ, hip.FP64)
    s.set_box((0, 0, 0), (L, L, L), (True, True, True))
    s.set_lj(1.0, 1.0, 2.5)
    pos = W.Stream(5).uniform(3 * n).reshape(n, 3) * L
    s.upload(pos, np.zeros((n, 3)), np.ones(n), np.full(n, 1.25))
    del pos
    with pytest.raises(hip.MddError) as e:
        s.thermo_noalloc()
    assert e.value.code == -1 and "8388608" in str(e.value)
    assert s.thermo()["ke"] == 0.0  # the allocating entry point still serves it
