"""Restart snapshots of the GPU state in the reference's RestartData formats:
the uploaded snapshot equals the state it was taken from, and a run continued
from it follows the uninterrupted run.

The reference's restart holds no contact history (ContactHistoryStore is not
per-atom data) and no thermostat variable, so the DEM case restarts at a
moment without active contacts and the LJ case runs without a thermostat.
The continued run differs from the uninterrupted one only in summation order
(the first rebuild after the upload orders particles differently)."""
import sys

import numpy as np
import pytest

from mddem_amd import hip
from mddem_amd import particles_io as io
from mddem_amd import workloads as W
from tests.helpers import rel_err

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _announce(request):
    print(f"\n[gpu case] {request.node.name}", flush=True)
    sys.stdout.flush()
    yield


def _min_image_err(pa, pb, L):
    # the restarted run wraps at its first rebuild, the original only at its
    # next one (lazy wrap): compare positions modulo the periodic box
    d = pb - pa
    d -= L * np.round(d / L)
    return float(np.abs(d).max() / np.max(L))


def _by_tag(s):
    st = s.state()
    o = np.argsort(st["tag"])
    f, t = s.forces()
    out = dict(pos=st["pos"][o], vel=st["vel"][o], force=f[o])
    if "omega" in st:
        out.update(omega=st["omega"][o], quat=st["quat"][o], torque=t[o])
    return out


@pytest.mark.parametrize("fmt", ["bincode", "json"])
def test_dem_restart_round_trip_and_continuation(tmp_path, fmt):
    w = W.dem_gas(n=500)
    a = hip.from_workload(w, hip.FP64, box_exit=False)
    a.run(100)
    for _ in range(40):  # a moment without active contacts (no history to lose)
        if a.thermo()["ncontacts"] == 0:
            break
        a.run(5)
    assert a.thermo()["ncontacts"] == 0
    step = a.thermo()["step"]
    snap = io.restart_from_system(a, step, w.dt, density=2500.0)
    path = str(tmp_path / io.restart_file_name(step, 0, fmt))
    io.write_restart(path, snap, fmt)
    b = hip.from_workload(w, hip.FP64, box_exit=False)
    io.upload_restart(b, io.read_restart(path))
    sa, sb = _by_tag(a), _by_tag(b)
    for k in sa:
        assert np.array_equal(sa[k], sb[k]), k
    a.run(100)
    b.run(100)
    sa, sb = _by_tag(a), _by_tag(b)
    assert _min_image_err(sa["pos"], sb["pos"], w.hi - w.lo) < 1e-9
    for k in ("vel", "omega"):
        assert rel_err(sb[k], sa[k]) < 1e-9, k


def test_lj_restart_continuation(tmp_path):
    w = W.lj_fcc(ncell=5, thermostat=False)
    a = hip.from_workload(w, hip.FP64, box_exit=False)
    a.run(150)
    snap = io.restart_from_system(a, 150, w.dt)
    path = str(tmp_path / io.restart_file_name(150, 0, "json"))
    io.write_restart(path, snap, "json")
    b = hip.from_workload(w, hip.FP64, box_exit=False)
    io.upload_restart(b, io.read_restart(path))
    sa, sb = _by_tag(a), _by_tag(b)
    for k in sa:
        assert np.array_equal(sa[k], sb[k]), k
    a.run(100)
    b.run(100)
    sa, sb = _by_tag(a), _by_tag(b)
    assert _min_image_err(sa["pos"], sb["pos"], w.hi - w.lo) < 1e-9
    assert rel_err(sb["vel"], sa["vel"]) < 1e-9


def test_restart_file_runs_the_same_on_oracle_and_gpu(tmp_path):
    # one RestartData file (bincode) read into the checker and into the HIP
    # path (read_restart, mddem_print/src/lib.rs:1222-1246), then the same
    # 300 steps on both: the file is the only state they share
    from tests.helpers import oracle_from_workload
    w = W.dem_gas(n=500)
    a = hip.from_workload(w, hip.FP64, box_exit=True)
    a.run(120)
    for _ in range(40):
        if a.thermo()["ncontacts"] == 0:
            break
        a.run(5)
    step = a.thermo()["step"]
    path = str(tmp_path / io.restart_file_name(step, 0, "bincode"))
    io.write_restart(path, io.restart_from_system(a, step, w.dt, density=2500.0), "bincode")
    data = io.read_restart(path)
    b = hip.from_workload(w, hip.FP64, box_exit=True)
    io.upload_restart(b, data)
    dem = io.unpack_dem(data["atom_data_buffers"][0])
    pos = np.stack([data["pos_x"], data["pos_y"], data["pos_z"]], axis=1)
    vel = np.stack([data["vel_x"], data["vel_y"], data["vel_z"]], axis=1)
    force = np.stack([data["force_x"], data["force_y"], data["force_z"]], axis=1)
    o = oracle_from_workload(w)
    o.set_atoms(pos, vel, np.asarray(data["mass"]), np.asarray(data["cutoff_radius"]),
                radius=dem["radius"], omega=dem["omega"], tag=np.asarray(data["tag"]),
                atype=np.asarray(data["atom_type"]))
    o.set_forces(force, dem["torque"])
    b.run(300)
    o.run(300)
    sb = _by_tag(b)
    ref = o.atoms()
    k = np.argsort(ref["tag"])
    assert _min_image_err(ref["pos"][k], sb["pos"], w.hi - w.lo) < 1e-12
    assert rel_err(sb["vel"], ref["vel"][k]) < 1e-9
    assert rel_err(sb["omega"], ref["omega"][k]) < 1e-6
    assert b.thermo()["rebuilds"] == o.rebuilds
