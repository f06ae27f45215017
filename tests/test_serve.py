"""Serving loop: admission policy and the job stream on the CPU; on the GPU,
every job's result is bitwise the same alone and in the mix."""
import numpy as np
import pytest

from mddem_amd import serve


def _job(jid, kind, n, arrival=0.0):
    return serve.Job(jid, kind, n, 100, arrival)


def test_admit_is_fifo_and_respects_budget_and_slots():
    q = [_job(0, "lj", 10_000), _job(1, "dem", 50_000), _job(2, "lj", 1_000)]
    budget = q[0].footprint + q[1].footprint - 1
    got = serve.admit(q, 0, 0, budget, slots=8)
    assert [j.jid for j in got] == [0]          # job 1 does not fit; job 2 may not overtake
    assert [j.jid for j in q] == [1, 2]
    got = serve.admit(q, 0, 0, budget, slots=8)
    assert [j.jid for j in got] == [1, 2] or [j.jid for j in got] == [1]
    q = [_job(k, "lj", 4000) for k in range(5)]
    got = serve.admit(q, 0, 6, 1 << 40, slots=8)
    assert len(got) == 2 and len(q) == 3        # two free slots


def test_stream_is_seeded_and_in_range():
    a = serve.make_stream(200, rate=5.0)
    b = serve.make_stream(200, rate=5.0)
    assert [(j.kind, j.n_target, j.steps, j.arrival) for j in a] == \
           [(j.kind, j.n_target, j.steps, j.arrival) for j in b]
    dem = [j for j in a if j.kind == "dem"]
    lj = [j for j in a if j.kind == "lj"]
    assert 0.55 < len(dem) / len(a) < 0.85
    assert all(1000 <= j.n_target <= 200_000 for j in dem)
    assert all(4000 <= j.n_target <= 500_000 for j in lj)
    assert all(1000 <= j.steps <= 50_000 for j in a)
    arr = np.array([j.arrival for j in a])
    assert np.all(np.diff(arr) > 0)
    assert 0.1 < np.mean(np.diff(arr)) < 0.4    # rate 5/s


@pytest.mark.gpu
def test_jobs_are_identical_alone_and_in_the_mix():
    from mddem_amd import hip
    jobs = [serve.Job(0, "dem", 2500, 300, 0.0, dict(seed=11)),
            serve.Job(1, "lj", 4000, 300, 0.0, dict(rho=0.85, temperature=1.0, seed=12)),
            serve.Job(2, "dem", 12_000, 250, 0.01, dict(seed=13)),
            serve.Job(3, "lj", 6912, 200, 0.02, dict(rho=0.82, temperature=0.8, seed=14)),
            serve.Job(4, "dem", 3000, 400, 0.03, dict(seed=15))]
    for j in jobs:
        j.workload = serve.build_workload(j)
    alone = [serve.run_job(j, hip.MIXED) for j in jobs]
    rep = serve.Server(hbm_budget=8 << 30, slots=8).run(jobs)
    assert rep["failed"] == 0
    assert rep["max_concurrent"] >= 2
    for j, a in zip(jobs, alone):
        assert j.digest["sha256"] == a["sha256"], j.jid
        assert j.digest["step"] == j.steps


@pytest.mark.gpu
def test_graph_rebuilds_with_uploads_and_copies_in_other_threads():
    # later jobs upload their state while earlier ones capture and replay
    # rebuild graphs on their own streams; every job must finish, match its
    # run alone bit for bit, and have replayed rebuild graphs
    from mddem_amd import hip
    jobs = []
    for k in range(10):
        if k % 2:
            jobs.append(serve.Job(k, "lj", 4000 + 2000 * k, 500, 0.005 * k,
                                  dict(rho=0.85, temperature=1.0, seed=40 + k)))
        else:
            jobs.append(serve.Job(k, "dem", 3000 + 1000 * k, 700, 0.005 * k, dict(seed=40 + k)))
    for j in jobs:
        j.workload = serve.build_workload(j)
    alone = [serve.run_job(j, hip.MIXED) for j in jobs]
    rep = serve.Server(hbm_budget=16 << 30, slots=10).run(jobs)
    assert rep["failed"] == 0, [j.error for j in jobs if j.error]
    assert rep["max_concurrent"] >= 3
    for j, a in zip(jobs, alone):
        assert j.digest["sha256"] == a["sha256"], j.jid
    assert sum(j.digest["graph_replays"] for j in jobs) > 0


@pytest.mark.gpu
def test_jobs_served_in_the_mix_match_the_oracle():
    # three jobs (two granular gases, one thermostatted LJ fluid) served
    # concurrently in FP64 with the reference's rebuild rule; each final
    # state against the CPU checker running that job's workload alone
    from mddem_amd import hip
    from tests.helpers import oracle_from_workload, rel_err
    # the gases run 3000 steps, by which about 50 spheres of each have collided
    jobs = [serve.Job(0, "dem", 1500, 3000, 0.0, dict(seed=21)),
            serve.Job(1, "lj", 2048, 300, 0.0, dict(rho=0.85, temperature=1.0, seed=22)),
            serve.Job(2, "dem", 2500, 3000, 0.0, dict(seed=23))]
    for j in jobs:
        j.workload = serve.build_workload(j)
    rep = serve.Server(hbm_budget=8 << 30, slots=4, precision=hip.FP64, box_exit=True,
                       keep_state=True).run(jobs)
    assert rep["failed"] == 0 and rep["max_concurrent"] >= 2
    for j in jobs:
        o = oracle_from_workload(j.workload)
        o.run(j.steps)
        ref = o.atoms()
        a, b = np.argsort(j.final["tag"]), np.argsort(ref["tag"])
        L = j.workload.hi - j.workload.lo
        dp = j.final["pos"][a] - ref["pos"][b]
        dp -= L * np.round(dp / L)
        assert np.abs(dp).max() < 1e-12 * L.max() * 1e3, j.jid
        assert rel_err(j.final["vel"][a], ref["vel"][b]) < 1e-9, j.jid
        assert j.final["rebuilds"] == o.rebuilds, j.jid
        if j.kind == "dem":
            v0 = j.workload.vel[np.argsort(j.workload.tag)]
            assert np.sum(np.linalg.norm(j.final["vel"][a] - v0, axis=1) > 1e-12) > 20, j.jid
