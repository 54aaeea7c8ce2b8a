"""Continuous-batching serve loop on device 0.

- every job's final coordinates and energies are identical (bitwise) to the
  same job run alone on the plain engine API;
- jobs submitted while others run are admitted and overlap with them (their
  uploads beside the block in flight, or after it when a pool must grow);
- the job state resident at any time stays within the HBM budget;
- jobs that break the shared schedule are refused at submission.
"""
import time

import numpy as np
import pytest

import sop_amd as S
from sop_amd import serving
from sop_amd import workloads as W

pytestmark = pytest.mark.gpu


def job_systems():
    out = []
    for k, (n, steps) in enumerate([(150, 120), (900, 300), (2500, 200), (64, 50), (1200, 400), (3000, 100),
                                    (400, 250), (700, 0), (1800, 150), (260, 500)]):
        s = S.make_chain_system(500 + k, n)
        s.params.seed = W.philox_key(9, k)
        s.params.T = 0.5 + 0.04 * k
        out.append((s, steps))
    return out


def alone(s, steps):
    with S.Engine(0) as e:
        i = e.add(s)
        if steps:
            e.run([i], steps)
        return e.coords(i), e.energy_eval(i)


def test_served_results_equal_alone_and_overlap():
    jobs = job_systems()
    with S.Engine(0) as eng:
        srv = serving.Server(eng)
        ids = []
        for s, steps in jobs[:5]:
            ids.append(srv.submit(s, steps))
        time.sleep(0.05)  # the first jobs are running when the rest arrive
        for s, steps in jobs[5:]:
            ids.append(srv.submit(s, steps))
        res = [srv.wait(j) for j in ids]
        st = srv.stats()
        srv.stop()
    # the late jobs were uploaded beside a running block, or put off until the
    # engine was idle because a pool had to grow (both paths give the same results)
    assert st["jobs_admitted"] == len(jobs)
    assert st["jobs_uploaded_beside"] + st["jobs_deferred"] > 0, st
    for (s, steps), (r, x) in zip(jobs, res):
        assert r["status"] == 0
        xr, er = alone(s, steps)
        assert np.array_equal(x, xr), (s.n, steps)
        assert np.array_equal(r["energies"], er[:6])
    # continuous batching: some job was admitted while another was resident
    iv = sorted((r["t_admit"], r["t_done"]) for r, _ in res)
    assert any(iv[k + 1][0] < iv[k][1] for k in range(len(iv) - 1))


def test_admission_beside_a_block_and_after_pool_growth():
    # A (short) and B (long) size the pools; once A has retired, C fits A's
    # freed ranges and uploads beside B's running block; D is larger than
    # any free range, needs a pool to grow, and waits for the idle engine
    a, b = S.make_chain_system(611, 3000), S.make_chain_system(612, 64)
    c, d = S.make_chain_system(613, 2000), S.make_chain_system(614, 8000)
    systems, steps = (a, b, c, d), (100, 6000, 100, 100)
    for k, s in enumerate(systems):
        s.params.seed = W.philox_key(10, k)
    with S.Engine(0) as eng:
        srv = serving.Server(eng)
        ids = [srv.submit(a, steps[0]), srv.submit(b, steps[1])]
        srv.wait(ids[0], with_coords=False)
        ids.append(srv.submit(c, steps[2]))
        while srv.stats()["jobs_admitted"] < 3:
            time.sleep(0.001)
        ids.append(srv.submit(d, steps[3]))
        res = [srv.wait(j) for j in ids]
        st = srv.stats()
        srv.stop()
    print("admission paths:", st, flush=True)
    assert all(r["status"] == 0 for r, _ in res)
    assert res[1][0]["t_done"] > res[3][0]["t_admit"]  # D was admitted while B ran
    assert st["jobs_uploaded_beside"] >= 1, st
    assert st["jobs_deferred"] >= 1, st
    for s, n, (r, x) in zip(systems, steps, res):
        xr, er = alone(s, n)
        assert np.array_equal(x, xr), s.n
        assert np.array_equal(r["energies"], er[:6])


def test_large_served_jobs_equal_alone():
    # jobs above the LDS size (global gathers on the side stream, four waves
    # per tile in a small batch), up to the served stream's 20,000-bead
    # maximum, served beside small ones
    systems = [(S.make_chain_system(620 + k, n), st) for k, (n, st) in
               enumerate([(7000, 150), (300, 400), (12000, 100), (1500, 250), (20000, 60)])]
    for k, (s, _) in enumerate(systems):
        s.params.seed = W.philox_key(11, k)
    with S.Engine(0) as eng:
        srv = serving.Server(eng)
        ids = [srv.submit(s, st) for s, st in systems]
        res = [srv.wait(j) for j in ids]
        srv.stop()
    for (s, st), (r, x) in zip(systems, res):
        assert r["status"] == 0
        xr, er = alone(s, st)
        assert np.array_equal(x, xr), s.n
        assert np.array_equal(r["energies"], er[:6])


def test_admission_respects_hbm_budget():
    jobs = job_systems()
    with S.Engine(0) as eng:
        probe = serving.Server(eng)
        sizes = [probe.job_bytes(s) for s, _ in jobs]
        probe.stop()
        budget = int(sorted(sizes)[-1] + sorted(sizes)[-2])  # two largest jobs
        srv = serving.Server(eng, hbm_budget_bytes=budget)
        ids = [srv.submit(s, steps) for s, steps in jobs]
        res = [srv.wait(j, with_coords=False)[0] for j in ids]
        srv.stop()
    events = []
    for r, b in zip(res, sizes):
        assert r["status"] == 0
        events += [(r["t_admit"], b), (r["t_done"], -b)]
    events.sort(key=lambda ev: (ev[0], ev[1]))
    resident = peak = 0
    for _, b in events:
        resident += b
        peak = max(peak, resident)
    assert peak <= budget


def test_schedule_mismatch_is_refused():
    with S.Engine(0) as eng:
        srv = serving.Server(eng)
        a = S.make_chain_system(1, 100)
        srv.submit(a, 10)
        b = S.make_chain_system(2, 100)
        b.params.n_pl = 5
        with pytest.raises(S.SopError):
            srv.submit(b, 10)
        c = S.make_chain_system(3, 100)
        c.step0 = 7
        with pytest.raises(S.SopError):
            srv.submit(c, 10)
        srv.stop()


def test_loaded_server_whole_cu_tier_equals_alone():
    # more tiles than two LDS launches' wave slots (a loaded server): jobs of
    # 5,057-10,112 beads take the whole-CU LDS tier beside the two-blocks-per-
    # unit one; every result still equals the job run alone
    rng = np.random.default_rng(12)
    sizes = [int(v) for v in rng.integers(1500, 3000, 320)] + [6000, 7500, 9000, 12000]
    systems = []
    for k, n in enumerate(sizes):
        s = S.make_chain_system(900 + k % 40, n)
        s.params.seed = W.philox_key(13, k)
        s.params.T = 0.55 + 0.001 * k
        systems.append(s)
    steps = 200
    with S.Engine(0) as eng:
        srv = serving.Server(eng, hbm_budget_bytes=64 << 30)
        try:
            ids = [srv.submit(s, steps) for s in systems]
            res = [srv.wait(j) for j in ids]
        finally:
            srv.stop()
    assert all(r["status"] == 0 for r, _ in res)
    check = [0, 150, len(sizes) - 4, len(sizes) - 3, len(sizes) - 2, len(sizes) - 1]
    with S.Engine(0) as e:
        for k in check:
            i = e.add(systems[k])
            e.run([i], steps)
            assert np.array_equal(res[k][1], e.coords(i)), (k, sizes[k])
            assert np.array_equal(res[k][0]["energies"], e.energy_eval(i)[:6]), k
            e.remove(i)
