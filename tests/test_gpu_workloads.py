"""The batched and served workloads at their own shapes, against the oracle.

C3  the 512-replica batch (N log-uniform in [200, 5000], T uniform in
    [0.5, 0.9], distinct Philox keys) runs 200 steps in one batch, crossing
    neighbor-list snapshots every 50 steps and pair-list rebuilds from shell
    rows and from the cell grid.  Eight replicas are then checked against the
    oracle at their coordinates: the smallest and the largest, two that share
    one staged LDS range, the one closest below the 5,056-bead LDS limit (C3
    has no replica above it), one whose pair list was rebuilt from the cell
    grid after the first build, and two more (all at step 200).  Lists bit-exact as sorted sets,
    forces within 1e-5 of the largest force, and, after the whole batch runs
    1,000 more steps, trajectories within 1e-4 A RMSD of the FP64 oracle run
    from the same coordinates and Philox stream.
Shared staged ranges: ten consecutive systems of 64-500 beads, every other one
    with its coordinates in the other buffer (odd parity), share staged LDS
    ranges in a batch; each equals its standalone run bitwise, and one of odd
    parity matches the oracle.
C4  64 jobs with the served stream's sizes (log-normal, median 1,500, clipped
    to [100, 20,000]) and one 20,000-bead job, submitted in four waves to one
    server: every result equals the job run alone bitwise, and two jobs match
    the oracle trajectory.  Step counts are 50-400 per job (the stream's
    1e4-1e6 steps would take minutes per job).
"""
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import sop_amd as S
from sop_amd import serving
from sop_amd import workloads as W
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu

LDS_BEADS = 5056  # default lds_beads_max: largest system staged in LDS


def pair_set(a):
    a = np.asarray(a, np.int64).reshape(-1, 2)
    return set(map(tuple, a.tolist()))


def check_lists_forces(eng, sid, s, x):
    """Pair and neighbor lists of system sid rebuilt at x bit-exact against the
    oracle's, and its forces on that pair list within 1e-5 of the largest."""
    o = Oracle(s)
    (nla, nlr), (pla, plr) = o.lists(x, grid=True)
    nat = np.stack([s.nat_i, s.nat_j], 1)
    eng.update_pair_list(sid)
    assert pair_set(eng.get_list(sid, S.LIST_PL_REP)) == pair_set(plr), s.n
    assert pair_set(eng.get_list(sid, S.LIST_PL_NATIVE)) == pair_set(nat[pla]), s.n
    fr, _, st = o.forces(x, (pla, plr))
    assert st == 0
    f = eng.force_eval(sid)
    err = np.linalg.norm(f - fr, axis=1).max() / np.linalg.norm(fr, axis=1).max()
    assert err <= 1e-5, (s.n, err)
    eng.update_neighbor_list(sid)
    assert pair_set(eng.get_list(sid, S.LIST_NL_REP)) == pair_set(nlr), s.n
    assert pair_set(eng.get_list(sid, S.LIST_NL_NATIVE)) == pair_set(nat[nla]), s.n


def oracle_final(s, x0, step0, nsteps):
    a5, var = S.params_derived(s.params)
    return Oracle(s).run(x0, step0, nsteps, s.params.seed, a5, var, s.params.n_nl, s.params.n_pl, grid=True)


def rmsd(a, b):
    return float(np.sqrt(((a - b) ** 2).sum(1).mean()))


def shared_items(sizes, per, lds_cap=LDS_BEADS):
    """Groups of consecutive systems (arena order) that the engine stages as
    one LDS range: adjacent systems join an item while its tiles stay within
    `per` and its beads within lds_cap; a system of more than `per` tiles
    gets items of its own."""
    groups, cur, tiles, beads = [], [], 0, 0
    for k, n in enumerate(sizes):
        nt = -(-n // 64)
        if nt > per:
            if cur:
                groups.append(cur)
            groups.append([k])
            cur, tiles, beads = [], 0, 0
            continue
        if cur and tiles + nt <= per and beads + 64 * nt <= lds_cap:
            cur.append(k)
            tiles += nt
            beads += 64 * nt
        else:
            if cur:
                groups.append(cur)
            cur, tiles, beads = [k], nt, 64 * nt
    if cur:
        groups.append(cur)
    return groups


def test_c3_batch_sampled_replicas_match_oracle():
    K0 = 200
    systems = W.c3(512)
    sizes = np.array([s.n for s in systems])
    with S.Engine(0) as eng:
        ids = [eng.add(s) for s in systems]
        eng.run(ids, K0)
        assert all(eng.step_of(i) == K0 for i in ids)
        # the full batch has more tiles than wave slots: one wave per tile, 16 tiles per item
        groups = shared_items(sizes, 16)
        pair = next(g for g in groups if len(g) >= 2)[:2]
        regrid = [r for r in range(len(systems)) if eng.list_stats(ids[r]).grid_builds >= 2]
        assert regrid, f"no replica rebuilt its shell rows from the cell grid within {K0} steps"
        below = int(np.argmax(np.where(sizes <= LDS_BEADS, sizes, -1)))
        pick = [int(np.argmin(sizes)), int(np.argmax(sizes)), *pair, below, regrid[0]]
        rng = np.random.default_rng(7)
        while len(set(pick)) < 8:
            pick.append(int(rng.integers(len(systems))))
        pick = sorted(set(pick))
        print("C3 replicas checked:", [(r, int(sizes[r])) for r in pick], "sharing:", pair, flush=True)
        x0 = {r: eng.coords(ids[r]) for r in pick}
        for r in pick:
            check_lists_forces(eng, ids[r], systems[r], x0[r])
        eng.run(ids, 1000)
        got = {r: eng.coords(ids[r]) for r in pick}
    with ThreadPoolExecutor(8) as ex:
        ref = dict(zip(pick, ex.map(lambda r: oracle_final(systems[r], x0[r], K0, 1000), pick)))
    for r in pick:
        d = rmsd(got[r], ref[r])
        print(f"replica {r} ({sizes[r]} beads): 1000-step rmsd {d:.2e} A", flush=True)
        assert d <= 1e-4, (r, sizes[r], d)


def test_systems_sharing_a_staged_item_equal_alone():
    sizes = [64, 100, 130, 200, 64, 300, 90, 180, 64, 450]
    systems = []
    for k, n in enumerate(sizes):
        s = S.make_chain_system(700 + k, n)
        s.params.seed = W.philox_key(12, k)
        s.params.T = 0.5 + 0.04 * k
        systems.append(s)
    # a small batch runs two waves per tile: items of up to 8 tiles
    assert sum(len(g) > 1 for g in shared_items(sizes, 8)) >= 3
    odd = set(range(1, len(sizes), 2))
    steps = 137

    def prepare(e, k):
        i = e.add(systems[k])
        if k in odd:  # one step moves the coordinates to the other buffer; the counter goes back to 0
            e.run([i], 1)
            e.set_rng(i, systems[k].params.seed, 0)
        return i

    with S.Engine(0) as e:
        ids = [prepare(e, k) for k in range(len(sizes))]
        x0 = {k: e.coords(ids[k]) for k in range(len(sizes))}
        e.run(ids, steps)
        batched = [e.coords(i) for i in ids]
    with S.Engine(0) as e:
        for k in range(len(sizes)):
            i = prepare(e, k)
            e.run([i], steps)
            assert np.array_equal(e.coords(i), batched[k]), (k, sizes[k])
            e.remove(i)
    k = 5
    assert k in odd
    s = systems[k]
    a5, var = S.params_derived(s.params)
    xr = Oracle(s).run(x0[k], 0, steps, s.params.seed, a5, var, s.params.n_nl, s.params.n_pl)
    assert rmsd(batched[k], xr) <= 1e-4


def test_c4_served_jobs_equal_alone_and_oracle():
    jobs = W.c4_jobs(rate=200.0, duration=10.0, seed=4)[:63]
    rng = np.random.default_rng(44)
    systems = [W.c4_system(j, 4) for j in jobs]
    big = S.make_chain_system(4 * 1000003 + 99999, 20000)
    big.params.seed = W.philox_key(4, 99999)
    big.params.T = 0.7
    systems.append(big)
    steps = [int(v) for v in np.round(np.exp(rng.uniform(np.log(50), np.log(400), len(systems))))]
    sizes = [s.n for s in systems]
    print("C4 job sizes:", sorted(sizes), flush=True)
    assert len(systems) >= 64 and max(sizes) == 20000
    with S.Engine(0) as eng:
        srv = serving.Server(eng)
        ids = []
        try:
            for w in range(4):  # four waves: later jobs arrive while earlier ones run
                for k in range(w * 16, (w + 1) * 16):
                    ids.append(srv.submit(systems[k], steps[k]))
                time.sleep(0.02)
            res = [srv.wait(j) for j in ids]
            st = srv.stats()
        finally:
            srv.stop()
    print("serve loop:", st, flush=True)
    assert st["jobs_admitted"] == len(systems)
    with S.Engine(0) as e:
        for k, (s, (r, x)) in enumerate(zip(systems, res)):
            assert r["status"] == 0, (k, r)
            i = e.add(s)
            e.run([i], steps[k])
            assert np.array_equal(x, e.coords(i)), (k, s.n, steps[k])
            assert np.array_equal(r["energies"], e.energy_eval(i)[:6]), k
            e.remove(i)
    small = [k for k in np.argsort(sizes) if sizes[k] >= 200][:2]
    for k in small:
        s = systems[k]
        d = rmsd(res[k][1], oracle_final(s, s.xyz, 0, steps[k]))
        print(f"job {k} ({s.n} beads, {steps[k]} steps): rmsd {d:.2e} A", flush=True)
        assert d <= 1e-4, (k, d)
