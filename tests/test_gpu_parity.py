"""Parity of the HIP path (libsop on device 0) with the oracle.

Tolerances (from the engine's specification):
  lists        bit-exact: identical sorted sets of (i<j) pairs
  forces       every bead: |F_gpu - F_ref| <= 1e-5 * max_i |F_ref,i|  (FP32 forces
               vs the FP64 oracle on the same coordinates and pair list)
  energies     FP64 sums of FP32 terms: each term within 2e-6 of the FP64 oracle,
               relative to the sum of the magnitudes of all terms
  contacts     exact integer count; Rg^2 relative 1e-9
  trajectory   RMSD <= 1e-4 A after 1,000 steps from the identical Philox stream
  batching     a system's coordinates are bitwise identical run alone, in a batch,
               on the plain per-system path and with or without hipGraphs
"""
import os

import numpy as np
import pytest

import sop_amd as S
from oracle import oracle as O
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def pair_set(a):
    a = np.asarray(a, np.int64).reshape(-1, 2)
    return set(map(tuple, a.tolist()))


def oracle_sets(o, s, xyz):
    (nla, nlr), (pla, plr) = o.lists(xyz)
    nat = np.stack([s.nat_i, s.nat_j], 1)
    return {S.LIST_NL_NATIVE: pair_set(nat[nla]), S.LIST_NL_REP: pair_set(nlr),
            S.LIST_PL_NATIVE: pair_set(nat[pla]), S.LIST_PL_REP: pair_set(plr)}, (pla, plr)


def c1():
    z = np.load(os.path.join(GOLD, "c1_system.npz"))
    return S.System(xyz=z["xyz"], params=S.params_config(), bond_r0=z["bond_r0"], nat_i=z["nat_i"],
                    nat_j=z["nat_j"], nat_r0=z["nat_r0"])


def multi_chain():
    """Three chains, mixed bead types, intra- and inter-chain natives."""
    xs, bs = [], []
    for k, n in enumerate((60, 70, 50)):
        x, b = S.generate_chain(10 + k, n)
        xs.append(x + np.array([k * 9.0, 0, 0], np.float32))
        bs.append(b)
    x = np.concatenate(xs)
    ch = np.repeat([0, 1, 2], (60, 70, 50)).astype(np.int32)
    ni, nj, nr = S.native_contacts(x, ch, 8.0)
    rng = np.random.default_rng(3)
    p = S.params_config()
    p.seed = 77
    return S.System(xyz=x, params=p, chain=ch, type=rng.integers(1, 3, len(x)).astype(np.int32),
                    bond_r0=np.concatenate(bs), nat_i=ni, nat_j=nj, nat_r0=nr)


def reference_box():
    """Reference parameters: 3.2x / 2.5x cutoffs, pair list every step, boxl 500."""
    s = S.make_chain_system(21, 200, params=S.params_reference())
    s.params.zeta = 50.0  # stable overdamped step at h = 0.02
    return s


def c2():
    return S.make_chain_system(2, 10000)


def tight_skin():
    """Neighbor cutoff 0.05 A above the pair cutoff: beads leave the skin
    within a few steps, so pair lists come from the materialized neighbor list."""
    s = S.make_chain_system(31, 400)
    s.params.nl_abs, s.params.pl_abs = 16.05, 16.0
    return s


def inverted_cutoffs():
    """Neighbor cutoff below the pair cutoff: PL = NL(14 A, snapshot) & (16 A, now)."""
    s = S.make_chain_system(32, 300)
    s.params.nl_abs, s.params.pl_abs = 14.0, 16.0
    return s


CASES = {"c1": c1, "multi": multi_chain, "refbox": reference_box, "c2": c2, "tight": tight_skin,
         "inverted": inverted_cutoffs}


@pytest.fixture(scope="module")
def eng():
    e = S.Engine(0)
    yield e
    e.close()


def test_selftest_wave_helpers_and_normals(eng):
    bad, nrm = eng.selftest()
    assert not bad.any(), bad
    ref = np.array([O.normals(0x9ABCDEF012345678, b, 1000 + b) for b in range(64)])
    assert np.abs(nrm - ref).max() < 2e-5


@pytest.mark.parametrize("case", list(CASES))
def test_lists_bit_exact(eng, case):
    s = CASES[case]()
    sid = eng.add(s)
    try:
        want, _ = oracle_sets(Oracle(s), s, s.xyz)
        for which, ref in want.items():
            got = eng.get_list(sid, which)
            assert len(got) == len(ref), (which, len(got), len(ref))
            assert pair_set(got) == ref, which
    finally:
        eng.remove(sid)


def force_check(f, fr):
    scale = np.linalg.norm(fr, axis=1).max()
    err = np.linalg.norm(f - fr, axis=1).max() / scale
    assert err <= 1e-5, err
    return err


@pytest.mark.parametrize("case", list(CASES))
def test_forces_energies_observables(eng, case):
    s = CASES[case]()
    o = Oracle(s)
    sid = eng.add(s)
    try:
        _, pl = oracle_sets(o, s, s.xyz)
        fr, er, st = o.forces(s.xyz, pl)
        assert st == 0
        force_check(eng.force_eval(sid), fr)
        e = eng.energy_eval(sid)
        scale = np.abs(er[:4]).sum()
        assert np.all(np.abs(e[:5] - er[:5]) <= 2e-6 * scale), (e, er)
        assert abs(e[S.E_VIRIAL] - er[5]) <= 2e-6 * max(abs(er[5]), scale)
        q, c, rg = eng.observables(sid)
        qr, cr, rgr = o.observables(s.xyz)
        assert c == cr and q == pytest.approx(qr, abs=0)
        assert rg == pytest.approx(rgr, rel=1e-9)
    finally:
        eng.remove(sid)


def test_random_force_is_var_times_philox_normals(eng):
    s = c1()
    s.step0 = 12345
    sid = eng.add(s)
    try:
        f0 = eng.force_eval(sid, with_noise=False)
        f1 = eng.force_eval(sid, with_noise=True)
        a5, var = S.params_derived(s.params)
        ref = np.array([O.normals(s.params.seed, i, 12345) for i in range(s.n)]) * np.float64(var)
        assert np.abs((f1 - f0) - ref).max() < 1e-5 * var * 4
    finally:
        eng.remove(sid)


@pytest.mark.parametrize("case,steps", [("c1", 1000), ("multi", 1000), ("refbox", 300), ("c2", 1000),
                                        ("tight", 200), ("inverted", 200)])
def test_trajectory_matches_oracle(eng, case, steps):
    s = CASES[case]()
    sid = eng.add(s)
    try:
        eng.run([sid], steps)
        assert eng.step_of(sid) == steps
        x = eng.coords(sid)
        a5, var = S.params_derived(s.params)
        xr = Oracle(s).run(s.xyz, 0, steps, s.params.seed, a5, var, s.params.n_nl, s.params.n_pl)
        rmsd = float(np.sqrt(((x - xr) ** 2).sum(1).mean()))
        assert rmsd <= 1e-4, rmsd
    finally:
        eng.remove(sid)


def test_golden_c1_forces_and_100_steps(eng):
    z = np.load(os.path.join(GOLD, "c1_expected.npz"))
    s = c1()
    sid = eng.add(s)
    try:
        assert pair_set(eng.get_list(sid, S.LIST_PL_REP)) == pair_set(z["pl_rep"])
        force_check(eng.force_eval(sid), z["forces"])
        eng.run([sid], 100)
        x = eng.coords(sid)
        assert np.sqrt(((x - z["x100"]) ** 2).sum(1).mean()) <= 1e-4
    finally:
        eng.remove(sid)


def test_explicit_list_updates_after_moving_beads(eng):
    s = multi_chain()
    sid = eng.add(s)
    try:
        rng = np.random.default_rng(5)
        x = s.xyz + rng.normal(0, 0.25, s.xyz.shape).astype(np.float32)
        eng.set_coords(sid, x)
        eng.update_neighbor_list(sid)
        eng.update_pair_list(sid)
        want, pl = oracle_sets(Oracle(s), s, x)
        for which, ref in want.items():
            assert pair_set(eng.get_list(sid, which)) == ref, which
        fr, _, _ = Oracle(s).forces(x, pl)
        force_check(eng.force_eval(sid), fr)
    finally:
        eng.remove(sid)


def batch_systems():
    out = []
    for k, n in enumerate((200, 777, 1500, 64, 3000)):
        s = S.make_chain_system(100 + k, n)
        s.params.seed = 1000 + k
        s.params.T = 0.5 + 0.1 * k
        out.append(s)
    return out


def test_batched_alone_plain_graph_bitwise():
    steps = 137  # crosses graph blocks and a partial block
    systems = batch_systems()
    with S.Engine(0) as a:
        ids = [a.add(s) for s in systems]
        a.run(ids, steps)
        batched = [a.coords(i) for i in ids]
    with S.Engine(0) as b:
        for s, ref in zip(systems, batched):
            i = b.add(s)
            b.run([i], steps)
            assert np.array_equal(b.coords(i), ref)
            b.remove(i)
        ids = [b.add(s) for s in systems]
        b.run_plain(ids, steps)
        for i, ref in zip(ids, batched):
            assert np.array_equal(b.coords(i), ref)
    with S.Engine(0, use_graphs=False) as c:
        ids = [c.add(s) for s in systems]
        c.run(ids, 60)
        c.run(ids, steps - 60)
        for i, ref in zip(ids, batched):
            assert np.array_equal(c.coords(i), ref)


@pytest.mark.parametrize("lpb", [4, 8, 32])
def test_lanes_per_bead_variants(lpb):
    s = multi_chain()
    o = Oracle(s)
    _, pl = oracle_sets(o, s, s.xyz)
    fr, _, _ = o.forces(s.xyz, pl)
    with S.Engine(0, lanes_per_bead=lpb) as e:
        sid = e.add(s)
        force_check(e.force_eval(sid), fr)
        e.run([sid], 50)
        a5, var = S.params_derived(s.params)
        xr = o.run(s.xyz, 0, 50, s.params.seed, a5, var, s.params.n_nl, s.params.n_pl)
        assert np.sqrt(((e.coords(sid) - xr) ** 2).sum(1).mean()) <= 1e-4


def test_pair_list_overflow_is_reported_not_written():
    with S.Engine(0, pl_cap_max=16) as e:
        with pytest.raises(S.SopError) as ei:
            e.add(c1())
        assert ei.value.status == S._abi.SOP_ERR_CAPACITY


def test_overstretched_bond_is_reported():
    s = c1()
    x = s.xyz.copy()
    u = (x[64] - x[63]) / np.linalg.norm(x[64] - x[63])
    x[64:] += (3.0 * u).astype(np.float32)  # bond 63-64 stretched by 3 A > R0
    s.xyz = x
    with S.Engine(0) as e:
        sid = e.add(s)
        with pytest.raises(S.SopError) as ei:
            e.run([sid], 1)
        assert ei.value.status == S._abi.SOP_ERR_NUMERIC


def test_extent_near_half_box_is_refused():
    s = c1()
    s.params.boxl = 40.0
    with S.Engine(0) as e:
        with pytest.raises(S.SopError) as ei:
            e.add(s)
        assert ei.value.status == S._abi.SOP_ERR_PBC


def test_timing_and_autotune_entry_points():
    systems = batch_systems()[:3]
    with S.Engine(0) as e:
        ids = [e.add(s) for s in systems]
        ms = e.time_step_kernel(ids, 5)
        assert ms > 0
        lanes, grid = e.autotune(ids)
        assert lanes in (4, 8, 16) and grid > 0
        e.run(ids, 20)


def test_autotune_leaves_no_trace_and_settings_transfer():
    # the trials run from a device snapshot: coordinates, lists and counters
    # are as before; the chosen settings, applied with sop_tune_set at the
    # same point of the same run on a fresh engine, give the same trajectories
    # bitwise
    systems = batch_systems()
    with S.Engine(0) as a:
        ids = [a.add(s) for s in systems]
        a.run(ids, 30)
        before = [a.coords(i) for i in ids]
        tuned = a.autotune_ex(ids)
        print("tuned:", tuned, flush=True)
        assert tuned["trials"] > 5 and tuned["block_ms_best"] <= tuned["block_ms_start"] * 1.0001
        assert all(np.array_equal(a.coords(i), x) for i, x in zip(ids, before))
        assert all(a.step_of(i) == 30 for i in ids)
        a.run(ids, 70)
        after = [a.coords(i) for i in ids]
    with S.Engine(0) as b:
        ids = [b.add(s) for s in systems]
        b.run(ids, 30)
        b.tune_set(tuned)
        b.run(ids, 70)
        for i, x in zip(ids, after):
            assert np.array_equal(b.coords(i), x)


@pytest.mark.parametrize("n", [1, 2, 3, 4])
def test_tiny_chains_match_oracle(eng, n):
    # one bead (no terms at all), one bond, one i,i+2 term, one possible native:
    # tiles padded far past the chain, rows of padding only
    s = S.make_chain_system(70 + n, n)
    sid = eng.add(s)
    try:
        o = Oracle(s)
        _, lists = oracle_sets(o, s, s.xyz)
        fr, _, st = o.forces(s.xyz, lists)
        assert st == 0
        f = eng.force_eval(sid)
        assert np.abs(f - fr).max() <= 1e-5 * max(1.0, np.abs(fr).max())
        eng.run([sid], 50)
        a5, var = S.params_derived(s.params)
        xr = o.run(s.xyz, 0, 50, s.params.seed, a5, var, s.params.n_nl, s.params.n_pl)
        x = eng.coords(sid)
        assert np.sqrt(((x - xr) ** 2).sum(1).mean()) <= 1e-4
    finally:
        eng.remove(sid)


def test_malformed_systems_are_rejected(eng):
    def chain10(**kw):
        s = S.make_chain_system(80, 10)
        for k, v in kw.items():
            setattr(s, k, v)
        return s

    bad = {
        "no beads": S.System(xyz=np.zeros((0, 3), np.float32), params=S.params_config()),
        "native inside |i-j| <= 2": chain10(nat_i=np.array([0], np.int32), nat_j=np.array([2], np.int32),
                                            nat_r0=np.array([5.0], np.float32)),
        "duplicate native": chain10(nat_i=np.array([0, 0], np.int32), nat_j=np.array([5, 5], np.int32),
                                    nat_r0=np.array([5.0, 5.0], np.float32)),
        "native j out of range": chain10(nat_i=np.array([0], np.int32), nat_j=np.array([10], np.int32),
                                         nat_r0=np.array([5.0], np.float32)),
        "bead type 3": chain10(type=np.full(10, 3, np.int32)),
        "decreasing chain ids": chain10(chain=np.array([1] * 5 + [0] * 5, np.int32)),
    }
    for what, s in bad.items():
        with pytest.raises(S.SopError) as ei:
            eng.add(s)
        assert ei.value.status == S._abi.SOP_ERR_INVALID, what
    # more native partners in one 64-bead tile than the step kernel's 16-bit
    # native offsets reach: each of beads 0..63 paired with the 1,030 beads
    # 67..1,096 (65,920 natives in tile 0), refused as a capacity error
    big = S.make_chain_system(82, 1100)
    ii, jj = np.meshgrid(np.arange(64, dtype=np.int32), np.arange(67, 1097, dtype=np.int32), indexing="ij")
    big.nat_i, big.nat_j = ii.ravel(), jj.ravel()
    big.nat_r0 = np.full(big.nat_i.size, 6.0, np.float32)
    with pytest.raises(S.SopError) as ei:
        eng.add(big)
    assert ei.value.status == S._abi.SOP_ERR_CAPACITY
    s = S.make_chain_system(80, 10)
    s.params.n_nl = 0
    with pytest.raises(S.SopError) as ei:
        eng.add(s)
    assert ei.value.status == S._abi.SOP_ERR_INVALID
    # the engine is still usable
    sid = eng.add(S.make_chain_system(81, 10))
    try:
        eng.run([sid], 5)
        assert eng.step_of(sid) == 5
    finally:
        eng.remove(sid)
