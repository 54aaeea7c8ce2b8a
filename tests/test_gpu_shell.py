"""Pair-list shell rows on device 0.

The cell grid writes shell rows (pairs within rc_pl + shell of the snapshot
pos_shell); later pair-list rebuilds filter those rows while every bead has
moved less than shell/2.  Either way the pair list must be the oracle's, bit
for bit (identical sorted sets of i<j pairs), on the same coordinates.
"""
import numpy as np
import pytest

import sop_amd as S
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu


def pair_set(a):
    return set(map(tuple, np.asarray(a, np.int64).reshape(-1, 2).tolist()))


def oracle_pl(s, x):
    (nla, nlr), (pla, plr) = Oracle(s).lists(np.float32(x))
    nat = np.stack([s.nat_i, s.nat_j], 1)
    return pair_set(nat[pla]), pair_set(plr)


def check_pl(eng, sid, s, x):
    nat, rep = oracle_pl(s, x)
    assert pair_set(eng.get_list(sid, S.LIST_PL_REP)) == rep
    assert pair_set(eng.get_list(sid, S.LIST_PL_NATIVE)) == nat


def test_shell_rebuilds_give_the_oracle_pair_list():
    s = S.make_chain_system(7, 1500)
    with S.Engine(0) as a, S.Engine(0, pl_shell=-1) as b:
        ia, ib = a.add(s), b.add(s)
        for k in range(4):
            a.run([ia], 10)
            x = a.coords(ia)
            print(f"shell round {k}", flush=True)
            a.update_pair_list(ia)   # forced rebuild at x: filtered from the shell rows
            check_pl(a, ia, s, x)
            b.set_coords(ib, x)
            b.update_pair_list(ib)   # no shell: straight from the cell grid
            check_pl(b, ib, s, x)
        sa, sb = a.list_stats(ia), b.list_stats(ib)
        assert sa.shell_builds >= 4 and sa.grid_builds >= 1, (sa.shell_builds, sa.grid_builds)
        assert sb.shell_builds == 0


def test_large_moves_rebuild_the_shell_rows():
    s = S.make_chain_system(8, 600)
    with S.Engine(0) as a:
        ia = a.add(s)
        st0 = a.list_stats(ia)
        rng = np.random.default_rng(2)
        x = s.xyz.copy()
        x[::7] += rng.normal(0.0, 1.0, x[::7].shape).astype(np.float32)  # some beads move > shell/2
        a.set_coords(ia, x)
        a.update_pair_list(ia)
        check_pl(a, ia, s, x)
        st1 = a.list_stats(ia)
        assert st1.grid_builds == st0.grid_builds + 1 and st1.shell_builds == st0.shell_builds
        # a small move keeps the shell rows
        y = x + rng.normal(0.0, 0.05, x.shape).astype(np.float32)
        a.set_coords(ia, y)
        a.update_pair_list(ia)
        check_pl(a, ia, s, y)
        st2 = a.list_stats(ia)
        assert st2.shell_builds == st1.shell_builds + 1 and st2.grid_builds == st1.grid_builds


def test_trajectory_with_and_without_shell_agree():
    s = S.make_chain_system(9, 800)
    with S.Engine(0) as a, S.Engine(0, pl_shell=-1) as b:
        ia, ib = a.add(s), b.add(s)
        a.run([ia], 300)
        b.run([ib], 300)
        xa, xb = a.coords(ia), b.coords(ib)
        # same lists at every rebuild; only the summation order of the rows differs
        assert np.sqrt(((xa - xb) ** 2).sum(1).mean()) <= 1e-4
        assert a.list_stats(ia).shell_builds > 0


def test_batch_mixing_shell_grid_and_neighbor_list_builds():
    # one batch whose rebuilds mix the three kinds: shell-row filters, grid
    # builds (large moves) and neighbor-list builds (a 0.05 A skin); the
    # shell-built systems are filtered beside the cell-grid chain and the
    # others after it.  Each system ends bitwise as it does alone; the tight
    # one also matches the oracle's trajectory
    normal = S.make_chain_system(41, 1500)
    tight = S.make_chain_system(31, 400)
    tight.params.nl_abs, tight.params.pl_abs = 16.05, 16.0
    small = S.make_chain_system(42, 300)
    small.params.T = 0.9
    systems = [normal, tight, small]
    steps = 120
    with S.Engine(0) as e:
        ids = [e.add(s) for s in systems]
        e.run(ids, steps)
        got = [e.coords(i) for i in ids]
        st = e.list_stats(ids[1])
        assert st.nl_builds > 0, st.nl_builds
        assert e.list_stats(ids[0]).shell_builds > 0
    with S.Engine(0) as e:
        for s, x in zip(systems, got):
            i = e.add(s)
            e.run([i], steps)
            assert np.array_equal(e.coords(i), x)
            e.remove(i)
    a5, var = S.params_derived(tight.params)
    xr = Oracle(tight).run(tight.xyz, 0, steps, tight.params.seed, a5, var, tight.params.n_nl, tight.params.n_pl)
    assert np.sqrt(((got[1] - xr) ** 2).sum(1).mean()) <= 1e-4
