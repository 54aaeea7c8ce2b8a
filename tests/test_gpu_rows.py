"""Row capacities and large batches on device 0.

- moving a system's shell rows and pair-list tiles to another capacity
  (sop_system_reserve_rows, and the growth after a run whose rows came within
  an eighth of the capacity) changes no result: coordinates bitwise equal to
  an engine that never moved them, the same lists;
- a system whose first lists overflow the starting capacity gets a fitted one
  at add: lists bit-exact against the oracle, one grid build counted;
- a batch of more than 2,048 systems rebuilding from the cell grid together
  gives each system the result it has alone.
"""
import numpy as np
import pytest

import sop_amd as S
from sop_amd import workloads as W
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu


def pair_set(a):
    a = np.asarray(a, np.int64).reshape(-1, 2)
    return set(map(tuple, a.tolist()))


def system(seed, n):
    s = S.make_chain_system(seed, n)
    s.params.seed = W.philox_key(21, seed)
    s.params.T = 0.7
    return s


def test_moving_rows_changes_no_result():
    s = system(801, 3000)
    with S.Engine(0) as ref:
        r = ref.add(s)
        ref.run([r], 160)
        want = ref.coords(r)
        want_pl = pair_set(ref.get_list(r, S.LIST_PL_REP))
    with S.Engine(0) as e:
        i = e.add(s)
        cap0, longest = e.rows(i)
        print("rows at add:", cap0, longest, flush=True)
        assert longest <= cap0
        e.run([i], 70)
        e.reserve_rows(i, cap0 + 128)  # larger: every row moves
        assert e.rows(i)[0] == cap0 + 128
        e.run([i], 80)
        # a capacity the longest row is within an eighth of (but still below
        # it by a margin): the next run ends by growing it
        _, longest = e.rows(i)
        tight = int(longest / 0.875 - 1) // 16 * 16
        assert tight >= longest + 8, (tight, longest)
        e.reserve_rows(i, tight)
        assert e.rows(i)[0] == tight
        e.run([i], 10)
        grown, longest2 = e.rows(i)
        print("rows:", tight, "->", grown, "longest", longest2, flush=True)
        assert grown > tight
        assert np.array_equal(e.coords(i), want)
        assert pair_set(e.get_list(i, S.LIST_PL_REP)) == want_pl


def test_first_lists_longer_than_the_starting_capacity():
    s = system(802, 1200)
    s.params.nl_abs, s.params.pl_abs = 32.0, 24.0  # first rows longer than the starting 384 entries
    with S.Engine(0) as e:
        i = e.add(s)
        cap, longest = e.rows(i)
        print("rows:", cap, longest, flush=True)
        assert longest > 384 and cap > longest
        st = e.list_stats(i)
        assert st.grid_builds == 1 and st.shell_builds == 0, (st.grid_builds, st.shell_builds)
        (nla, nlr), (pla, plr) = Oracle(s).lists(s.xyz)
        nat = np.stack([s.nat_i, s.nat_j], 1)
        assert pair_set(e.get_list(i, S.LIST_PL_REP)) == pair_set(plr)
        assert pair_set(e.get_list(i, S.LIST_PL_NATIVE)) == pair_set(nat[pla])


def test_batch_of_more_than_2048_systems():
    # no shell rows: every pair-list rebuild of every system comes from the cell grid
    systems = [system(3000 + k, 64 + (k % 5) * 20) for k in range(2100)]
    for s in systems:
        s.params.n_pl = 5
    with S.Engine(0, pl_shell=-1.0) as e:
        ids = [e.add(s) for s in systems]
        e.run(ids, 12)
        got = {k: e.coords(ids[k]) for k in (0, 1000, 2047, 2048, 2099)}
        st = e.list_stats(ids[2099])
        assert st.grid_builds >= 3, st.grid_builds
    with S.Engine(0, pl_shell=-1.0) as e:
        for k, x in got.items():
            i = e.add(systems[k])
            e.run([i], 12)
            assert np.array_equal(e.coords(i), x), k
            e.remove(i)
