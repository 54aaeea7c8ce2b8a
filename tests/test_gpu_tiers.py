"""The step kernel's work-item tiers on device 0.

A system is stepped from coordinates staged in LDS at two blocks per compute
unit (up to 5,056 beads), staged in a whole compute unit's LDS (up to 10,112
beads), or gathered from global memory (larger, or with LDS staging turned
off).  The whole-CU and global-gather launches run on a second stream beside
the two-blocks-per-unit launch.  The
arithmetic is the same in every tier, so trajectories and pair lists must be
bitwise identical.
"""
import numpy as np
import pytest

import sop_amd as S

pytestmark = pytest.mark.gpu


def test_tiers_are_bitwise_identical():
    # one system per tier: LDS, whole-CU LDS, global gathers (the last two
    # run on the side stream beside the first)
    systems = [S.make_chain_system(31, 3000), S.make_chain_system(32, 7000), S.make_chain_system(33, 11000)]
    with S.Engine(0) as a, S.Engine(0, lds_beads_max=-1) as b:
        ia = [a.add(s) for s in systems]
        ib = [b.add(s) for s in systems]
        print("tiers: 120 steps", flush=True)
        a.run(ia, 120)
        b.run(ib, 120)
        for x, y in zip(ia, ib):
            assert np.array_equal(a.coords(x), b.coords(y))
            assert np.array_equal(a.get_list(x, S.LIST_PL_REP), b.get_list(y, S.LIST_PL_REP))
        fa = a.force_eval(ia[1])
        fb = b.force_eval(ib[1])
        assert np.array_equal(fa, fb)


def test_tier_boundaries_are_bitwise_identical():
    # sizes on either side of each tier's limit: 5,056 beads is the largest
    # system staged at two blocks per compute unit and 5,057 the smallest on
    # the whole-CU tier; 10,112 is the largest whole-CU system and 10,113
    # gathers from global memory.  All against global gathers throughout.
    sizes = (5056, 5057, 10112, 10113)
    systems = [S.make_chain_system(50 + k, n) for k, n in enumerate(sizes)]
    with S.Engine(0) as a, S.Engine(0, lds_beads_max=-1) as b:
        ia = [a.add(s) for s in systems]
        ib = [b.add(s) for s in systems]
        print("tier boundaries: 60 steps", flush=True)
        a.run(ia, 60)
        b.run(ib, 60)
        for n, x, y in zip(sizes, ia, ib):
            assert np.array_equal(a.coords(x), b.coords(y)), n
            assert np.array_equal(a.get_list(x, S.LIST_PL_REP), b.get_list(y, S.LIST_PL_REP)), n


def test_split_and_whole_tiles_are_bitwise_identical():
    # a small batch runs each 64-bead tile on two waves, half of the rounds
    # each; a batch of more tiles than 16 per compute unit (4,536 here) runs
    # one wave per tile.  Each bead is computed the same way either way.
    s = S.make_chain_system(41, 4000)
    with S.Engine(0) as a, S.Engine(0) as b:
        ia = a.add(s)
        ib = [b.add(s) for _ in range(72)]
        print("split vs whole tiles: 60 steps", flush=True)
        a.run([ia], 60)
        b.run(ib, 60)
        xa = a.coords(ia)
        for k in (0, 35, 71):
            assert np.array_equal(xa, b.coords(ib[k]))
        assert np.array_equal(a.force_eval(ia), b.force_eval(ib[5]))


def test_global_gather_batch_of_16_and_32_bit_rows():
    # systems above 65,536 beads keep 32-bit pair-list entries, smaller ones
    # 16-bit: in one global-gather batch the shell filter's waves stage rows
    # of both widths at once, each in its own part of the block's LDS
    systems = [S.make_chain_system(41, 70000)] + [S.make_chain_system(42 + k, 2000 + 700 * k) for k in range(3)]
    with S.Engine(0, lds_beads_max=-1) as e:
        ids = [e.add(s) for s in systems]
        print("mixed widths: 30 steps", flush=True)
        e.run(ids, 30)
        got = [e.coords(i) for i in ids]
    with S.Engine(0, lds_beads_max=-1) as e:
        for s, x in zip(systems, got):
            i = e.add(s)
            e.run([i], 30)
            assert np.array_equal(e.coords(i), x), s.n
            e.remove(i)
