"""The large multi-chain assembly on device 0 at full size: 1,000 chains of
500 beads (500,000 beads) on a 10x10x10 lattice at 60 A, open boundaries,
intra-chain natives only, repulsion inside and between chains.

The oracle's all-pairs neighbor list is out of reach at this size; its
grid-built list (the same arrays, pinned to the all-pairs one on the CPU)
gives the lists, the forces and a 20-step trajectory here.  The assembly must
also step bitwise the same alone and batched beside smaller systems.
"""
import numpy as np
import pytest

import sop_amd as S
from sop_amd import workloads as W
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu


def keys(pairs, n):
    a = np.asarray(pairs, np.int64).reshape(-1, 2)
    return np.sort(np.minimum(a[:, 0], a[:, 1]) * n + np.maximum(a[:, 0], a[:, 1]))


@pytest.fixture(scope="module")
def assembly():
    print("assembly: 1,000 chains", flush=True)
    s = W.c5()
    print("assembly: oracle lists", flush=True)
    return s, Oracle(s).lists(s.xyz, grid=True)


def test_assembly_lists_forces_energies(assembly):
    s, ((nla, nlr), (pla, plr)) = assembly
    nat = np.stack([s.nat_i, s.nat_j], 1)
    # inter-chain pairs are present: the chains do interact
    assert (s.chain[plr[:, 0]] != s.chain[plr[:, 1]]).sum() > 0
    with S.Engine(0, nl_cap_max=1024) as e:
        sid = e.add(s)
        for which, ref in ((S.LIST_PL_NATIVE, nat[pla]), (S.LIST_PL_REP, plr), (S.LIST_NL_NATIVE, nat[nla]),
                           (S.LIST_NL_REP, nlr)):
            print(f"assembly: list {which}", flush=True)
            got = e.get_list(sid, which)
            assert len(got) == len(ref), (which, len(got), len(ref))
            assert np.array_equal(keys(got, s.n), keys(ref, s.n)), which
        print("assembly: forces", flush=True)
        fr, er, st = Oracle(s).forces(s.xyz, (pla, plr))
        assert st == 0
        f = e.force_eval(sid)
        err = np.linalg.norm(f - fr, axis=1).max() / np.linalg.norm(fr, axis=1).max()
        assert err <= 1e-5, err
        en = e.energy_eval(sid)
        scale = np.abs(er[:4]).sum()
        assert np.all(np.abs(en[:5] - er[:5]) <= 2e-6 * scale), (en, er)


def test_assembly_trajectory_alone_and_in_a_batch(assembly):
    s, _ = assembly
    a5, var = S.params_derived(s.params)
    small = [S.make_chain_system(60 + k, n) for k, n in enumerate((300, 2000, 7000))]
    with S.Engine(0) as a, S.Engine(0) as b:
        ia = a.add(s)
        ib = b.add(s)
        others = [b.add(t) for t in small]
        print("assembly: 20 steps alone", flush=True)
        a.run([ia], 20)
        print("assembly: 20 steps in a batch", flush=True)
        b.run([ib] + others, 20)
        x = a.coords(ia)
        assert np.array_equal(x, b.coords(ib))
    print("assembly: oracle 20 steps", flush=True)
    xr = Oracle(s).run(s.xyz, 0, 20, s.params.seed, a5, var, s.params.n_nl, s.params.n_pl, grid=True)
    rmsd = float(np.sqrt(((x - xr) ** 2).sum(1).mean()))
    assert rmsd <= 1e-4, rmsd
