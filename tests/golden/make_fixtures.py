"""Writes the committed fixtures of tests/golden/ (run from the repo root).

c1_system.npz   the 128-bead workload chain (generator seed 1, natives < 8 A)
c1_expected.npz oracle outputs on it: lists, FP64 forces/energies at the
                native state, and the FP64 coordinates after 100 steps
                (Philox key = params.seed = 1, benchmark parameters)

The reference ships no fixtures of its own (no tests, decks or outputs), so
these are this engine's regression vectors.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import sop_amd as S  # noqa: E402
from oracle.oracle import Oracle  # noqa: E402


def main():
    s = S.make_chain_system(1, 128)
    np.savez(os.path.join(HERE, "c1_system.npz"), xyz=s.xyz, bond_r0=s.bond_r0, nat_i=s.nat_i, nat_j=s.nat_j,
             nat_r0=s.nat_r0)
    o = Oracle(s)
    (nla, nlr), (pla, plr) = o.lists(s.xyz)
    f, e, st = o.forces(s.xyz, (pla, plr))
    a5, var = S.params_derived(s.params)
    x100 = o.run(s.xyz, 0, 100, s.params.seed, a5, var, s.params.n_nl, s.params.n_pl)
    np.savez(os.path.join(HERE, "c1_expected.npz"), nl_att=nla, nl_rep=nlr, pl_att=pla, pl_rep=plr, forces=f,
             energies=e, x100=x100)


if __name__ == "__main__":
    main()
