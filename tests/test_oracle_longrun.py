"""Pins of the long-run comparison procedure (tests/longrun_stats.py) on the
oracle alone: two oracle sets with disjoint Philox keys agree; a set run 10 %
hotter (the noise amplitude var = sqrt(2 T zeta / h) off by 5 %) is rejected.
The device test (test_gpu_longrun.py) applies the same procedure to the HIP
path against the oracle."""
import longrun_stats as L
from sop_amd import workloads as W

R, NSTEPS, NUP, BURN = 16, 20000, 100, 40


def test_procedure_accepts_same_dynamics_and_rejects_wrong_temperature():
    s = W.c1()
    ka = [W.philox_key(101, r) for r in range(R)]
    kb = [W.philox_key(102, r) for r in range(R)]
    a = L.oracle_set(s, ka, NSTEPS, NUP)
    b = L.oracle_set(s, kb, NSTEPS, NUP)
    hot = L.oracle_set(s, kb, NSTEPS, NUP, T=1.1 * s.params.T)
    same = L.compare(a, b, BURN)
    assert L.agree(same), same
    wrong = L.compare(a, hot, BURN)
    assert not L.agree(wrong), wrong
    # the energy is the sharpest probe of the temperature
    assert abs(wrong["E_total"]["z"]) > 3 and wrong["E_total"]["p"] < 0.01, wrong
