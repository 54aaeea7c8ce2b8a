"""Long-run statistics of the HIP path against the FP64 oracle.

The 128-bead chain (T = 0.6, dt = 0.05, zeta = 50, lists 32 / 16 A every
50 / 10 steps) runs 50,000 steps in 32 replicas on the device, one batch with
distinct Philox keys, and 32 replicas in the FP64 oracle on the host with a
disjoint set of keys.  Every 100 steps both record the total energy, Rg^2 and
Q (sop.cu:394-396, 528-611).  After a burn-in of 10,000 steps the two sets of
distributions must agree (tests/longrun_stats.py):

  means  |mean_gpu - mean_oracle| <= 3 standard errors (block averages, one
         block per replica, so the error accounts for autocorrelation);
  KS     two-sample Kolmogorov-Smirnov p > 0.01 on samples spaced by at least
         2 tau_int + 1.

Disjoint keys make the sets independent: with the same keys the device and
oracle trajectories would coincide for their first thousands of steps and the
comparison would not test the distributions.  The same procedure rejects an
oracle run at T = 0.66 against one at T = 0.6 (test_oracle_longrun.py).
"""
import numpy as np
import pytest

import sop_amd as S
from sop_amd import workloads as W

import longrun_stats as L

pytestmark = pytest.mark.gpu

R, NSTEPS, NUP, BURN = 32, 50000, 100, 100


def gpu_set(system, keys):
    out = np.zeros((len(keys), NSTEPS // NUP, 4))
    with S.Engine(0) as e:
        ids = []
        for k in keys:
            s = S.System(xyz=system.xyz, params=S.copy_params(system.params), bond_r0=system.bond_r0,
                         nat_i=system.nat_i, nat_j=system.nat_j, nat_r0=system.nat_r0)
            s.params.seed = int(k)
            ids.append(e.add(s))
        for m in range(NSTEPS // NUP):
            e.run(ids, NUP)
            for r, i in enumerate(ids):
                q, c, rg2 = e.observables(i)
                out[r, m] = (e.energy_eval(i)[S.E_TOTAL], rg2, q, c)
    return out


def test_long_run_distributions_match_oracle():
    s = W.c1()
    ka = [W.philox_key(201, r) for r in range(R)]
    kb = [W.philox_key(202, r) for r in range(R)]
    a = gpu_set(s, ka)
    b = L.oracle_set(s, kb, NSTEPS, NUP, threads=16)
    res = L.compare(a, b, BURN)
    for name, r in res.items():
        print(name, {k: (round(v, 4) if isinstance(v, float) else v) for k, v in r.items()}, flush=True)
    for name, r in res.items():
        assert abs(r["z"]) <= r["zmax"], (name, r)
        assert r["p"] > r["alpha"], (name, r)
