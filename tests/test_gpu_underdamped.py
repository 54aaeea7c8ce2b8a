"""Underdamped dynamics (sop.cu:421-484) on device 0 against the oracle.

  trajectory   RMSD <= 1e-4 A after 1,000 steps (reference h = 2.5e-3, zeta =
               0.05) from the identical Philox stream; velocities after the
               run within 1e-3 of the largest oracle velocity component
  free beads   velocity variance and lag-1 correlation of 20,000 free beads
               match the closed form of the discrete recursion
  increments   the last step's increment equals x(k) - x(k-1) to float rounding
  batching     overdamped and underdamped systems in one batch: coordinates and
               velocities bitwise identical to each run alone
"""
import numpy as np
import pytest

import sop_amd as S
from oracle.oracle import Oracle

from test_oracle_underdamped import free_beads, stationary

pytestmark = pytest.mark.gpu


def ud_chain(seed, n, h=2.5e-3, zeta=0.05, T=0.6):
    s = S.make_chain_system(seed, n)
    s.params.dynamics = S._abi.DYN_UNDERDAMPED
    s.params.h, s.params.zeta, s.params.T = h, zeta, T
    return s


@pytest.mark.parametrize("n,steps", [(128, 1000), (2000, 1000)])
def test_underdamped_trajectory_matches_oracle(n, steps):
    s = ud_chain(40 + n, n)
    rng = np.random.default_rng(n)
    s.vel = rng.normal(0, 0.7, (n, 3)).astype(np.float32)
    a = S.params_underdamped(s.params)
    _, var = S.params_derived(s.params)
    with S.Engine(0) as e:
        i = e.add(s)
        e.run([i], steps)
        x, v = e.coords(i), e.velocities(i)
    xr, vr = Oracle(s).run_ud(s.xyz, 0, steps, s.params.seed, a, var, 1, s.params.n_nl, s.params.n_pl, v0=s.vel)
    rmsd = float(np.sqrt(((x - xr) ** 2).sum(1).mean()))
    assert rmsd <= 1e-4, rmsd
    assert np.abs(v - vr).max() <= 1e-3 * np.abs(vr).max()


def test_free_bead_velocity_statistics():
    s = free_beads(20000, spacing=40.0)
    a = S.params_underdamped(s.params)
    _, var = S.params_derived(s.params)
    V, rho = stationary(a, var)
    with S.Engine(0) as e:
        i = e.add(s)
        e.run([i], 60)
        vs, cs = [], []
        for _ in range(4):
            e.run([i], 19)
            v0 = e.velocities(i).astype(np.float64)
            e.run([i], 1)
            v1 = e.velocities(i).astype(np.float64)
            vs.append(v1)
            cs.append((v0 * v1).mean())
    vs = np.concatenate(vs).ravel()
    assert vs.var() == pytest.approx(V, rel=5 * np.sqrt(2.0 / vs.size))
    assert np.mean(cs) / V == pytest.approx(rho, abs=5 / np.sqrt(vs.size))


@pytest.mark.parametrize("dyn", [S._abi.DYN_OVERDAMPED, S._abi.DYN_UNDERDAMPED])
def test_increments_are_the_last_displacement(dyn):
    s = S.make_chain_system(51, 500)
    s.params.dynamics = dyn
    if dyn == S._abi.DYN_UNDERDAMPED:
        s.params.h, s.params.zeta = 0.01, 0.5
    with S.Engine(0) as e:
        i = e.add(s)
        e.run([i], 23)
        x0 = e.coords(i).astype(np.float64)
        e.run([i], 1)
        x1 = e.coords(i).astype(np.float64)
        d = e.increments(i).astype(np.float64)
    ulp = np.spacing(np.abs(x1).astype(np.float32)).astype(np.float64)
    assert np.all(np.abs((x1 - x0) - d) <= 2 * ulp)
    assert np.abs(d).max() > 0


def test_mixed_dynamics_batch_bitwise():
    systems = []
    for k, n in enumerate((300, 1200, 90, 2500)):
        s = S.make_chain_system(60 + k, n)
        s.params.seed = 500 + k
        if k % 2:
            s.params.dynamics = S._abi.DYN_UNDERDAMPED
            s.params.h, s.params.zeta = 0.01, 0.5 + k
        systems.append(s)
    steps = 77
    with S.Engine(0) as a:
        ids = [a.add(s) for s in systems]
        a.run(ids, steps)
        ref = [(a.coords(i), a.velocities(i)) for i in ids]
    with S.Engine(0) as b:
        for s, (x, v) in zip(systems, ref):
            i = b.add(s)
            b.run([i], steps)
            assert np.array_equal(b.coords(i), x)
            assert np.array_equal(b.velocities(i), v)
            b.remove(i)
