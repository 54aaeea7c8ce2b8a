"""Pins of the underdamped integrator (sop.cu:421-484, param.cpp:104-108) in the
oracle and the coefficient routine of the library (CPU only):

  coefficients  zeta = 0 reduces a1..a4 to velocity Verlet (h, h^2/2, 1/h, h/2);
                the general values follow the reference's float/double order
  zeta = 0      the scheme is velocity Verlet: total energy is conserved and a
                run with reversed final velocities returns to its start
  free beads    the velocity recursion v' = A v + B xi + C xi' (A = a3 a1,
                B = a3 a2 var, C = a4 var) has stationary variance
                (B^2 + C^2 + 2ABC) / (1 - A^2) and lag-1 covariance A V + B C
  T < 1e-5      velocities are never updated (steepest-descent branch)
"""
import numpy as np
import pytest

import sop_amd as S
from oracle.oracle import Oracle


def ud_params(h=0.01, zeta=0.5, T=0.6, nl=1000.0, pl=1000.0):
    p = S.params_config()
    p.dynamics = S._abi.DYN_UNDERDAMPED
    p.h, p.zeta, p.T = h, zeta, T
    p.nl_abs, p.pl_abs = nl, pl
    p.n_nl, p.n_pl = 10**9, 10**9
    return p


def test_coefficients_reduce_to_velocity_verlet_at_zero_friction():
    for h in (2.5e-3, 0.01, 0.05):
        p = ud_params(h=h, zeta=0.0)
        a = S.params_underdamped(p)
        hf = np.float32(h)
        want = np.array([hf, hf * (hf / np.float32(2)), np.float32(1.0 / np.float64(hf)), hf / np.float32(2)],
                        np.float32)
        assert np.array_equal(a, want), (h, a, want)


def test_coefficients_follow_reference_rounding():
    # param.cpp:104-108 with FLOAT = float: float*float products round to
    # float, the double literals promote the rest; a5 = h/zeta in float.
    p = ud_params(h=2.5e-3, zeta=0.05)
    a = S.params_underdamped(p)
    h, z = np.float32(p.h), np.float32(p.zeta)
    halfh = np.float32(h / np.float32(2))
    hz = np.float32(h * z)
    want = np.array([
        np.float32(np.float64(h) * (1.0 - np.float64(np.float32(z * halfh)))),
        np.float32(h * halfh),
        np.float32((1.0 - np.float64(hz) / 2.0 + np.float64(np.float32(hz * hz)) / 4.0) / np.float64(h)),
        np.float32(np.float64(halfh) * (1.0 - np.float64(hz) / 2.0)),
    ], np.float32)
    assert np.array_equal(a, want)
    # and they are the textbook values to float precision
    hd, zd = float(h), float(z)
    ref = [hd * (1 - zd * hd / 2), hd * hd / 2, (1 - hd * zd / 2 + (hd * zd) ** 2 / 4) / hd, hd / 2 * (1 - hd * zd / 2)]
    assert np.allclose(a, ref, rtol=2e-7, atol=0)


def total_energy(o, x, v):
    pl = o.lists(np.asarray(x, np.float32))[1]
    e = o.forces(np.asarray(x, np.float64), pl)[1]
    return e[4] + 0.5 * (np.asarray(v, np.float64) ** 2).sum()


def test_zero_friction_conserves_energy():
    s = S.make_chain_system(21, 60)
    s.params = ud_params(h=0.005, zeta=0.0, T=1.0)
    o = Oracle(s)
    a = S.params_underdamped(s.params)
    _, var = S.params_derived(s.params)
    assert var == 0.0
    rng = np.random.default_rng(5)
    v0 = rng.normal(0, 0.5, (s.n, 3)).astype(np.float32)
    x, v = s.xyz, v0
    e0 = total_energy(o, s.xyz, v0)
    es = []
    for seg in range(10):
        x, v = o.run_ud(np.float32(x), 200 * seg, 200, 3, a, var, 1, 10**9, 10**9, v0=np.float32(v))
        es.append(total_energy(o, x, v))
    es = np.array(es)
    ke0 = 0.5 * (v0.astype(np.float64) ** 2).sum()
    # velocity Verlet: bounded O(h^2) energy error, no drift
    assert np.abs(es - e0).max() < 2e-3 * ke0, (e0, es)
    # with friction the same run loses energy, so the check above can fail
    s.params.zeta = 2.0
    a2 = S.params_underdamped(s.params)
    xf, vf = o.run_ud(s.xyz, 0, 2000, 3, a2, 0.0, 1, 10**9, 10**9, v0=v0)
    assert total_energy(o, xf, vf) < e0 - 0.2 * ke0


def test_zero_friction_is_time_reversible():
    s = S.make_chain_system(22, 40)
    s.params = ud_params(h=0.005, zeta=0.0, T=1.0)
    o = Oracle(s)
    a = S.params_underdamped(s.params)
    rng = np.random.default_rng(6)
    v0 = rng.normal(0, 0.5, (s.n, 3)).astype(np.float32)
    x1, v1 = o.run_ud(s.xyz, 0, 200, 3, a, 0.0, 1, 10**9, 10**9, v0=v0)
    # the float32 hand-over of x1, v1 bounds what comes back
    x2, v2 = o.run_ud(np.float32(x1), 0, 200, 3, a, 0.0, 1, 10**9, 10**9, v0=-np.float32(v1))
    assert np.abs(x2 - s.xyz.astype(np.float64)).max() < 1e-4
    assert np.abs(v2 + v0.astype(np.float64)).max() < 1e-3
    # a one-step offset in the velocity hand-over would not come back
    assert np.abs(x1 - s.xyz).max() > 0.1


def free_beads(n, spacing=40.0):
    k = int(np.ceil(n ** (1 / 3)))
    g = np.stack(np.meshgrid(np.arange(k), np.arange(k), np.arange(k), indexing="ij"), -1).reshape(-1, 3)[:n]
    p = ud_params(h=0.02, zeta=5.0, T=0.8, nl=32.0, pl=16.0)
    return S.System(xyz=spacing * g.astype(np.float32), params=p, chain=np.arange(n, dtype=np.int32))


def stationary(a, var):
    a1, a2, a3, a4 = (float(t) for t in a)
    A, B, C = a3 * a1, a3 * a2 * float(var), a4 * float(var)
    V = (B * B + C * C + 2 * A * B * C) / (1 - A * A)
    return V, (A * V + B * C) / V


def test_free_bead_velocity_variance_and_correlation():
    s = free_beads(1500)
    o = Oracle(s)
    a = S.params_underdamped(s.params)
    _, var = S.params_derived(s.params)
    V, rho = stationary(a, var)
    assert V == pytest.approx(s.params.T, rel=0.15)  # near equipartition at h zeta = 0.1
    x, v = s.xyz, None
    step = 0
    samples, pairs = [], []
    x, v = o.run_ud(x, step, 60, 9, a, var, 1, 50, 10)  # burn-in, 6 relaxation times
    step += 60
    for k in range(16):
        x, v = o.run_ud(np.float32(x), step, 29, 9, a, var, 1, 50, 10, v0=np.float32(v))
        step += 29
        x, v1 = o.run_ud(np.float32(x), step, 1, 9, a, var, 1, 50, 10, v0=np.float32(v))
        step += 1
        samples.append(v1)
        pairs.append((v, v1))
        v = v1
    vs = np.concatenate(samples).ravel()
    n = vs.size
    assert vs.var() == pytest.approx(V, rel=5 * np.sqrt(2.0 / n))
    c = np.mean([(p0 * p1).mean() for p0, p1 in pairs]) / V
    assert c == pytest.approx(rho, abs=5 / np.sqrt(n))


def test_frozen_velocities_below_temperature_threshold():
    s = S.make_chain_system(23, 30)
    s.params = ud_params(h=0.01, zeta=0.5, T=0.0)
    o = Oracle(s)
    a = S.params_underdamped(s.params)
    v0 = np.full((s.n, 3), 0.25, np.float32)
    x1, v1 = o.run_ud(s.xyz, 0, 50, 3, a, 0.0, 0, 10**9, 10**9, v0=v0)
    assert np.array_equal(v1, v0.astype(np.float64))
    x1u, v1u = o.run_ud(s.xyz, 0, 50, 3, a, 0.0, 1, 10**9, 10**9, v0=v0)
    assert not np.allclose(v1u, v0)
