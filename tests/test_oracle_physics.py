"""Pins of the oracle's force field, list predicates, integrator and observables
(CPU only).  Each check is chosen so that a plausible mistake in the oracle (a
dropped term, a wrong sign or index, a transposed table, FMA in the predicate,
single instead of double rounding of the 3.2 factor) fails it:

  energies   closed-form values at special points of each term
  forces     F = -grad E by central differences (FP64); Newton's third law;
             translation / rotation invariance
  virial     W = -dE(lambda x)/dlambda at lambda = 1 (pair-potential identity)
  lists      constructed geometries with known membership (strict <), the
             double-then-float rounding of 3.2*r0 (sop.cu:746), minimum-image
             ties (rint half-even); the grid-built neighbor list (for
             systems too large for the all-pairs loop) equals the all-pairs
             one array for array
  integrator noise off: one step equals x + a5 F; free beads: <dr^2> = 6 (T/zeta) t
  Rg, Q      pair-sum Rg^2 equals the centroid form; Q at scaled native states
"""
import math

import numpy as np
import pytest

import sop_amd as S
from oracle import oracle as O
from oracle.oracle import Oracle


def mk(xyz, chain=None, types=None, bond_r0=None, nat=(), params=None):
    xyz = np.asarray(xyz, np.float32).reshape(-1, 3)
    n = len(xyz)
    nat = list(nat)
    p = S.copy_params(params if params is not None else S.params_config())
    return S.System(xyz=xyz, params=p, chain=chain, type=types, bond_r0=bond_r0,
                    nat_i=np.array([a for a, b, r in nat], np.int32),
                    nat_j=np.array([b for a, b, r in nat], np.int32),
                    nat_r0=np.array([r for a, b, r in nat], np.float32))


def energy(o, x, pl=None):
    if pl is None:
        pl = o.lists(np.asarray(x, np.float32))[1]
    return o.forces(np.asarray(x, np.float64), pl)[1]


# ------------------------------------------------------------ closed forms

def test_fene_closed_form():
    p = S.params_config()
    # two bonded beads, bond r0 = 3.8, separation 4.8: dev = 1
    s = mk([[0, 0, 0], [4.8, 0, 0]], bond_r0=np.array([3.8, 3.8], np.float32), params=p)
    o = Oracle(s)
    e = energy(o, s.xyz)
    dev = np.float64(np.float32(4.8)) - np.float64(np.float32(3.8))
    want = -0.5 * 20.0 * 4.0 * math.log(1.0 - dev * dev / 4.0)  # Kremer-Grest FENE
    assert e[S.E_BOND] == pytest.approx(want, rel=1e-12)
    s0 = mk([[0, 0, 0], [3.8, 0, 0]], bond_r0=np.array([3.8, 3.8], np.float32), params=p)
    f, e0, _ = Oracle(s0).forces(s0.xyz, Oracle(s0).lists(s0.xyz)[1])
    assert abs(e0[S.E_BOND]) < 1e-12 and np.abs(f).max() < 1e-5


def test_fene_overstretch_flagged():
    s = mk([[0, 0, 0], [5.9, 0, 0]], bond_r0=np.array([3.8, 3.8], np.float32))
    _, _, st = Oracle(s).forces(s.xyz, Oracle(s).lists(s.xyz)[1])
    assert st == -3


def test_soft_sphere_i_i2_at_sigma():
    p = S.params_config()  # sigma_ss 5.4, eps_ss 1
    # bend so that beads 0 and 2 sit exactly 5.4 apart
    x = np.array([[0, 0, 0], [2.7, 2.65, 0], [5.4, 0, 0]], np.float64)
    s = mk(x, bond_r0=np.array([np.linalg.norm(x[1] - x[0])] * 3, np.float32), params=p)
    o = Oracle(s)
    e = energy(o, np.float32(x))
    d = np.float64(np.float32(5.4))
    assert e[S.E_ANGLE] == pytest.approx((np.float64(np.float32(5.4)) / d) ** 6, rel=1e-6)


def test_native_lj_minimum_and_repulsion_at_sigma():
    p = S.params_config()
    # 4-bead chain, native pair (0,3) at its native distance -> E = -eps
    x = np.array([[0, 0, 0], [3.8, 0, 0], [3.8, 3.8, 0], [0, 3.8, 0]], np.float32)
    b = np.full(4, 3.8, np.float32)
    s = mk(x, bond_r0=b, nat=[(0, 3, 3.8)], params=p)
    o = Oracle(s)
    f, e, _ = o.forces(x, o.lists(x)[1])
    assert e[S.E_NATIVE] == pytest.approx(-1.0, abs=1e-12)
    # without the native the same pair repels with eps((s/r)^12 + (s/r)^6) = 2 at r = sigma
    s2 = mk(x, bond_r0=b, params=p)
    e2 = energy(Oracle(s2), x)
    assert e2[S.E_REP] == pytest.approx(2.0, rel=1e-6)
    # (sigma/r)^6 form: c12 = 0 -> eps at r = sigma
    p6 = S.copy_params(p)
    p6.rep_c12 = 0.0
    s3 = mk(x, bond_r0=b, params=p6)
    assert energy(Oracle(s3), x)[S.E_REP] == pytest.approx(1.0, rel=1e-6)


def test_type_tables_are_indexed_by_both_types():
    p = S.params_reference()
    p.boxl = 0.0
    x = np.array([[0, 0, 0], [3.8, 0, 0], [7.6, 0, 0], [20, 0, 0]], np.float32)
    s = mk(x, chain=np.array([0, 0, 0, 1], np.int32), types=np.array([1, 1, 1, 2], np.int32),
           bond_r0=np.full(4, 3.8, np.float32), params=p)
    o = Oracle(s)
    nl, pl = o.lists(x)
    # sigma(1,2) = 5.4: pair list cut 2.5*5.4 = 13.5 A keeps only bead 2 (12.4 A away)
    assert sorted(map(tuple, pl[1].tolist())) == [(2, 3)]
    e = o.forces(x, pl)[1][S.E_REP]
    r = 12.4
    assert e == pytest.approx((5.4 / r) ** 12 + (5.4 / r) ** 6, rel=1e-5)


# ---------------------------------------------------- gradient identities

def random_system(seed, n=60, nchains=2, params=None):
    sysd = S.make_chain_system(seed, n, params=params)
    ch = np.repeat(np.arange(nchains), int(math.ceil(n / nchains)))[:n].astype(np.int32)
    keep = [(i, j, r) for i, j, r in zip(sysd.nat_i, sysd.nat_j, sysd.nat_r0)
            if not (ch[i] == ch[j] and abs(int(i) - int(j)) <= 2)]
    rng = np.random.default_rng(seed)
    x = sysd.xyz + rng.normal(0, 0.15, sysd.xyz.shape).astype(np.float32)
    return mk(x, chain=ch, types=rng.integers(1, 3, n).astype(np.int32), bond_r0=sysd.bond_r0, nat=keep,
              params=sysd.params)


@pytest.mark.parametrize("seed", [1, 2])
def test_forces_are_minus_energy_gradient(seed):
    s = random_system(seed)
    o = Oracle(s)
    pl = o.lists(s.xyz)[1]
    x = s.xyz.astype(np.float64)
    f, _, _ = o.forces(x, pl)
    h = 1e-6
    rng = np.random.default_rng(seed)
    for i in rng.choice(len(x), 12, replace=False):
        for d in range(3):
            xp, xm = x.copy(), x.copy()
            xp[i, d] += h
            xm[i, d] -= h
            g = (o.forces(xp, pl)[1][S.E_TOTAL] - o.forces(xm, pl)[1][S.E_TOTAL]) / (2 * h)
            assert -g == pytest.approx(f[i, d], rel=2e-5, abs=2e-5)


def test_newton_third_law_and_invariance():
    s = random_system(3)
    o = Oracle(s)
    pl = o.lists(s.xyz)[1]
    x = s.xyz.astype(np.float64)
    f, e, _ = o.forces(x, pl)
    assert np.abs(f.sum(0)).max() < 1e-9 * np.abs(f).max()
    f2, e2, _ = o.forces(x + np.array([3.0, -7.0, 11.0]), pl)
    assert np.allclose(f, f2, rtol=1e-9, atol=1e-9) and e2[S.E_TOTAL] == pytest.approx(e[S.E_TOTAL], rel=1e-12)
    th = 0.7
    R = np.array([[math.cos(th), -math.sin(th), 0], [math.sin(th), math.cos(th), 0], [0, 0, 1]])
    f3, e3, _ = o.forces(x @ R.T, pl)
    assert np.allclose(f3, f @ R.T, rtol=1e-8, atol=1e-8)
    assert e3[S.E_TOTAL] == pytest.approx(e[S.E_TOTAL], rel=1e-12)


def test_virial_is_scaling_derivative():
    s = random_system(4)
    o = Oracle(s)
    pl = o.lists(s.xyz)[1]
    x = s.xyz.astype(np.float64)
    w = o.forces(x, pl)[1][S.E_VIRIAL]
    h = 1e-6
    dE = (o.forces(x * (1 + h), pl)[1][S.E_TOTAL] - o.forces(x * (1 - h), pl)[1][S.E_TOTAL]) / (2 * h)
    assert w == pytest.approx(-dE, rel=1e-5)


# ----------------------------------------------------------------- lists

def test_line_membership_is_strict():
    # beads on a line 4 A apart: d^2 = 16 (j-i)^2 exactly; absolute cutoffs
    # NL 16 A, PL 12 A; |i-j| > 2 inside a chain
    p = S.params_config()
    p.nl_abs, p.pl_abs = 16.0, 12.0
    x = np.zeros((10, 3), np.float32)
    x[:, 0] = 4.0 * np.arange(10)
    s = mk(x, bond_r0=np.full(10, 4.0, np.float32), params=p)
    o = Oracle(s)
    nl, pl = o.lists(x)
    assert sorted(map(tuple, nl[1].tolist())) == [(i, i + 3) for i in range(7)]  # 12 < 16, 16 excluded
    assert pl[1].shape[0] == 0  # 12 is not < 12
    # two chains: across chains |i-j| <= 2 interacts
    s2 = mk(x, chain=np.array([0] * 5 + [1] * 5, np.int32), bond_r0=np.full(10, 4.0, np.float32), params=p)
    nl2 = Oracle(s2).nl(x)[1]
    want = {(i, i + 3) for i in range(7)} | {(3, 5), (4, 5), (4, 6)}
    assert set(map(tuple, nl2.tolist())) == want


def test_natives_leave_the_repulsive_list():
    p = S.params_config()
    x = np.zeros((8, 3), np.float32)
    x[:, 0] = 4.0 * np.arange(8)
    s = mk(x, bond_r0=np.full(8, 4.0, np.float32), nat=[(0, 3, 12.0), (2, 7, 20.0)], params=p)
    o = Oracle(s)
    nl, pl = o.lists(x)
    assert sorted(nl[0].tolist()) == [0, 1]       # both within 32 A
    assert sorted(pl[0].tolist()) == [0]          # 20 A is outside the 16 A pair cut
    rep = set(map(tuple, nl[1].tolist()))
    assert (0, 3) not in rep and (2, 7) not in rep and (1, 4) in rep


def test_cutoff_factor_rounds_once_from_double():
    # find r0 where float(3.2 * double(r0)) differs from float32(3.2) * r0
    rng = np.random.default_rng(0)
    cands = rng.uniform(3.0, 9.0, 20000).astype(np.float32)
    dbl = (np.float64(3.2) * cands.astype(np.float64)).astype(np.float32)
    sgl = np.float32(3.2) * cands
    idx = np.nonzero(dbl != sgl)[0]
    assert len(idx) > 0
    for k in idx[:50]:
        assert O.cutoff(3.2, 0.0, float(cands[k])) == dbl[k]
    assert O.cutoff(0.0, 16.0, 3.8) == np.float32(16.0)


def test_predicate_has_no_fma():
    # d2 is formed as ((dx*dx) + (dy*dy)) + dz*dz in float with each product
    # rounded; pick coordinates where a fused multiply-add would differ
    rng = np.random.default_rng(1)
    found = 0
    for _ in range(20000):
        a = rng.uniform(-20, 20, 3).astype(np.float32)
        b = rng.uniform(-20, 20, 3).astype(np.float32)
        d = b - a
        plain = (d[0] * d[0] + d[1] * d[1]) + d[2] * d[2]
        fused = np.float32(np.float64(d[2]) * d[2] + np.float64(d[0] * d[0] + d[1] * d[1]))
        if plain != fused:
            x = np.stack([a, b]).reshape(-1).astype(np.float32)
            assert np.float32(O.lib().orc_dist2_f(x.ctypes.data_as(O.C.POINTER(O.C.c_float)), 0, 1, 0.0)) == plain
            found += 1
            if found > 20:
                break
    assert found > 0


def test_min_image_ties_round_half_even():
    L = np.float32(100.0)
    assert O.min_image_f(50.0, L) == np.float32(50.0)     # rint(0.5) = 0
    assert O.min_image_f(-50.0, L) == np.float32(-50.0)
    assert O.min_image_f(150.0, L) == np.float32(-50.0)   # rint(1.5) = 2
    assert O.min_image_f(250.0, L) == np.float32(50.0)    # rint(2.5) = 2
    assert O.min_image_f(49.0, L) == np.float32(49.0)
    assert O.min_image_f(51.0, L) == np.float32(-49.0)
    assert O.min_image_f(51.0, 0.0) == np.float32(51.0)   # no box


def grid_cases():
    rng = np.random.default_rng(17)
    # a 3,000-bead globule cut into 7 chains, two bead types, factor cutoffs
    # (3.2 x sigma per type pair) in an open box
    p = S.params_reference()
    p.boxl = 0.0
    yield random_system(41, n=3000, nchains=7, params=p)
    # the same with the absolute 32 A cutoff
    yield random_system(42, n=2000, nchains=3)
    # clusters far apart in a 3 km-wide cloud: the grid widens its cells
    c = rng.uniform(-1500, 1500, (30, 1, 3))
    x = (c + rng.normal(0, 8.0, (30, 10, 3))).reshape(-1, 3)
    yield mk(x, chain=np.arange(300, dtype=np.int32), bond_r0=np.full(300, 3.8, np.float32))
    # a lattice at 8 A spacing from -100 A: pairs at exactly 32 A (dropped,
    # strict <) and pairs on both sides of cell faces at exact multiples
    g = np.arange(7, dtype=np.float32) * 8.0 - 100.0
    x = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    yield mk(x, chain=np.arange(len(x), dtype=np.int32), bond_r0=np.full(len(x), 3.8, np.float32))
    # a single bead
    yield mk(np.zeros((1, 3)), bond_r0=np.full(1, 3.8, np.float32))


def test_grid_neighbor_list_equals_brute_force():
    # orc_nl_grid must return orc_nl's arrays exactly: same pairs, same
    # row-major order (a stencil missing a neighbour cell, a cell computed
    # from the wrong corner or an unsorted row fails this)
    for k, s in enumerate(grid_cases()):
        o = Oracle(s)
        a = o.nl(s.xyz)
        b = o.nl(s.xyz, grid=True)
        assert np.array_equal(a[0], b[0]), k
        assert np.array_equal(a[1], b[1]), (k, len(a[1]), len(b[1]))
        assert k == 4 or len(a[1]) > 0, k


def test_grid_neighbor_list_refuses_a_periodic_box():
    s = random_system(3, n=200, params=S.params_reference())
    assert s.params.boxl > 0
    with pytest.raises(AssertionError):
        Oracle(s).nl(s.xyz, grid=True)


def test_grid_trajectory_equals_brute_force():
    s = random_system(44, n=400, nchains=3)
    a5, var = S.params_derived(s.params)
    o = Oracle(s)
    xa = o.run(s.xyz, 0, 120, 5, a5, var, 50, 10)
    xb = o.run(s.xyz, 0, 120, 5, a5, var, 50, 10, grid=True)
    assert np.array_equal(xa, xb)


# ------------------------------------------------------------ integrator

def test_noise_free_step_is_a5_times_force():
    s = random_system(5)
    s.params.T = 0.0
    o = Oracle(s)
    a5, var = S.params_derived(s.params)
    assert var == 0.0
    x1 = o.run(s.xyz, 0, 1, 7, a5, var, 50, 10)
    f, _, _ = o.forces(s.xyz, o.lists(s.xyz)[1])
    assert np.allclose(x1, s.xyz.astype(np.float64) + np.float64(a5) * f, rtol=0, atol=1e-12)


def test_free_beads_diffuse_with_T_over_zeta():
    # single-bead chains 100 A apart never interact; <dr^2> = 6 D t, D = T/zeta
    n = 400
    g = np.stack(np.meshgrid(np.arange(8), np.arange(8), np.arange(7), indexing="ij"), -1).reshape(-1, 3)[:n]
    p = S.params_config()
    s = mk(100.0 * g, chain=np.arange(n, dtype=np.int32), params=p)
    o = Oracle(s)
    a5, var = S.params_derived(p)
    nsteps = 200
    x = o.run(s.xyz, 0, nsteps, 11, a5, var, 50, 10)
    msd = ((x - s.xyz.astype(np.float64)) ** 2).sum(1).mean()
    D = p.T / p.zeta
    want = 6 * D * nsteps * p.h
    assert msd == pytest.approx(want, rel=0.08)


def test_one_step_composition():
    s = random_system(6)
    o = Oracle(s)
    a5, var = S.params_derived(s.params)
    step0, seed = 40, 99
    x1 = o.run(s.xyz, step0, 1, seed, a5, var, 50, 10)
    f, _, _ = o.forces(s.xyz, o.lists(s.xyz)[1])
    xi = np.array([O.normals(seed, i, step0) for i in range(s.n)])
    want = s.xyz.astype(np.float64) + np.float64(a5) * (f + np.float64(var) * xi)
    assert np.allclose(x1, want, rtol=0, atol=1e-12)


# ----------------------------------------------------------- observables

def test_rg_pair_sum_equals_centroid_form():
    s = random_system(7, n=80)
    q, c, rg = Oracle(s).observables(s.xyz)
    x = s.xyz.astype(np.float64)
    assert rg == pytest.approx(((x - x.mean(0)) ** 2).sum(1).mean(), rel=1e-10)


def test_q_at_scaled_native_state():
    sysd = S.make_chain_system(8, 100)
    o = Oracle(sysd)
    q, c, _ = o.observables(sysd.xyz)
    assert c == len(sysd.nat_i) and q == 1.0
    cm = sysd.xyz.mean(0)
    assert o.observables((sysd.xyz - cm) * 1.2 + cm)[0] == 1.0
    assert o.observables((sysd.xyz - cm) * 1.3 + cm)[0] == 0.0


# ------------------------------------------------- FP32 port vs FP64 oracle

def test_port_tracks_oracle():
    s = random_system(9, n=100)
    a5, var = S.params_derived(s.params)
    port = O.Port(s)
    port.run(0, 100, 3, a5, var, 50, 10)
    x = Oracle(s).run(s.xyz, 0, 100, 3, a5, var, 50, 10)
    rmsd = np.sqrt(((port.x.reshape(-1, 3) - x) ** 2).sum(1).mean())
    assert rmsd < 1e-4
