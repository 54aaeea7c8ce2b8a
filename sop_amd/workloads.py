"""Seeded synthetic workloads of the five benchmark configurations.

C1  one 128-bead chain (generator seed 1)
C2  one 10,000-bead compact globule, ~50k native contacts (seed 2)
C3  512 replicas, N log-uniform in [200, 5000], T uniform in [0.5, 0.9],
    distinct Philox keys (seed 3)
C4  served stream: Poisson arrivals, N log-normal (median 1500, sigma 0.9)
    clipped to [100, 20000], steps log-uniform, T uniform in [0.5, 0.9] (seed 4)
C5  1000 chains x 500 beads on a 10x10x10 lattice at 60 A, intra-chain natives
    only, no periodic box, no two beads of different chains within 4 A (seed 5)

Chains are compact self-avoiding walks (3.8 A bonds, no non-bonded bead within
4 A) confined to the globule volume of 140 A^3 per bead; natives are the pairs
with |i-j| > 2 closer than 8 A in that state, r0 = their distance.
"""
import numpy as np

from . import System, copy_params, make_chain_system, native_contacts, params_config, generate_chain


def philox_key(config_seed: int, replica: int) -> int:
    """splitmix64 of (config seed, replica id): independent of batch position."""
    z = ((config_seed & 0xFFFFFFFF) << 32 | (replica & 0xFFFFFFFF)) & 0xFFFFFFFFFFFFFFFF
    z = (z + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
    return z ^ (z >> 31)


def c1():
    s = make_chain_system(1, 128)
    s.params.seed = philox_key(1, 0)
    return s


def c2():
    s = make_chain_system(2, 10000)
    s.params.seed = philox_key(2, 0)
    return s


def c3_sizes(n=512, seed=3, lo=200, hi=5000):
    rng = np.random.default_rng(seed)
    sizes = np.round(np.exp(rng.uniform(np.log(lo), np.log(hi), n))).astype(int)
    temps = rng.uniform(0.5, 0.9, n)
    return sizes, temps


def c3(n=512, seed=3):
    sizes, temps = c3_sizes(n, seed)
    out = []
    for r, (nb, T) in enumerate(zip(sizes, temps)):
        s = make_chain_system(seed * 1000003 + r, int(nb))
        s.params.T = float(T)
        s.params.seed = philox_key(seed, r)
        out.append(s)
    return out


def c4_jobs(rate=200.0, duration=600.0, seed=4, steps_lo=1e4, steps_hi=1e6, sigma=0.9, max_jobs=None):
    """Arrival schedule: list of (t_arrival, n_beads, nsteps, T, job_seed)."""
    rng = np.random.default_rng(seed)
    t, out = 0.0, []
    while True:
        t += rng.exponential(1.0 / rate)
        if t > duration or (max_jobs is not None and len(out) >= max_jobs):
            break
        nb = int(np.clip(np.round(np.exp(rng.normal(np.log(1500.0), sigma))), 100, 20000))
        st = int(np.round(np.exp(rng.uniform(np.log(steps_lo), np.log(steps_hi)))))
        out.append((t, nb, st, float(rng.uniform(0.5, 0.9)), len(out)))
    return out


def c4_system(job, seed=4):
    t, nb, st, T, k = job
    s = make_chain_system(seed * 1000003 + k, nb)
    s.params.T = T
    s.params.seed = philox_key(seed, k)
    return s


def _rotation(rng):
    """Uniform random rotation matrix (QR of a Gaussian matrix, signs fixed)."""
    q, r = np.linalg.qr(rng.normal(size=(3, 3)))
    q = q * np.sign(np.diag(r))
    if np.linalg.det(q) < 0:
        q[:, 0] = -q[:, 0]
    return q


def c5(nchains=1000, per=500, spacing=60.0, seed=5, min_dist=4.0, tries=64):
    """Chains go onto the lattice in site order.  A chain with a bead closer
    than min_dist (the walk's own non-bonded minimum) to a bead of a chain
    already placed on one of the 26 neighbouring sites is turned about its
    centroid by a seeded random rotation, up to `tries` times, and otherwise
    regenerated from a new seed; sites two or more apart are >= 120 A away,
    beyond the reach of two confined 500-bead walks."""
    from scipy.spatial import cKDTree

    g = int(round(nchains ** (1 / 3)))
    rng = np.random.default_rng(seed)
    placed = {}
    xs, bs, nis, njs, nrs, ch = [], [], [], [], [], []
    off = 0
    for c in range(nchains):
        site = (c % g, (c // g) % g, c // (g * g))
        near = [placed[(site[0] + a, site[1] + b, site[2] + d)]
                for a in (-1, 0, 1) for b in (-1, 0, 1) for d in (-1, 0, 1)
                if (site[0] + a, site[1] + b, site[2] + d) in placed]
        tree = cKDTree(np.concatenate(near)) if near else None
        gen_seed = seed * 1000003 + c
        while True:
            x0, b = generate_chain(gen_seed, per)
            x0 = x0.astype(np.float64)
            x0 -= x0.mean(0)
            for t in range(tries):
                y = (x0 if t == 0 else x0 @ _rotation(rng).T) + spacing * np.array(site, np.float64)
                x = y.astype(np.float32)
                if tree is None or not np.isfinite(tree.query(x, distance_upper_bound=min_dist)[0]).any():
                    break
            else:
                gen_seed += 7919 * nchains
                continue
            break
        # bond r0 = the bond lengths as placed (rotation moves them by float rounding)
        b = b.copy()
        b[:-1] = np.linalg.norm(x[1:].astype(np.float64) - x[:-1], axis=1).astype(np.float32)
        placed[site] = x
        ni, nj, nr = native_contacts(x, None, 8.0)
        xs.append(x)
        bs.append(b)
        nis.append(ni + off)
        njs.append(nj + off)
        nrs.append(nr)
        ch.append(np.full(per, c, np.int32))
        off += per
    p = params_config()
    p.seed = philox_key(seed, 0)
    return System(xyz=np.concatenate(xs), params=p, chain=np.concatenate(ch), bond_r0=np.concatenate(bs),
                  nat_i=np.concatenate(nis), nat_j=np.concatenate(njs), nat_r0=np.concatenate(nrs))
