"""Long-run statistical comparison of two sets of overdamped trajectories.

Each set is an array of samples of shape (replicas, samples, 4) holding
(E_total, Rg^2, Q, contacts) recorded every nup steps, as the reference's
update.out rows (sop.cu:394-396) and calculate_observables (sop.cu:528-611)
report them.  The replicas of one set are independent (distinct Philox keys);
the two sets use disjoint keys, so they are independent samples of the same
process when the dynamics agree.

Two checks per observable, on the samples after the burn-in:
  means  block averages, one block per replica: the difference of the two
         sets' means over their replica means is within `zmax` standard
         errors, the error from the spread of the replica means;
  KS     two-sample Kolmogorov-Smirnov test on decorrelated samples (every
         `stride`-th sample of each replica, stride >= 2 tau_int + 1 with the
         integrated autocorrelation time of the residuals about the set's
         mean time profile), p-value above `alpha`.
The sets may drift together (a chain unfolding from its native state): both
checks compare the same time windows of the same process.
"""
import numpy as np
from scipy import stats

OBS = ("E_total", "Rg2", "Q")


def tau_int(res, cmax=0.05, wmax=None):
    """Integrated autocorrelation time (in samples) of residual series res
    (replicas, samples): 1 + 2 sum of the pooled autocorrelation up to the
    first lag where it falls below cmax."""
    r = res - res.mean(1, keepdims=True)
    v = (r * r).mean()
    if v <= 0:
        return 0.0
    m = r.shape[1]
    wmax = wmax or m // 4
    t = 1.0
    for k in range(1, wmax):
        c = (r[:, :-k] * r[:, k:]).mean() / v
        if c < cmax:
            break
        t += 2.0 * c
    return t


def compare(a, b, burn, zmax=3.0, alpha=0.01):
    """Returns {observable: dict(z, p, stride, mean_a, mean_b, se)} for sample
    sets a and b of shape (replicas, samples, >=3)."""
    out = {}
    for k, name in enumerate(OBS):
        xa, xb = a[:, burn:, k], b[:, burn:, k]
        ma, mb = xa.mean(1), xb.mean(1)
        se = np.sqrt(ma.var(ddof=1) / len(ma) + mb.var(ddof=1) / len(mb))
        z = (ma.mean() - mb.mean()) / se if se > 0 else (0.0 if ma.mean() == mb.mean() else np.inf)
        ta = tau_int(xa - xa.mean(0, keepdims=True))
        tb = tau_int(xb - xb.mean(0, keepdims=True))
        stride = int(np.ceil(2.0 * max(ta, tb) + 1.0))
        p = stats.ks_2samp(xa[:, ::stride].ravel(), xb[:, ::stride].ravel()).pvalue
        out[name] = {"z": float(z), "p": float(p), "stride": stride, "mean_a": float(ma.mean()),
                     "mean_b": float(mb.mean()), "se": float(se), "zmax": zmax, "alpha": alpha,
                     "n_ks": int(xa[:, ::stride].size)}
    return out


def agree(res):
    return all(abs(r["z"]) <= r["zmax"] and r["p"] > r["alpha"] for r in res.values())


def oracle_set(system, keys, nsteps, nup, threads=8, T=None):
    """Sample sets of the FP64 oracle, one replica per Philox key, run on
    host threads.  T overrides the temperature (the noise amplitude var)."""
    from concurrent.futures import ThreadPoolExecutor

    import sop_amd as S
    from oracle.oracle import Oracle

    p = S.copy_params(system.params)
    if T is not None:
        p.T = T
    a5, var = S.params_derived(p)
    o = Oracle(system)

    def one(key):
        return o.run_sample(system.xyz, 0, nsteps, int(key), a5, var, p.n_nl, p.n_pl, nup)[1]

    with ThreadPoolExecutor(threads) as ex:
        return np.stack(list(ex.map(one, keys)))
