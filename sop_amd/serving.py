"""Continuous-batching server (libsop serve loop) and the served benchmark leg.

A Server owns an Engine in serving mode: jobs (System + step count) are
submitted as they arrive; the engine's serve thread admits them by HBM budget,
runs them in the resident batch in hipGraph blocks, and retires each as it
finishes.  Latency is arrival (submit) to result available on the host.
"""
import ctypes as C
import threading
import time

import numpy as np

from . import Engine, System, SopError, _abi, _check, _ptr, lib, params_config
from . import workloads as W


class Server:
    def __init__(self, engine: Engine, hbm_budget_bytes=0, max_jobs=0, block_steps=0):
        self.eng = engine
        o = _abi.ServeOpts()
        o.hbm_budget_bytes, o.max_jobs, o.block_steps = hbm_budget_bytes, max_jobs, block_steps
        engine._c(lib().sop_serve_start(engine._h, C.byref(o)))
        self._n = {}

    def job_bytes(self, system: System) -> int:
        d = system.desc()
        return lib().sop_serve_job_bytes(self.eng._h, C.byref(d))

    def submit(self, system: System, nsteps: int) -> int:
        d = system.desc()
        j = C.c_int64()
        _check(lib().sop_serve_submit(self.eng._h, C.byref(d), nsteps, C.byref(j)))
        self._n[j.value] = system.n
        return j.value

    def poll(self, job, with_coords=False):
        """Returns (done, result dict, coords or None)."""
        r = _abi.JobResult()
        x = np.zeros((self._n[job], 3), np.float32) if with_coords else None
        st = lib().sop_serve_poll(self.eng._h, job, C.byref(r), _ptr(x, C.c_float) if x is not None else None)
        if st == _abi.SOP_ERR_BUSY:
            return False, None, None
        _check(st)
        res = {"status": r.status, "n_beads": r.n_beads, "steps": r.steps, "t_submit": r.t_submit,
               "t_admit": r.t_admit, "t_done": r.t_done, "energies": np.array(r.energies[:])}
        return True, res, x

    def wait(self, job, timeout=600.0, with_coords=True):
        t0 = time.time()
        while True:
            done, res, x = self.poll(job, with_coords)
            if done:
                return res, x
            if time.time() - t0 > timeout:
                raise TimeoutError(f"job {job} not done after {timeout}s")
            time.sleep(0.001)

    def stats(self) -> dict:
        st = _abi.ServeStats()
        self.eng._c(lib().sop_serve_stats_get(self.eng._h, C.byref(st)))
        return {f: getattr(st, f) for f, _ in _abi.ServeStats._fields_}

    def stop(self):
        self.eng._c(lib().sop_serve_stop(self.eng._h))


def bench_served(seconds, log, rate=200.0, steps_lo=1e3, steps_hi=1e4, templates=256, seed=4,
                 budget_gib=32.0, drain_timeout=120.0, lanes_per_bead=4, block_steps=0, window=None,
                 resident=None, resident_steps=0, tune=None):
    """Served workload: Poisson arrivals at `rate` jobs/s for `seconds`; job
    sizes log-normal (median 1500, sigma 0.9) clipped to [100, 20000]; steps
    log-uniform in [steps_lo, steps_hi]; T uniform in [0.5, 0.9]; fresh Philox
    key per job.  Chains come from `templates` pre-generated compact walks of
    the drawn sizes (job k uses template k % templates).  The engine's step
    kernel uses `lanes_per_bead` lanes per bead; the server runs `block_steps`
    steps per graph replay (0: the jobs' n_nl).

    window=None: wait until every job is done (or drain_timeout) and report
    latency percentiles and the completed jobs' bead-steps/s.  window=W: the
    measurement ends W seconds after the first arrival, whatever is still
    queued or running (an offered load above capacity never drains): the
    server's capacity is the bead-steps it ran in the window (the serve loop's
    own count, finished jobs or not) / W; latencies are those of the jobs
    completed in the window, and the rest are reported as censored.

    resident: a System submitted before the stream with resident_steps steps
    (a large assembly run alongside the served stream); its steps/s from
    admission to completion is reported.  tune: settings of Engine.autotune_ex
    applied to the serving engine before it starts."""
    jobs = W.c4_jobs(rate=rate, duration=seconds, seed=seed, steps_lo=steps_lo, steps_hi=steps_hi)
    t0 = time.perf_counter()
    base = [W.c4_system(jobs[k], seed) for k in range(min(templates, len(jobs)))]
    log(f"served: {len(jobs)} arrivals over {seconds}s, steps in [{steps_lo:.0f}, {steps_hi:.0f}], "
        f"{len(base)} chain templates in {time.perf_counter() - t0:.1f}s")
    eng = Engine(0, lanes_per_bead=lanes_per_bead)
    if tune:
        eng.tune_set(tune)
    srv = None
    res_out = None
    try:
        srv = Server(eng, hbm_budget_bytes=int(budget_gib * (1 << 30)), block_steps=block_steps)
        res_id = srv.submit(resident, resident_steps) if resident is not None else None
        ids = []
        start = time.perf_counter()
        clock0 = lib().sop_clock_seconds()
        stop_at = start + window if window is not None else None
        for (t, nb, nst, T, k) in jobs:
            now = time.perf_counter() - start
            if stop_at is not None and start + t > stop_at:
                break
            if t > now:
                time.sleep(t - now)
            s = base[k % len(base)]
            s.params.T = T
            s.params.seed = W.philox_key(seed, k)
            ids.append((srv.submit(s, nst), s.n, nst))
        log(f"served: {len(ids)} jobs submitted at {time.perf_counter() - start:.1f}s")
        lat, done_beadsteps, t_last = [], 0, clock0
        pending = list(ids)
        deadline = stop_at if stop_at is not None else time.perf_counter() + drain_timeout
        res_done = res_id is None
        while (pending or not res_done) and time.perf_counter() < deadline:
            keep = []
            for (j, n, nst) in pending:
                done, r, _ = srv.poll(j)
                if done:
                    if r["status"] != 0:
                        msg = lib().sop_last_error(eng._h).decode(errors="replace")
                        raise SopError(r["status"], f"served job {j} ({n} beads, {nst} steps) failed: {msg}")
                    lat.append(r["t_done"] - r["t_submit"])
                    done_beadsteps += n * nst
                    t_last = max(t_last, r["t_done"])
                else:
                    keep.append((j, n, nst))
            pending = keep
            if not res_done:
                done, r, _ = srv.poll(res_id)
                if done:
                    res_done = True
                    if r["status"] != 0:
                        raise SopError(r["status"], "resident job failed")
                    dt = r["t_done"] - r["t_admit"]
                    res_out = {"n_beads": resident.n, "steps": resident_steps, "seconds": dt,
                               "steps_per_s": resident_steps / dt, "bead_steps_per_s": resident.n * resident_steps / dt}
            if pending or not res_done:
                time.sleep(0.01)
        t_end = lib().sop_clock_seconds()
        loop = srv.stats()
    finally:
        # stop the serve thread before the engine goes, whatever happened above
        try:
            if srv is not None:
                srv.stop()
        except SopError as ex:
            log(f"served: the serve loop stopped with an error: {ex}")
        finally:
            eng.close()
    lat = np.array(lat)
    offered = sum(n * nst for (_, n, nst) in ids) / seconds if seconds > 0 else 0.0
    out = {"jobs": len(ids), "completed": int(len(lat)), "censored": len(ids) - int(len(lat)),
           "arrival_rate": rate, "arrival_seconds": seconds, "window_s": window,
           "steps_range": [steps_lo, steps_hi], "templates": len(base), "block_steps": block_steps,
           "p50_latency_s": float(np.percentile(lat, 50)) if len(lat) else None,
           "p99_latency_s": float(np.percentile(lat, 99)) if len(lat) else None,
           "offered_bead_steps_per_s": offered,
           "loop": {k: (round(v, 4) if isinstance(v, float) else v) for k, v in loop.items()}}
    if window is None:
        wall = t_last - clock0
        out["bead_steps_per_s"] = done_beadsteps / wall if wall > 0 else None
        out["wall_s"] = wall
    else:
        out["capacity_bead_steps_per_s"] = loop["bead_steps"] / (t_end - clock0)
        out["completed_bead_steps_per_s"] = done_beadsteps / (t_end - clock0)
        out["wall_s"] = t_end - clock0
    if resident is not None:
        out["resident"] = res_out
    p50, p99 = out["p50_latency_s"], out["p99_latency_s"]
    rate_s = out.get("bead_steps_per_s") or out.get("capacity_bead_steps_per_s")
    log(f"served: {out['completed']}/{out['jobs']} done ({out['censored']} censored), "
        f"p50 {p50 if p50 is None else round(p50, 3)}s p99 {p99 if p99 is None else round(p99, 3)}s, "
        f"{rate_s if rate_s is None else round(rate_s / 1e9, 3)}e9 bead-steps/s (offered {offered:.4g})"
        + (f", resident {res_out}" if resident is not None else ""))
    return out
