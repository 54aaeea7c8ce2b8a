#!/usr/bin/env python3
"""Throughput of the SOP overdamped hot path on one MI355X.

One bench step = one overdamped step (lists rebuilt on schedule, fused
force + Philox + move) of every replica of the batched workload:
512 independent chains, N log-uniform in [200, 5000], T uniform in [0.5, 0.9],
distinct Philox keys, 32 A / 16 A neighbor / pair lists every 50 / 10 steps
(the batched-replica configuration).  Inputs are generated on the host and are
resident in HBM before the timed region.

Printed on stdout: one JSON line.  Also measured in the same run:
  plain        the same replicas, one system per launch, no graphs
  roofline     HBM roofline of the fused step kernel (algorithmic bytes per
               launch / average launch time from HIP events), beside a
               measured device copy rate; traffic = HBM bytes per launch
               and VALU instructions from child runs under rocprofv3 --pmc
  served       continuous-batching request loop at a load it sustains
               (steps in [1e3, 1e4]; p50/p99 job latency)
  served_c4    the served stream at its own step range [1e4, 1e6]: offered
               load above capacity, measured over a bounded window (capacity,
               p50/p99 of completed jobs, censored count)
  assembly     the 500,000-bead, 1,000-chain assembly alone, resident in the
               server alone, and beside the served stream
  cpu_baseline FP32 port of the reference's host method on the host cores
               (one GPU's CPU share, and one core)
Progress goes to stderr.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "bead-steps/s batched+served vs plain per-system on MI355X; p50/p99 job latency"
HBM_PEAK = 8.0e12
VALU_PEAK = 256 * 64 * 2.4e9  # FP32 lane-ops/s: 256 CUs x 64 lanes x 2.4 GHz, no packed FP32


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def c3_engine(S, args, lanes=None, grid=None, tiles=None):
    return S.Engine(0, lanes_per_bead=lanes or args.lanes, lds_beads_max=args.lds_beads, pl_shell=args.pl_shell,
                    pl_cap_max=args.pl_cap, step_tiles=args.step_tiles if tiles is None else tiles,
                    step_blocks=args.step_blocks if grid is None else grid)


def algorithmic_bytes(eng, ids, systems, n_pl, n_nl):
    """Per step (force kernel) and per bead-step (whole schedule), with the
    measured half-list lengths: 24N + 4B + 8Pa + 4Pr for the force step;
    + (12N + 8Qa + 4Qr + 8Pa + 4Pr)/n_pl + (28N + 8Qa + 4Qr)/n_nl for rebuilds."""
    step = 0.0
    full = 0.0
    beads = 0
    builds = [0, 0, 0]
    for i, s in zip(ids, systems):
        st = eng.list_stats(i)
        builds[0] += st.grid_builds
        builds[1] += st.shell_builds
        builds[2] += st.nl_builds
        n = s.n
        b = n - 1
        pa, pr, qa, qr = st.pl_native, st.pl_rep, st.nl_native, st.nl_rep
        f = 24 * n + 4 * b + 8 * pa + 4 * pr
        step += f
        full += f + (12 * n + 8 * qa + 4 * qr + 8 * pa + 4 * pr) / n_pl + (28 * n + 8 * qa + 4 * qr) / n_nl
        beads += n
    log(f"pair-list rebuilds over all systems: {builds[0]} from the cell grid, {builds[1]} from shell rows, "
        f"{builds[2]} from the neighbor list")
    return step, full / beads


def cpu_baseline(systems, steps=1000, warm=100, every=8, one_core_every=64):
    """The oracle's FP32 port of the reference host method on the host cores:
    every `every`-th replica on the box's CPU share for one GPU (16 threads,
    or nproc if fewer), and every `one_core_every`-th on one thread.  Each
    times `steps` steps after `warm` warm-up steps (bead-steps/s)."""
    from concurrent.futures import ThreadPoolExecutor

    import sop_amd as S
    from oracle import oracle as O
    nproc = os.cpu_count() or 1

    def timed(sample, threads):
        ports = [O.Port(s) for s in sample]

        def go(k, st0, nst):
            s = sample[k]
            a5, var = S.params_derived(s.params)
            ports[k].run(st0, nst, s.params.seed, a5, var, s.params.n_nl, s.params.n_pl)

        with ThreadPoolExecutor(threads) as ex:
            list(ex.map(lambda k: go(k, 0, warm), range(len(ports))))
            t0 = time.perf_counter()
            list(ex.map(lambda k: go(k, warm, steps), range(len(ports))))
            dt = time.perf_counter() - t0
        for p in ports:
            p.close()
        beads = sum(s.n for s in sample)
        return beads * steps / dt, beads, dt

    sample = systems[::every]
    threads = max(1, min(16, nproc, len(sample)))
    v_all, b_all, dt_all = timed(sample, threads)
    one = systems[::one_core_every]
    v_one, b_one, dt_one = timed(one, 1)
    return {"value": v_all, "unit": "bead-steps/s", "cores": threads, "kind": "port", "nproc": nproc,
            "sample": f"{len(sample)} replicas of the workload (every {every}th, {b_all} beads), {steps} steps after "
                      f"{warm} warm-up, FP32 port of the reference host method (O(N^2) master-list filter, "
                      f"half-list force loops, overdamped step), {threads} threads over replicas (the CPU share "
                      f"of one GPU on a {nproc}-thread host)",
            "seconds": dt_all,
            "one_core": {"value": v_one, "cores": 1, "seconds": dt_one,
                         "sample": f"{len(one)} replicas (every {one_core_every}th, {b_one} beads), {steps} steps "
                                   f"after {warm} warm-up, one thread"}}


def copy_calibration(nbytes):
    """A plain device-to-device copy of known size (the runtime's copy kernel),
    so that FETCH_SIZE / WRITE_SIZE counters can be scaled to bytes."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(nbytes)) == 0 and hip.hipMalloc(C.byref(b), C.c_size_t(nbytes)) == 0
    assert hip.hipMemset(a, 1, C.c_size_t(nbytes)) == 0
    for _ in range(3):
        assert hip.hipMemcpy(b, a, C.c_size_t(nbytes), 3) == 0  # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    log(f"copied {nbytes} bytes device-to-device 3 times")


def copy_bandwidth(nbytes=1 << 30, reps=5):
    """Measured device copy rate: read + write bytes of a device-to-device
    hipMemcpy of nbytes, timed with HIP events (best of reps), in GB/s."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    a, b = C.c_void_p(), C.c_void_p()
    ev0, ev1 = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(nbytes)) == 0 and hip.hipMalloc(C.byref(b), C.c_size_t(nbytes)) == 0
    assert hip.hipMemset(a, 1, C.c_size_t(nbytes)) == 0
    assert hip.hipEventCreate(C.byref(ev0)) == 0 and hip.hipEventCreate(C.byref(ev1)) == 0
    best = 1e30
    for _ in range(reps):
        hip.hipEventRecord(ev0, None)
        assert hip.hipMemcpyAsync(b, a, C.c_size_t(nbytes), 3, None) == 0
        hip.hipEventRecord(ev1, None)
        assert hip.hipEventSynchronize(ev1) == 0
        ms = C.c_float()
        hip.hipEventElapsedTime(C.byref(ms), ev0, ev1)
        best = min(best, ms.value)
    hip.hipEventDestroy(ev0)
    hip.hipEventDestroy(ev1)
    hip.hipFree(a)
    hip.hipFree(b)
    return 2.0 * nbytes / (best * 1e-3) / 1e9


def pmc_probe(counters, child_args, kernel_prefix, timeout=180):
    """Runs this script as a child under rocprofv3 --pmc and returns the mean
    value per dispatch of each counter over the dispatches whose kernel name
    starts with kernel_prefix (None when rocprofv3 is missing or the run gave
    no rows)."""
    import csv
    import glob
    import shutil
    import subprocess
    import tempfile
    prof = shutil.which("rocprofv3") or "/opt/rocm/bin/rocprofv3"
    if not os.path.exists(prof):
        return None
    d = tempfile.mkdtemp(prefix="sop_pmc_")
    cmd = [prof, "--pmc", *counters, "--output-format", "csv", "-d", d, "-o", "p", "--",
           sys.executable, os.path.abspath(__file__), *child_args]
    try:
        r = subprocess.run(cmd, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL, timeout=timeout)
        if r.returncode != 0:
            log(f"rocprofv3 {' '.join(counters)}: exit {r.returncode} (results read if complete)")
    except subprocess.TimeoutExpired:
        log(f"rocprofv3 {' '.join(counters)}: timed out")
    acc = {}
    for f in glob.glob(os.path.join(d, "**", "*counter_collection*.csv"), recursive=True):
        with open(f) as fh:
            for row in csv.DictReader(fh):
                if not row.get("Kernel_Name", "").startswith(kernel_prefix):
                    continue
                a = acc.setdefault(row["Counter_Name"], [0.0, 0])
                a[0] += float(row["Counter_Value"])
                a[1] += 1
    shutil.rmtree(d, ignore_errors=True)
    if not acc:
        return None
    out = {k: v[0] / v[1] for k, v in acc.items()}
    out["dispatches"] = max(v[1] for v in acc.values())
    return out


def traffic_probe(child_args, kernel_prefix):
    """HBM bytes per dispatch of the kernel (FETCH_SIZE and WRITE_SIZE, KiB
    per dispatch, each scaled by what the same counter reads for a device copy
    of known size) and its VALU and LDS instruction counts, from child runs of
    this script under rocprofv3 --pmc."""
    nb = 256 << 20
    copy_args = ["--calibrate-copy", str(nb)]
    cf = pmc_probe(["FETCH_SIZE"], copy_args, "__amd_rocclr_copyBuffer")
    cw = pmc_probe(["WRITE_SIZE"], copy_args, "__amd_rocclr_copyBuffer")
    kf = pmc_probe(["FETCH_SIZE", "SQ_INSTS_VALU", "SQ_INSTS_LDS", "SQ_WAVES"], child_args, kernel_prefix)
    kw = pmc_probe(["WRITE_SIZE"], child_args, kernel_prefix)
    if not (cf and cw and kf and kw):
        return None
    f_fetch = nb / (cf["FETCH_SIZE"] * 1024.0)
    f_write = nb / (cw["WRITE_SIZE"] * 1024.0)
    read = kf["FETCH_SIZE"] * 1024.0 * f_fetch
    write = kw["WRITE_SIZE"] * 1024.0 * f_write
    out = {"bytes": read + write, "read": read, "write": write, "fetch_scale": f_fetch, "write_scale": f_write,
           "valu_insts": kf.get("SQ_INSTS_VALU"), "lds_insts": kf.get("SQ_INSTS_LDS"), "waves": kf.get("SQ_WAVES"),
           "dispatches": kf["dispatches"]}
    # the read requests by size: FETCH_SIZE counts a 128-byte request as 128
    # bytes only through TCC_BUBBLE, which reads 0 on this device, so the
    # request counts give the reads independently of the copy's scale
    kr = pmc_probe(["TCC_EA0_RDREQ_sum", "TCC_EA0_RDREQ_64B_sum", "TCC_EA0_RDREQ_128B_sum"], child_args,
                   kernel_prefix)
    if kr:
        n_all, n64, n128 = kr["TCC_EA0_RDREQ_sum"], kr["TCC_EA0_RDREQ_64B_sum"], kr["TCC_EA0_RDREQ_128B_sum"]
        out["read_requests"] = {"all": n_all, "64B": n64, "128B": n128,
                                "bytes": 32.0 * max(0.0, n_all - n64 - n128) + 64.0 * n64 + 128.0 * n128}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--replicas", type=int, default=512)
    ap.add_argument("--plain-steps", type=int, default=50)
    ap.add_argument("--serve-seconds", type=float, default=20.0,
                    help="served leg at a sustainable load: arrivals for this long, steps log-uniform in "
                         "[1e3, 1e4] (0: skip)")
    ap.add_argument("--c4-window", type=float, default=30.0,
                    help="served leg at the stream's own step range [1e4, 1e6]: seconds of arrivals and "
                         "measurement (0: skip)")
    ap.add_argument("--c5-steps", type=int, default=200, help="steps of the 500,000-bead assembly legs (0: skip)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-traffic", action="store_true", help="skip the rocprofv3 counter probes")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--lanes", type=int, default=4)
    ap.add_argument("--step-tiles", type=int, default=0, help="tiles per step-kernel item (0: the engine's choice)")
    ap.add_argument("--step-blocks", type=int, default=0, help="step-kernel grid (0: the engine's default)")
    ap.add_argument("--lds-beads", type=int, default=0, help="LDS-resident step for systems up to this size "
                    "(0 = engine default, < 0 = off)")
    ap.add_argument("--serve-block", type=int, default=0,
                    help="steps per served graph replay (0: the jobs' neighbor-list interval)")
    ap.add_argument("--serve-lanes", type=int, default=0,
                    help="lanes per bead of the served legs' engines (0: the batched leg's, after autotune)")
    ap.add_argument("--pl-cap", type=int, default=0, help="per-bead pair-list capacity bound (0: the engine's)")
    ap.add_argument("--pl-shell", type=float, default=None,
                    help="pair-list shell width in A (default: the engine's, 3 A; < 0 = none)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="BYTES",
                    help="only run a device-to-device copy of BYTES (counter calibration for profiles)")
    ap.add_argument("--traffic-probe", action="store_true",
                    help="only run the batched workload's warm-up and steps (the counter probes' child run)")
    ap.add_argument("--tune", default=None, help="settings (JSON of Engine.autotune_ex) for --traffic-probe")
    args = ap.parse_args()
    if args.calibrate_copy:
        copy_calibration(args.calibrate_copy)
        return
    if args.gpus != 1:
        print(f"bench.py: this engine runs on one GPU; --gpus {args.gpus} is not supported "
              "(replicas and served jobs would be divided among GPUs, one engine per GPU, with no communication)",
              file=sys.stderr)
        sys.exit(2)

    import sop_amd as S
    from sop_amd import workloads as W

    t = time.perf_counter()
    systems = W.c3(args.replicas)
    beads = sum(s.n for s in systems)
    log(f"workload: {len(systems)} replicas, {beads} beads, generated in {time.perf_counter() - t:.1f}s")
    if args.traffic_probe:
        with c3_engine(S, args) as eng:
            if args.tune:
                eng.tune_set(json.loads(args.tune))
            ids = [eng.add(s) for s in systems]
            eng.run(ids, args.warmup)
            eng.run(ids, args.steps)
        return

    eng = c3_engine(S, args)
    t = time.perf_counter()
    ids = [eng.add(s) for s in systems]
    log(f"resident on device in {time.perf_counter() - t:.1f}s")
    lanes, grid, tiles = args.lanes, args.step_blocks, args.step_tiles
    tuned = tuned_served = None
    if not args.no_autotune:
        t = time.perf_counter()
        tuned = eng.autotune_ex(ids)
        lanes, grid = tuned["lanes_per_bead"], tuned["step_blocks"]
        log(f"autotune ({time.perf_counter() - t:.1f}s, {tuned['trials']} trials of {tuned['block_steps']}-step "
            f"blocks): {json.dumps(tuned)}")

    eng.run(ids, args.warmup)
    log(f"warmup {args.warmup} steps done")
    t0 = time.perf_counter()
    eng.run(ids, args.steps)
    dt = time.perf_counter() - t0
    value = beads * args.steps / dt
    log(f"timed {args.steps} steps: {dt:.3f}s -> {value:.4g} bead-steps/s")

    p = systems[0].params
    step_bytes, bytes_per_bead_step = algorithmic_bytes(eng, ids, systems, p.n_pl, p.n_nl)
    avg_ms = eng.time_step_kernel(ids, 20)
    achieved = step_bytes / (avg_ms * 1e-3)
    log(f"step kernel {avg_ms:.4f} ms/launch, {achieved / 1e9:.1f} GB/s algorithmic")

    plain_value = None
    if args.plain_steps > 0:
        t0 = time.perf_counter()
        eng.run_plain(ids, args.plain_steps)
        pdt = time.perf_counter() - t0
        plain_value = beads * args.plain_steps / pdt
        log(f"plain path {args.plain_steps} steps: {pdt:.3f}s -> {plain_value:.4g} bead-steps/s")
    eng.close()

    from sop_amd import serving
    slanes = args.serve_lanes or lanes
    if not args.no_autotune:
        # settings for the served mix: tuned on a resident batch of the stream's job sizes
        t = time.perf_counter()
        jobs = W.c4_jobs(rate=200.0, duration=2.0, seed=44)
        with S.Engine(0, lanes_per_bead=slanes) as et:
            tids = [et.add(W.c4_system(j, 44)) for j in jobs[:256]]
            tuned_served = et.autotune_ex(tids)
        slanes = tuned_served["lanes_per_bead"]
        log(f"autotune on 256 served-mix jobs ({time.perf_counter() - t:.1f}s): {json.dumps(tuned_served)}")
    served = served_c4 = None
    if args.serve_seconds > 0:
        served = serving.bench_served(args.serve_seconds, log, lanes_per_bead=slanes, block_steps=args.serve_block,
                                      tune=tuned_served)
        served["config"] = "reduced step range [1e3, 1e4] (a load one GPU sustains), drained to the last job"
    if args.c4_window > 0:
        served_c4 = serving.bench_served(args.c4_window, log, steps_lo=1e4, steps_hi=1e6, templates=1024,
                                         lanes_per_bead=slanes, block_steps=args.serve_block, window=args.c4_window,
                                         tune=tuned_served)
        served_c4["config"] = ("the stream's stated parameters: 200 jobs/s, steps log-uniform [1e4, 1e6]; offered "
                               "load above capacity, measured over a bounded window")
        served_c4["capacity_vs_batched"] = served_c4["capacity_bead_steps_per_s"] / value

    assembly = None
    if args.c5_steps > 0:
        t = time.perf_counter()
        c5 = W.c5()
        log(f"assembly: {c5.n} beads, {len(c5.nat_i)} natives, generated in {time.perf_counter() - t:.1f}s")
        with S.Engine(0, lanes_per_bead=slanes) as e5:
            if tuned_served:
                e5.tune_set(tuned_served)
            i5 = e5.add(c5)
            e5.run([i5], 50)
            t0 = time.perf_counter()
            e5.run([i5], args.c5_steps)
            d5 = time.perf_counter() - t0
        alone = {"bead_steps_per_s": c5.n * args.c5_steps / d5, "ms_per_step": d5 / args.c5_steps * 1e3,
                 "steps": args.c5_steps, "path": "batched engine, graph blocks"}
        log(f"assembly alone: {alone['ms_per_step']:.3f} ms/step, {alone['bead_steps_per_s']:.4g} bead-steps/s")
        rsteps = 25 * args.c5_steps
        res_alone = serving.bench_served(0.0, log, lanes_per_bead=slanes, resident=c5, resident_steps=rsteps,
                                         tune=tuned_served)
        res_mix = serving.bench_served(min(20.0, args.c4_window or 20.0), log, steps_lo=1e4, steps_hi=1e6,
                                       templates=1024, lanes_per_bead=slanes, window=min(20.0, args.c4_window or 20.0),
                                       resident=c5, resident_steps=rsteps, tune=tuned_served)
        assembly = {"n_beads": c5.n, "alone": alone, "served_alone": res_alone["resident"],
                    "served_with_stream": res_mix["resident"],
                    "stream_with_assembly": {k: res_mix[k] for k in ("jobs", "completed", "censored", "p50_latency_s",
                                                                     "p99_latency_s", "capacity_bead_steps_per_s",
                                                                     "window_s")}}

    copy_gbs = copy_bandwidth()
    log(f"device copy: {copy_gbs:.0f} GB/s (read + write)")
    traffic = None
    if not args.no_traffic:
        child = ["--traffic-probe", "--replicas", str(args.replicas), "--steps", "20", "--warmup", "10",
                 "--lanes", str(lanes), "--step-blocks", str(grid), "--step-tiles", str(tiles),
                 "--lds-beads", str(args.lds_beads), "--pl-cap", str(args.pl_cap)]
        if args.pl_shell is not None:
            child += ["--pl-shell", str(args.pl_shell)]
        if tuned:
            child += ["--tune", json.dumps(tuned)]
        traffic = traffic_probe(child, f"void sopk::k_step<{lanes}, 0>")
        if traffic:
            log(f"step kernel HBM traffic {traffic['bytes'] / 1e6:.1f} MB/launch "
                f"({traffic['bytes'] / step_bytes:.2f}x algorithmic), {traffic['valu_insts']:.4g} VALU instructions")
            rq = traffic.get("read_requests")
            if rq:
                log(f"step kernel reads by request size: {rq['128B']:.4g} x 128 B + {rq['64B']:.4g} x 64 B "
                    f"= {rq['bytes'] / 1e6:.1f} MB (FETCH_SIZE scaled: {traffic['read'] / 1e6:.1f} MB)")

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(systems)
        log(f"cpu baseline: {cpu['value']:.4g} bead-steps/s on {cpu['cores']} threads "
            f"(nproc {cpu['nproc']}), {cpu['one_core']['value']:.4g} on one")

    valu = None
    if traffic and traffic.get("valu_insts"):
        # wave64 VALU instructions x 64 lanes per launch / launch time, against
        # 256 CUs x 64 FP32 lanes x 2.4 GHz (no packed FP32)
        lane_ops = traffic["valu_insts"] * 64 / (avg_ms * 1e-3)
        valu = {"achieved": lane_ops / 1e12, "peak": VALU_PEAK / 1e12, "unit": "T lane-ops/s",
                "frac": lane_ops / VALU_PEAK, "insts_per_launch": traffic["valu_insts"]}
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "bead-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {
            "workload": "batched replicas: 512 chains, N log-uniform [200,5000], T uniform [0.5,0.9], "
                        "distinct Philox keys, lists 32/16 A every 50/10 steps, one resident batch",
            "replicas": len(systems), "beads": beads, "lanes_per_bead": lanes, "step_blocks": grid,
            "step_tiles": tiles, "parallelism": "replica batch on 1 GPU",
            "autotune": tuned, "autotune_served_mix": tuned_served,
        },
        "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK, "traffic": traffic["bytes"] if traffic else None,
                     "kernel": f"k_step<{lanes},0>", "kernel_ms": avg_ms,
                     "algorithmic_bytes_per_launch": step_bytes,
                     "measured_copy_GBps": copy_gbs, "frac_of_measured_copy": achieved / 1e9 / copy_gbs,
                     "traffic_detail": traffic, "valu": valu},
        "schedule_roofline": {"bytes_per_bead_step": bytes_per_bead_step,
                              "frac": value * bytes_per_bead_step / HBM_PEAK},
        "plain": {"value": plain_value, "steps": args.plain_steps,
                  "path": "one system per launch, no graphs, same kernels"} if plain_value else None,
        "batched_vs_plain": value / plain_value if plain_value else None,
        "served": served,
        "served_c4": served_c4,
        "assembly": assembly,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
