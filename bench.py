#!/usr/bin/env python3
"""Throughput of the SOP overdamped hot path on one MI355X.

One bench step = one overdamped step (lists rebuilt on schedule, fused
force + Philox + move) of every replica of the batched workload:
512 independent chains, N log-uniform in [200, 5000], T uniform in [0.5, 0.9],
distinct Philox keys, 32 A / 16 A neighbor / pair lists every 50 / 10 steps
(the batched-replica configuration).  Inputs are generated on the host and are
resident in HBM before the timed region.

Printed on stdout: one JSON line.  Also measured in the same run:
  plain       the same replicas, one system per launch, no graphs
  roofline    HBM roofline of the fused step kernel (algorithmic bytes per
              launch / average launch time from HIP events)
  served      continuous-batching request loop (p50/p99 job latency)
  cpu_baseline FP32 port of the reference's host method on the host cores
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


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


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


def cpu_baseline(systems, steps=3000, warm=20, max_threads=16, every=8):
    from concurrent.futures import ThreadPoolExecutor

    import sop_amd as S
    from oracle import oracle as O
    sample = systems[::every]
    ports = [O.Port(s) for s in sample]
    threads = max(1, min(max_threads, os.cpu_count() or 1, len(ports)))

    def go(k, st0, nst):
        s = sample[k]
        a5, var = S.params_derived(s.params)
        ports[k].run(st0, nst, s.params.seed, a5, var, s.params.n_nl, s.params.n_pl)

    with ThreadPoolExecutor(threads) as ex:
        list(ex.map(lambda k: go(k, 0, warm), range(len(ports))))
        t0 = time.perf_counter()
        list(ex.map(lambda k: go(k, warm, steps), range(len(ports))))
        dt = time.perf_counter() - t0
    beads = sum(s.n for s in sample)
    for p in ports:
        p.close()
    return {"value": beads * steps / dt, "unit": "bead-steps/s", "cores": threads, "kind": "port",
            "sample": f"{len(sample)} replicas of the workload (every {every}th, {beads} beads), {steps} steps after "
                      f"{warm} warm-up, FP32 port of the reference host method (O(N^2) master-list filter, "
                      f"half-list force loops, overdamped step), {threads} threads over replicas",
            "seconds": dt}


# HBM bytes per launch of the step kernel from rocprofv3 --pmc FETCH_SIZE and
# WRITE_SIZE passes over this workload (profiles/k_step_batched_512_replicas.md),
# each scaled by the counter's reading for a device copy of known size
# (FETCH_SIZE x 2.0, WRITE_SIZE x 1.0).  Keyed by (kernel, replicas, beads);
# any other run reports null.
TRAFFIC = {
    ("k_step<4,0>", 512, 737499): 324.49e6 + 18.10e6,
}


def measured_traffic(kernel, replicas, beads):
    return TRAFFIC.get((kernel, replicas, beads))


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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--replicas", type=int, default=512)
    ap.add_argument("--plain-steps", type=int, default=50)
    ap.add_argument("--serve-seconds", type=float, default=20.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--lanes", type=int, default=4)
    ap.add_argument("--lds-beads", type=int, default=0, help="LDS-resident step for systems up to this size "
                    "(0 = engine default, < 0 = off)")
    ap.add_argument("--serve-block", type=int, default=0,
                    help="steps per served graph replay (0: the jobs' neighbor-list interval)")
    ap.add_argument("--serve-lanes", type=int, default=0,
                    help="lanes per bead of the served leg's engine (0: the batched leg's, after autotune)")
    ap.add_argument("--pl-cap", type=int, default=0, help="per-bead pair-list capacity bound (0: the engine's)")
    ap.add_argument("--pl-shell", type=float, default=None,
                    help="pair-list shell width in A (default: the engine's, 3 A; < 0 = none)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="BYTES",
                    help="only run a device-to-device copy of BYTES (counter calibration for profiles)")
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

    eng = S.Engine(0, lanes_per_bead=args.lanes, lds_beads_max=args.lds_beads, pl_shell=args.pl_shell,
                   pl_cap_max=args.pl_cap)
    t = time.perf_counter()
    ids = [eng.add(s) for s in systems]
    log(f"resident on device in {time.perf_counter() - t:.1f}s")
    lanes, grid = args.lanes, 0
    if not args.no_autotune:
        lanes, grid = eng.autotune(ids)
        log(f"autotune: {lanes} lanes/bead, grid {grid}")

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
    traffic = measured_traffic(f"k_step<{lanes},0>", len(systems), beads)

    served = None
    if args.serve_seconds > 0:
        try:
            from sop_amd import serving
            served = serving.bench_served(args.serve_seconds, log, lanes_per_bead=args.serve_lanes or lanes,
                                          block_steps=args.serve_block)
        except (ImportError, NotImplementedError) as ex:
            log(f"served leg skipped: {ex}")

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(systems)
        log(f"cpu baseline: {cpu['value']:.4g} bead-steps/s on {cpu['cores']} threads")

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
            "replicas": len(systems), "beads": beads, "lanes_per_bead": lanes, "grid_blocks": grid,
            "parallelism": "replica batch on 1 GPU",
        },
        "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK, "traffic": traffic,
                     "kernel": f"k_step<{lanes},0>", "kernel_ms": avg_ms,
                     "algorithmic_bytes_per_launch": step_bytes},
        "schedule_roofline": {"bytes_per_bead_step": bytes_per_bead_step,
                              "frac": value * bytes_per_bead_step / HBM_PEAK},
        "plain": {"value": plain_value, "steps": args.plain_steps,
                  "path": "one system per launch, no graphs, same kernels"} if plain_value else None,
        "batched_vs_plain": value / plain_value if plain_value else None,
        "served": served,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
