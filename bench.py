#!/usr/bin/env python3
"""Benchmark of the MI355X halloc allocator (one JSON line on stdout).

Workload (the headline metric, 16 B malloc+free pairs/s): 2^20 lanes in
blocks of 256, a 512 MiB pool with the default options; one step is one
launch in which every lane does hamalloc(16), writes a 16 B pattern, reads it
back and calls hafree (tst/perf/priv-throughput/priv-throughput.cu:14-27 with
the write/verify of tst/corr/test/test.cu:84-95).  W untimed steps, then K
timed steps; the device is synchronized before and after the timed region.

Also reported:
  roofline         dominant kernel (the plain workload kernel) against HBM:
                   algorithmic bytes per pair = 4.05 reference-scheme atomics
                   x 8 B + 32 B payload = 64.4 B; traffic from the committed
                   rocprofv3 counter summary when present
  atomic_roofline  the same kernel against the measured atomic rate of the
                   chip (ha_probe_atomics, distinct cache lines)
  cpu_baseline     the host model of the scheme (oracle/), multithreaded, on a
                   bounded sample of the same workload, and glibc malloc/free
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

NTHREADS = 1 << 20
BS = 256
PAIR_ALGO_BYTES = 4.05 * 8 + 32.0      # reference-scheme atomics x 8 B + payload write + read
PAIR_ALGO_ATOMICS = 4.05
HBM_PEAK_GBS = 8000.0


def cpu_baseline(budget_s: float):
    """Host model (oracle) and glibc on the host cores; bounded sample."""
    from oracle import model as M
    M.build()
    cores = min(os.cpu_count() or 1, 16)
    opts = M.make_opts(nheads=8)
    # calibrate on a small sample, then size the sample to the budget
    r0, _ = M.bench_plain(opts, cores, NTHREADS, 1)
    iters = max(1, int(budget_s * 0.5 * r0 / NTHREADS))
    t0 = time.time()
    rate, bad = M.bench_plain(opts, cores, NTHREADS, iters)
    t_model = time.time() - t0
    g0, _ = M.bench_glibc(cores, NTHREADS, 1)
    giters = max(1, int(budget_s * 0.5 * g0 / NTHREADS))
    grate, gbad = M.bench_glibc(cores, NTHREADS, giters)
    return {
        "value": rate, "unit": "pairs/s", "cores": cores, "kind": "port",
        "sample": f"host model, 2^20 lanes x {iters} passes of malloc(16)/write/verify/free "
                  f"in 64-lane waves over {cores} threads ({t_model:.1f} s); violations {bad}",
        "glibc": {"value": grate, "unit": "pairs/s", "cores": cores,
                  "sample": f"2^20 lanes x {giters} passes of malloc(16)/write/verify/free; violations {gbad}"},
    }


def traffic_from_profiles():
    p = os.path.join(ROOT, "profiles", "plain_16B_counters.json")
    if not os.path.exists(p):
        return None
    try:
        return json.load(open(p)).get("hbm_bytes_per_launch")
    except (OSError, ValueError):
        return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--nheads", type=int, default=8)
    ap.add_argument("--cpu-budget", type=float, default=12.0, help="seconds for the CPU baseline sample")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-probe", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}): the allocator keeps per-device "
              "state and this build measures one device's share", file=sys.stderr)
        sys.exit(2)

    import halloc_amd as H
    import numpy as np

    probe = copy_gbps = None
    if not args.no_probe:
        probe = H.Halloc.probe_atomics()
        copy_gbps = H.Halloc.probe_copy(1 << 30, 5)
    H.Halloc.init(H.HallocOpts(memory=512 << 20, nheads=args.nheads))
    r = H.Halloc.run_plain(nthreads=NTHREADS, bs=BS, size=16, iters=1, warmup=args.warmup,
                           launches=args.steps)
    d = H.Halloc.dump()
    g = H.Halloc.geometry()
    leaked = int(((d[:g["nsbs"]] >> np.uint64(32)) != 0).sum())
    H.Halloc.shutdown()
    violations = r["errors"] + r["nulls"] + leaked

    pairs = NTHREADS * args.steps
    wall_s = r["wall_ms"] / 1e3
    kern_s = r["total_ms"] / 1e3
    value = pairs / wall_s
    avg_launch_s = kern_s / args.steps
    achieved = NTHREADS * PAIR_ALGO_BYTES / avg_launch_s / 1e9
    traffic = traffic_from_profiles()
    out = {
        "metric": "16B malloc+free pairs/s on 1 MI355X, batched vs plain; p99 serve latency",
        "value": value,
        "unit": "pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r["wall_ms"] / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic",
        "config": {"workload": "plain 16 B: 2^20 lanes (256/block), hamalloc(16) + 16 B write/verify + "
                               "hafree per lane per step; 512 MiB pool, default halloc_opts_t",
                   "pool_bytes": 512 << 20, "lanes": NTHREADS, "block": BS, "nheads": args.nheads},
        "kernel_ms_per_step": r["total_ms"] / args.steps,
        "launch_ms": [round(float(x), 4) for x in r["ms"]],
        "violations": violations,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic},
    }
    if copy_gbps is not None:
        out["roofline"]["copy_probe_gbps"] = copy_gbps
    if probe is not None:
        ach = NTHREADS * PAIR_ALGO_ATOMICS / avg_launch_s
        out["atomic_roofline"] = {"achieved": ach, "peak": float(probe[0]), "unit": "atomics/s",
                                  "frac": ach / float(probe[0]),
                                  "note": "reference-scheme atomics (4.05/pair) per second vs the measured "
                                          "rate of 64-bit atomics on distinct cache lines",
                                  "probe": {"lane_random": float(probe[0]), "wave_run": float(probe[1]),
                                            "one_word": float(probe[2])}}
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(args.cpu_budget)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
