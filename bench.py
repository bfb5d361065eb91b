#!/usr/bin/env python3
"""Benchmark of the MI355X halloc allocator (one JSON line on stdout).

Workload (the headline metric, 16 B malloc+free pairs/s): 2^20 lanes in
blocks of 256, a 512 MiB pool with the default options; one step is one
launch in which every lane does hamalloc(16), writes a 16 B pattern, reads it
back and calls hafree (tst/perf/priv-throughput/priv-throughput.cu:14-27 with
the write/verify of tst/corr/test/test.cu:84-95).  W untimed steps, then K
timed steps; the device is synchronized before and after the timed region.
The default W is 10: each launch's tickets sweep an eighth of every head
slab, so the first 8 launches meet the slabs' bit and nibble lines cold and
run about 20% slower than the steady state that follows.

Also reported (each on its own freshly initialized heap):
  batching         4096 random-graph builds (n_i log-uniform on [1k, 64k]
                   nodes; a 16 B header plus a Poisson(8) int32 adjacency
                   array per node), 4 GiB pool: plain per-call kernels vs
                   one persistent hipGraph launch using hamalloc_batch, with
                   the autotuned options (halloc_amd/tuned) when present
  serving          continuous-batching request loop: Poisson arrivals at
                   2000 req/ms, log-normal lengths (median 512, sigma 1, cap
                   32768 tokens of 16 B) growing in 256 B pages, <= 65536
                   live sequences, 16 GiB pool, watermark admission;
                   p50/p99 page-malloc latency and request latency
  phase            the reference's phase-throughput with its experiments'
                   flags: 2^20 threads, U{16..256} B, 5 allocation and 5
                   free phases of 8 chain steps, 512 MiB pool
  sweep            2^19 lanes, U{1..8} requests log-uniform on [16 B, 3 KiB]
                   plus 1% on (3 KiB, 64 KiB], random-order frees, 1 GiB
                   pool; per size class; the large path timed alone
  arena            both arena types (uncached default, cached_arena 1): a
                   caller's kernel making 64 passes of stores, loads and
                   updates over 2^18 blocks of 64 B, and the plain kernel
  roofline        dominant kernel (the plain workload kernel) against HBM:
                   algorithmic bytes per pair = 4.05 reference-scheme atomics
                   x 8 B + 32 B payload = 64.4 B; traffic from the committed
                   rocprofv3 counter summary when present
  atomics          atomics per pair: the reference scheme's 4.05, this
                   build's count from the code, and the measured TCC_ATOMIC
                   count per dispatch from the committed counter summary
  chain            the binding resource: the same grid making only the
                   kernel's dependent atomic round trips per wave
                   (ha_probe_chain), against the kernel's launch time
  cpu_baseline     the host model of the scheme (oracle/), multithreaded over
                   every usable host core, on a bounded sample of the same
                   workload, and glibc malloc/free
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


def host_cpu():
    """(usable cores, CPU model name) of this host."""
    try:
        cores = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        cores = os.cpu_count() or 1
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return cores, model


def cpu_baseline(budget_s: float):
    """Host model (oracle) and glibc on every usable host core; bounded sample."""
    from oracle import model as M
    M.build()
    cores, cpu_model = host_cpu()
    opts = M.make_opts(nheads=16)
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
        "nproc": os.cpu_count(), "cpu_model": cpu_model,
        "sample": f"host model, 2^20 lanes x {iters} passes of malloc(16)/write/verify/free "
                  f"in 64-lane waves over {cores} threads ({t_model:.1f} s); violations {bad}",
        "glibc": {"value": grate, "unit": "pairs/s", "cores": cores,
                  "sample": f"2^20 lanes x {giters} passes of malloc(16)/write/verify/free; violations {gbad}"},
    }


def pct(a, q):
    import numpy as np
    return float(np.percentile(np.asarray(a, dtype=np.float64), q)) if len(a) else None


def bench_batching(H, njobs=4096, memory=4 << 30, seed=0):
    """Both forms under the library defaults, the batched form under its
    autotuned options and the plain form under its own; the headline ratio
    puts the tuned batched form against the fastest plain run."""
    import numpy as np
    from halloc_amd import autotune
    from halloc_amd.workloads import graph_sizes
    cfg = autotune.load()
    pcfg = autotune.load_plain()
    nn = graph_sizes(np.random.default_rng(seed), njobs)
    H.Halloc.init(autotune.opts_of(cfg, memory))
    t = H.Halloc.run_graphs(nn, batch=cfg["batch"], waves=cfg["waves"], repeats=3)
    H.Halloc.shutdown()
    H.Halloc.init(autotune.opts_of(pcfg, memory))
    p = H.Halloc.run_graphs(nn, run_batch=False, repeats=3)
    H.Halloc.shutdown()
    # the library's default allocator options at the tuned launch shape (blocks
    # per call, resident waves: the workload's choice), and at the tuner's
    # starting shape
    d0 = autotune.DEFAULT
    H.Halloc.init(H.HallocOpts(memory=memory))
    d = H.Halloc.run_graphs(nn, batch=cfg["batch"], waves=cfg["waves"], repeats=3)
    d1 = H.Halloc.run_graphs(nn, batch=d0["batch"], waves=d0["waves"], run_plain=False, repeats=3)
    H.Halloc.shutdown()
    pairs = 2 * t["nodes"]  # a header and an adjacency array per node
    plain_best = min(t["plain_ms"], p["plain_ms"], d["plain_ms"])
    return {
        "workload": f"{njobs} random-graph builds, n_i log-uniform on [1k, 64k], seed {seed}, "
                    f"{memory >> 30} GiB pool",
        "nodes": t["nodes"], "pairs": pairs,
        "batched_ms": t["batch_ms"], "plain_ms": plain_best, "plain_launches": t["plain_launches"],
        "batched_pairs_per_s": pairs / (t["batch_ms"] / 1e3),
        "plain_pairs_per_s": pairs / (plain_best / 1e3),
        "batched_vs_plain": plain_best / t["batch_ms"],
        "tuned": cfg, "tuned_plain": {k: pcfg[k] for k in autotune.OPT_KEYS},
        "tuned_from_file": os.path.exists(autotune.TUNED),
        "tuned_opts": {"batched_ms": t["batch_ms"], "plain_ms": t["plain_ms"]},
        "plain_tuned_opts": {"plain_ms": p["plain_ms"]},
        "default_opts": {"batched_ms": d["batch_ms"], "plain_ms": d["plain_ms"],
                         "batched_vs_plain": d["plain_ms"] / d["batch_ms"],
                         "launch_shape": {"batch": cfg["batch"], "waves": cfg["waves"]},
                         "batched_ms_at_start_shape": d1["batch_ms"],
                         "start_shape": {"batch": d0["batch"], "waves": d0["waves"]}},
        "violations": t["nulls"] + t["errors"] + p["nulls"] + p["errors"] + d["nulls"] + d["errors"]
                      + d1["nulls"] + d1["errors"],
    }


def bench_serving(H, nreq=100000, memory=16 << 30, seed=0):
    import numpy as np
    from halloc_amd.workloads import pages_for, serving_requests
    arr, ln, tag = serving_requests(nreq, seed=seed)
    H.Halloc.init(H.HallocOpts(memory=memory))
    g = H.Halloc.geometry()
    wm = g["arena_bytes"] // 16
    r = H.Halloc.run_serve(arr, ln, tag, slots=65536, watermark=wm, sample_every=4)
    H.Halloc.shutdown()
    lat = r["lat_ns"]
    req_ms = (r["done_ns"].astype(np.float64) - arr.astype(np.float64)) / 1e6
    queue_us = (r["first_ns"].astype(np.float64) - arr.astype(np.float64)) / 1e3
    pages = r["pages"]
    return {
        "workload": f"{nreq} requests, Poisson 2000 req/ms, log-normal length (median 512, sigma 1, cap 32768 "
                    f"tokens x 16 B), 256 B pages, 65536 slots, {memory >> 30} GiB pool, watermark {wm >> 20} MiB",
        "completed": r["completed"], "pages": pages, "kernel_ms": r["kernel_ms"],
        "pairs_per_s": pages / (r["kernel_ms"] / 1e3),
        "p50_alloc_ns": pct(lat, 50), "p99_alloc_ns": pct(lat, 99), "alloc_samples": int(len(lat)),
        "p50_request_ms": pct(req_ms, 50), "p99_request_ms": pct(req_ms, 99),
        "p99_admission_wait_us": pct(queue_us, 99),
        "stalls": r["stalls"],
        "violations": int(r["nulls"]) + int(r["bad"].sum()) + int((r["produced"] != ln).sum())
                      + int(r["completed"] != nreq) + int(pages != int(pages_for(ln).sum())),
    }


def bench_phase(H, memory=512 << 20):
    """The phase workload as the reference's phase-throughput runs it (tst/perf/
    phase-throughput with the flags of its experiments, tst/exp/common.pl:
    -f0.95 -F0.05 -e0.91 -g5): 2^20 threads, one block of U{16..256} B each,
    alternating allocation and free phases (10 launches = 5 of each, 8 chain
    steps per launch); f / F / e are the reference's steady-state fractions
    (the share of threads holding a block at the end of an allocation and of
    a free phase, and the share acting per step), which fix the chain's
    probabilities (tst/common/common.cu:314-355).  The heap uses the library
    defaults: 32 head sets, adaptive."""
    from halloc_amd import harness
    o = harness.defaults_for("phase-throughput")
    o.parse(["-n1048576", "-s16", "-S256", "-duniform", "-l1", "-i8", "-t10", "-f0.95", "-F0.05",
             "-e0.91", "-g5", f"-m{memory}"])
    H.Halloc.init(H.HallocOpts(memory=memory))
    r = H.Halloc.run_test(harness.TESTS["phase-throughput"], o.nthreads, o.bs, o.ntries, o.nallocs, o.niters,
                          o.alloc_sz, o.max_alloc_sz, o.max_alloc_sh, harness.DISTRS[o.distr], o.period_mask,
                          o.group_sh, o.probabs, 0)
    H.Halloc.shutdown()
    ms = r["malloc_ms"] + r["free_ms"]
    pairs = (r["mallocs"] + r["frees"]) / 2
    return {"workload": "phase-throughput -n1048576 -s16 -S256 -l1 -i8 -t10 -f0.95 -F0.05 -e0.91 -g5, "
                        f"{memory >> 20} MiB pool, library defaults (32 head sets, adaptive)",
            "chain_probabilities": [[float(x) for x in row] for row in o.probabs],
            "pairs_per_s": pairs / (ms / 1e3), "mallocs": r["mallocs"], "frees": r["frees"],
            "nulls": r["nulls"], "kernel_ms": ms, "violations": r["nulls"] + r["checks_failed"]}


def bench_sweep(H, memory=1 << 30, seed=0, rounds=8):
    """2^19 lanes in `rounds` rounds, so that each round's live set (and its
    large requests on the device heap) fits the pool; per size class, as many
    blocks as fill half the slab heap."""
    import numpy as np
    from halloc_amd.workloads import sweep_sizes
    rng = np.random.default_rng(seed)
    # sizes over all 16 classes: head sets sized to the heap (4 at 1 GiB)
    H.Halloc.init(H.HallocOpts(memory=memory, nheads=0))
    arena = H.Halloc.geometry()["arena_bytes"]
    nheads = H.Halloc.geometry()["nheads"]
    nreq, ms, viol, nulls = 0, 0.0, 0, 0
    nbig, ms_big, nulls_big = 0, 0.0, 0
    for _ in range(rounds):
        sizes = sweep_sizes(rng, (1 << 19) // rounds)
        small = np.where(sizes <= 3072, sizes, 0).astype(np.uint32)
        big = np.where(sizes > 3072, sizes, 0).astype(np.uint32)
        order = rng.permutation(sizes.size).astype(np.uint32)
        r = H.Halloc.run_lists(small, order)
        nreq += int((small != 0).sum())
        ms += r["malloc_ms"] + r["free_ms"]
        viol += r["errors"]
        nulls += r["nulls"]
        b = H.Halloc.run_lists(big, order)
        nbig += int((big != 0).sum())
        ms_big += b["malloc_ms"] + b["free_ms"]
        viol += b["errors"]
        nulls_big += b["nulls"]
    out = {"workload": f"2^19 lanes in {rounds} rounds, U{{1..8}} requests log-uniform on [16 B, 3 KiB] + 1% "
                       f"on (3 KiB, 64 KiB] (device heap, timed apart), random-order frees, "
                       f"{memory >> 30} GiB pool, {nheads} head sets (nheads 0)",
           "requests": nreq, "pairs_per_s": nreq / (ms / 1e3), "nulls": nulls,
           "large_requests": nbig, "large_path_pairs_per_s": nbig / (ms_big / 1e3), "large_path_nulls": nulls_big,
           "violations": viol + nulls, "per_class_pairs_per_s": {}}
    for bsz in (16, 24, 32, 48, 64, 96, 128, 192, 256, 384, 512, 768, 1024, 1536, 2048, 3072):
        n = min(1 << 19, arena // 2 // bsz // 4)
        sz = np.full((4, n), bsz, dtype=np.uint32)
        c = H.Halloc.run_lists(sz, None)
        out["per_class_pairs_per_s"][str(bsz)] = sz.size / ((c["malloc_ms"] + c["free_ms"]) / 1e3)
        out["violations"] += c["nulls"] + c["errors"]
    H.Halloc.shutdown()
    return out


def bench_arena(H, memory=512 << 20, n=1 << 18, words=8, iters=64, reps=5):
    """What the arena's memory type costs the caller's own kernels, and the
    plain 16 B kernel, in both arena types: n lanes each own a block of
    8 * words bytes (16 MiB in all, so the blocks fit the eight L2s) and make
    `iters` passes of stores, loads or updates over its words
    (ha_run_touch)."""
    out = {"workload": f"{n} lanes x {8 * words} B blocks, {iters} passes of stores / loads / updates per "
                       f"launch, mean of {reps} launches; plain 16 B kernel, 2^20 lanes, 10 launches; "
                       f"{memory >> 20} MiB pool, library defaults"}
    for name, cached in (("uncached", 0), ("cached", 1)):
        H.Halloc.init(H.HallocOpts(memory=memory, cached_arena=cached))
        r = H.Halloc.run_touch(n, words, iters, reps)
        p = H.Halloc.run_plain(warmup=10, launches=10)
        H.Halloc.shutdown()
        nbytes = n * words * 8 * iters
        out[name] = {"store_ms": r["store_ms"], "load_ms": r["load_ms"], "update_ms": r["update_ms"],
                     "store_gbps": nbytes / (r["store_ms"] * 1e6), "load_gbps": nbytes / (r["load_ms"] * 1e6),
                     "update_gbps": 2 * nbytes / (r["update_ms"] * 1e6),
                     "plain_16B_us": p["total_ms"] * 1e3 / 10,
                     "violations": r["nulls"] + p["errors"] + p["nulls"]}
    return out


def plain_counters():
    """The committed counter summary of the plain kernel (profiles/), or {}."""
    p = os.path.join(ROOT, "profiles", "plain_16B_counters.json")
    try:
        return json.load(open(p))
    except (OSError, ValueError):
        return {}


# Atomics per 64-lane wave of the plain kernel, counted from the code.  Every
# wave is one 16 B class group of 64 lanes, so the group's tickets stay
# multiples of 64 and its blocks fill one aligned 64-block run: malloc takes a
# ticket (1), reserves the run with one OR on its bit word (1), adds to the
# slab counter (1) and records the block lengths with one OR per nibble word
# (4); free clears the nibbles (4) and the bits (1) the same way and subtracts
# from the counter (1).
CODE_ATOMICS_PER_WAVE = 1 + 1 + 1 + 4 + 4 + 1 + 1
# dependent round trips of one iteration of a wave: head load with the
# ticket, probe, counter (malloc); descriptor, nibbles, counter (free); the
# payload write and read-back overlap the rest
CHAIN_ROUNDS = 6


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    # 100 launches of about 38 us: the synchronization that ends the timed
    # region (about 14 us) weighs on the per-step time as 1/steps
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    # head sets of the plain 16 B kernel: 32, the library default and the plain
    # form's tuned count (one size class on 96 slabs leaves room for them)
    ap.add_argument("--nheads", type=int, default=32)
    ap.add_argument("--cpu-budget", type=float, default=12.0, help="seconds for the CPU baseline sample")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-probe", action="store_true")
    ap.add_argument("--no-extras", action="store_true", help="only the plain 16 B workload")
    ap.add_argument("--extras", default="batching,serving,phase,sweep,arena",
                    help="comma-separated workloads to run after the plain one")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}): the allocator keeps per-device "
              "state and this build measures one device's share", file=sys.stderr)
        sys.exit(2)

    import halloc_amd as H
    import numpy as np

    probe = copy_gbps = chain = None
    if not args.no_probe:
        probe = H.Halloc.probe_atomics()
        copy_gbps = H.Halloc.probe_copy(1 << 30, 5)
        chain = H.Halloc.probe_chain((0, 1, 2, 4, CHAIN_ROUNDS, 8, 16), nthreads=NTHREADS)
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
    counters = plain_counters()
    traffic = counters.get("hbm_bytes_per_launch")
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
                               f"hafree per lane per step; 512 MiB pool, default halloc_opts_t, "
                               f"{args.nheads} head sets",
                   "pool_bytes": 512 << 20, "lanes": NTHREADS, "block": BS, "nheads": args.nheads},
        "kernel_ms_per_step": r["total_ms"] / args.steps,
        # a second pass after the timed one, an event pair around each launch
        "launch_ms_event_pairs": [round(float(x), 4) for x in r["ms"]],
        "violations": violations,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic},
    }
    if copy_gbps is not None:
        out["roofline"]["copy_probe_gbps"] = copy_gbps
    measured = counters.get("tcc_atomics_per_pair")
    out["atomics"] = {
        "reference_scheme_per_pair": PAIR_ALGO_ATOMICS,
        "this_build_per_pair_by_code": CODE_ATOMICS_PER_WAVE / 64,
        "measured_per_pair": measured,
        "measured_rate": measured * NTHREADS / avg_launch_s if measured else None,
        "source": "TCC_ATOMIC per dispatch / 2^20 pairs (profiles/plain_16B_counters.json)",
    }
    if probe is not None:
        out["atomics"]["probe_rates"] = {"lane_random": float(probe[0]), "wave_run": float(probe[1]),
                                         "one_word": float(probe[2])}
    if chain is not None:
        # the bound of a chain of dependent atomics per wave: the same grid,
        # every wave making CHAIN_ROUNDS dependent round trips and nothing else
        t = chain[CHAIN_ROUNDS] * 1e-3
        out["chain"] = {"rounds_per_iteration": CHAIN_ROUNDS, "probe_launch_s": t,
                        "round_trip_s": (chain[16] - chain[0]) * 1e-3 / 16,
                        "kernel_launch_s": avg_launch_s, "frac": t / avg_launch_s,
                        "probe_ms_by_rounds": {str(k): v for k, v in chain.items()}}
    extras = [] if args.no_extras else [e for e in args.extras.split(",") if e]
    for e in extras:
        if e not in ("batching", "serving", "phase", "sweep", "arena"):
            print(f"unknown workload {e!r}", file=sys.stderr)
            sys.exit(2)
    if "batching" in extras:
        out["batching"] = bench_batching(H)
        out["batched_vs_plain"] = out["batching"]["batched_vs_plain"]
    if "serving" in extras:
        out["serving"] = bench_serving(H)
        out["p99_serve_alloc_ns"] = out["serving"]["p99_alloc_ns"]
    if "phase" in extras:
        out["phase"] = bench_phase(H)
    if "sweep" in extras:
        out["sweep"] = bench_sweep(H)
    if "arena" in extras:
        out["arena"] = bench_arena(H)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(args.cpu_budget)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
