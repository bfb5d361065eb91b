"""Autotuner for the batched persistent graph-build workload.

Coordinate descent over the allocator options a batched workload can choose
(slab size, busy/roomy/sparse thresholds, probe stride, head sets) and the
workload's own launch shape (blocks per wave call of hamalloc_batch, resident
waves).  Each candidate runs the batched hipGraph form on the device; a
candidate that returns NULL or fails its checks is rejected.  The result is a
JSON file the benchmark and callers load with `load()`.

    python -m halloc_amd.autotune --jobs 1024 --out halloc_amd/tuned/graphs_mi355x.json
"""
from __future__ import annotations

import argparse
import json
import math
import os
import time

import numpy as np

from halloc_amd import Halloc, HallocError, HallocOpts
from halloc_amd.workloads import graph_sizes

TUNED = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuned", "graphs_mi355x.json")

SPACE = {
    "batch": [4, 8, 16],
    "waves": [1024, 2048, 4096, 8192],
    "sb_sz_sh": [21, 22, 23, 24],
    "nheads": [1, 2, 4, 8, 16, 32],
    "busy_fraction": [0.7, 0.835, 0.9],
    "roomy_fraction": [0.4, 0.6],
    "sparse_fraction": [0.005, 0.012, 0.05],
    "hash_step_frac": [0.0, 0.01, 0.05, 0.11, 0.25],
}
DEFAULT = {"batch": 8, "waves": 2048, "sb_sz_sh": 22, "nheads": 16, "busy_fraction": 0.835,
           "roomy_fraction": 0.6, "sparse_fraction": 0.012, "hash_step_frac": 0.0}
OPT_KEYS = ("sb_sz_sh", "nheads", "busy_fraction", "roomy_fraction", "sparse_fraction", "hash_step_frac")


def opts_of(cfg: dict, memory: int) -> HallocOpts:
    return HallocOpts(memory=memory, **{k: cfg[k] for k in OPT_KEYS})


def valid(cfg: dict) -> bool:
    return cfg["sparse_fraction"] <= cfg["roomy_fraction"] <= cfg["busy_fraction"]


def evaluate(cfg: dict, nnodes: np.ndarray, memory: int, repeats: int = 2) -> float:
    """Best batched-form time (ms) of the graph jobs under cfg; inf if rejected."""
    if not valid(cfg):
        return math.inf
    try:
        Halloc.init(opts_of(cfg, memory))  # rejects options with no size table
    except HallocError:
        return math.inf
    try:
        r = Halloc.run_graphs(nnodes, batch=cfg["batch"], waves=cfg["waves"], run_plain=False,
                              repeats=repeats)
    except HallocError:
        return math.inf
    finally:
        Halloc.shutdown()
    if r["nulls"] or r["errors"]:
        return math.inf
    return r["batch_ms"]


def tune(nnodes: np.ndarray, memory: int = 4 << 30, sweeps: int = 2, start: dict | None = None, log=print,
         space: dict | None = None):
    """Coordinate descent from `start` (default options) over `space` (default
    SPACE); returns (best cfg, best ms, history)."""
    cur = dict(start or DEFAULT)
    best = evaluate(cur, nnodes, memory)
    hist = [(dict(cur), best)]
    log(f"start {cur} -> {best:.3f} ms")
    for sweep in range(sweeps):
        improved = False
        for key, values in (space or SPACE).items():
            for v in values:
                if v == cur[key]:
                    continue
                cand = dict(cur, **{key: v})
                t = evaluate(cand, nnodes, memory)
                hist.append((cand, t))
                if t < best:
                    best, cur, improved = t, cand, True
                    log(f"sweep {sweep} {key}={v} -> {t:.3f} ms (new best)")
        if not improved:
            break
    return cur, best, hist


def load(path: str = TUNED) -> dict:
    """The tuned configuration, or the defaults if no tuning file exists."""
    if os.path.exists(path):
        with open(path) as f:
            return dict(DEFAULT, **json.load(f)["best"])
    return dict(DEFAULT)


def main():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--jobs", type=int, default=1024)
    ap.add_argument("--memory", type=int, default=4 << 30)
    ap.add_argument("--sweeps", type=int, default=2)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--out", default=TUNED)
    args = ap.parse_args()
    nnodes = graph_sizes(np.random.default_rng(args.seed), args.jobs)
    t0 = time.time()
    best, ms, hist = tune(nnodes, args.memory, args.sweeps, log=lambda m: print(m, flush=True))
    base = hist[0][1]
    out = {"workload": f"{args.jobs} random-graph builds, n_i log-uniform on [1k, 64k], seed {args.seed}, "
                       f"{args.memory >> 20} MiB pool, batched hipGraph form",
           "best": best, "best_ms": ms, "default_ms": base, "evaluations": len(hist),
           "seconds": round(time.time() - t0, 1),
           "history": [{"cfg": c, "ms": (None if math.isinf(t) else t)} for c, t in hist]}
    os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
    with open(args.out, "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps({"best": best, "best_ms": ms, "default_ms": base}))


if __name__ == "__main__":
    main()
