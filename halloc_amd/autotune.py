"""Autotuner for the random-graph build workload, batched and plain forms.

Coordinate descent over the allocator options a workload can choose (slab
size, busy/roomy/sparse thresholds, probe stride, head sets) and, for the
batched persistent hipGraph form, its launch shape (blocks per wave call of
hamalloc_batch, resident waves).  The plain per-call form is tuned on its own
over the allocator options, so that each form is compared at its best.  Each
candidate runs on the device; one that returns NULL or fails its checks is
rejected.  The result is a JSON file the benchmark and callers load with
`load()` (batched) and `load_plain()` (plain).

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
# the library's default options and the search's starting launch shape
DEFAULT = {"batch": 8, "waves": 2048, "sb_sz_sh": 22, "nheads": 32, "busy_fraction": 0.835,
           "roomy_fraction": 0.6, "sparse_fraction": 0.012, "hash_step_frac": 0.0}
OPT_KEYS = ("sb_sz_sh", "nheads", "busy_fraction", "roomy_fraction", "sparse_fraction", "hash_step_frac")
SPACE_PLAIN = {k: v for k, v in SPACE.items() if k in OPT_KEYS}  # the plain form has no launch shape


def opts_of(cfg: dict, memory: int) -> HallocOpts:
    return HallocOpts(memory=memory, **{k: cfg[k] for k in OPT_KEYS})


def valid(cfg: dict) -> bool:
    return cfg["sparse_fraction"] <= cfg["roomy_fraction"] <= cfg["busy_fraction"]


def evaluate(cfg: dict, nnodes: np.ndarray, memory: int, repeats: int = 2, form: str = "batched") -> float:
    """Best time (ms) of the graph jobs in one form ("batched" or "plain")
    under cfg; inf if rejected."""
    if not valid(cfg):
        return math.inf
    try:
        Halloc.init(opts_of(cfg, memory))  # rejects options with no size table
    except HallocError:
        return math.inf
    try:
        r = Halloc.run_graphs(nnodes, batch=cfg["batch"], waves=cfg["waves"], run_plain=form == "plain",
                              run_batch=form == "batched", repeats=repeats)
    except HallocError:
        return math.inf
    finally:
        Halloc.shutdown()
    if r["nulls"] or r["errors"]:
        return math.inf
    return r["batch_ms"] if form == "batched" else r["plain_ms"]


def tune(nnodes: np.ndarray, memory: int = 4 << 30, sweeps: int = 2, start: dict | None = None, log=print,
         space: dict | None = None, form: str = "batched"):
    """Coordinate descent for one form from `start` (default options) over
    `space` (default SPACE, or SPACE_PLAIN for the plain form); returns (best
    cfg, best ms, history)."""
    if space is None:
        space = SPACE if form == "batched" else SPACE_PLAIN
    cur = dict(start or DEFAULT)
    best = evaluate(cur, nnodes, memory, form=form)
    hist = [(dict(cur), best)]
    log(f"start {cur} -> {best:.3f} ms")
    for sweep in range(sweeps):
        improved = False
        for key, values in space.items():
            for v in values:
                if v == cur[key]:
                    continue
                cand = dict(cur, **{key: v})
                t = evaluate(cand, nnodes, memory, form=form)
                hist.append((cand, t))
                if t < best:
                    best, cur, improved = t, cand, True
                    log(f"sweep {sweep} {key}={v} -> {t:.3f} ms (new best)")
        if not improved:
            break
    return cur, best, hist


def load(path: str = TUNED, key: str = "best") -> dict:
    """The tuned configuration of the batched form, or the defaults if no
    tuning file exists."""
    if os.path.exists(path):
        with open(path) as f:
            d = json.load(f)
        if key in d:
            return dict(DEFAULT, **d[key])
    return dict(DEFAULT)


def load_plain(path: str = TUNED) -> dict:
    """The tuned allocator options of the plain form (defaults if absent)."""
    return load(path, "best_plain")


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
    say = lambda m: print(m, flush=True)
    best, ms, hist = tune(nnodes, args.memory, args.sweeps, log=say)
    pbest, pms, phist = tune(nnodes, args.memory, args.sweeps, log=say, form="plain")
    base, pbase = hist[0][1], phist[0][1]
    enc = lambda h: [{"cfg": c, "ms": (None if math.isinf(t) else t)} for c, t in h]
    out = {"workload": f"{args.jobs} random-graph builds, n_i log-uniform on [1k, 64k], seed {args.seed}, "
                       f"{args.memory >> 20} MiB pool; batched hipGraph form and plain per-call form",
           "best": best, "best_ms": ms, "default_ms": base,
           "best_plain": {k: pbest[k] for k in OPT_KEYS}, "best_plain_ms": pms, "default_plain_ms": pbase,
           "evaluations": len(hist) + len(phist), "seconds": round(time.time() - t0, 1),
           "history": enc(hist), "history_plain": enc(phist)}
    os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
    with open(args.out, "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps({"best": best, "best_ms": ms, "default_ms": base, "best_plain": out["best_plain"],
                      "best_plain_ms": pms, "default_plain_ms": pbase}))


if __name__ == "__main__":
    main()
