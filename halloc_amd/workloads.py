"""Synthetic workload generators for the drivers in libhalloc_amd.so.

Each returns plain numpy arrays; the drivers take them through the C ABI.
"""
from __future__ import annotations

import numpy as np

TOKEN_BYTES = 16   # one decoded token
PAGE_BYTES = 256   # one hamalloc page: 16 tokens
MAX_TOKENS = 32768


def serving_requests(nreq: int, rate_per_ms: float = 2000.0, median: float = 512.0, sigma: float = 1.0,
                     cap: int = MAX_TOKENS, seed: int = 0):
    """Poisson arrivals at rate_per_ms requests per ms; lengths log-normal with
    the given median and sigma, clipped to [1, cap] tokens; a random tag per
    request.  Returns (arrival_ns uint64 sorted, length uint32, tag uint32)."""
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1e6 / rate_per_ms, size=nreq)
    arrival = np.floor(np.cumsum(gaps) - gaps[0]).astype(np.uint64)
    length = np.clip(np.rint(rng.lognormal(np.log(median), sigma, size=nreq)), 1, cap).astype(np.uint32)
    tag = rng.integers(0, 1 << 32, size=nreq, dtype=np.uint64).astype(np.uint32)
    return arrival, length, tag


def pages_for(length) -> np.ndarray:
    """256 B pages a sequence of `length` tokens occupies."""
    ln = np.asarray(length, dtype=np.int64)
    return (ln + PAGE_BYTES // TOKEN_BYTES - 1) // (PAGE_BYTES // TOKEN_BYTES)


def sweep_sizes(rng, n, per_max=8, large_frac=0.01, lo=16, hi=3072, large_hi=65536):
    """(per_max, n) request sizes: lane i makes U{1..per_max} requests, each
    log-uniform on [lo, hi] bytes, or with probability large_frac uniform on
    (hi, large_hi]; 0 marks no request."""
    counts = rng.integers(1, per_max + 1, n)
    logs = rng.uniform(np.log(lo), np.log(hi), (per_max, n))
    sizes = np.floor(np.exp(logs)).astype(np.int64).clip(lo, hi)
    large = rng.random((per_max, n)) < large_frac
    sizes[large] = rng.integers(hi + 1, large_hi + 1, int(large.sum()))
    sizes[np.arange(per_max)[:, None] >= counts[None, :]] = 0
    return sizes.astype(np.uint32)


def graph_sizes(rng, njobs, lo=1024, hi=65536):
    """Node counts of the random-graph jobs: log-uniform on [lo, hi]."""
    return np.floor(np.exp(rng.uniform(np.log(lo), np.log(hi), njobs))).astype(np.uint32)
