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
