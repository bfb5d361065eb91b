"""ctypes wrapper of the host model of the halloc scheme (TEST INFRASTRUCTURE ONLY).

Only tests/, ``__graft_entry__.smoke()`` and the CPU-baseline leg of bench.py
import this module.  It is the checker, never the thing measured or shipped:
the product path (``halloc_amd``) never imports it.

The model restates canonizer/halloc's allocation path (src/halloc.cu,
src/slab.cuh, src/slab.h, src/sbset.*, src/size-info.h, src/utils.cu) and the
large region that replaces the reference's delegation to the CUDA device heap;
see oracle/halloc_model.hpp for the function-by-function citations.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "build", "libhalloc_model.so")

NSIZES = 16
SZ_NONE = 0xFFFFFFFF
LARGE_SB = 0xFFFFFFFE   # slab field of a large block's record (class field: 16 + its class)
REC_NULL = 0xFFFFFFFF
NSETS = 1 + 2 * NSIZES + 4
MAXB = 16
SIZE_TABLE_FIELDS = ("nchunks", "chunk_sz", "chunk_id", "ncib", "sparse", "hash_step",
                     "roomy", "busy", "block_sz", "nblocks", "tmod", "chunk_bytes")


class HmOpts(ctypes.Structure):
    _fields_ = [("memory", ctypes.c_uint64),
                ("halloc_fraction", ctypes.c_double),
                ("busy_fraction", ctypes.c_double),
                ("roomy_fraction", ctypes.c_double),
                ("sparse_fraction", ctypes.c_double),
                ("sb_sz_sh", ctypes.c_int32),
                ("nheads", ctypes.c_int32),
                ("hash_mode", ctypes.c_int32),
                ("fixed_sets", ctypes.c_int32),
                ("hash_step_frac", ctypes.c_double)]


def make_opts(memory=512 << 20, halloc_fraction=0.75, busy_fraction=0.835, roomy_fraction=0.6,
              sparse_fraction=0.012, sb_sz_sh=22, nheads=1, hash_mode=1, hash_step_frac=0.0,
              fixed_sets=0) -> HmOpts:
    """Defaults follow src/halloc.h:30-32; fixed_sets 0 = adaptive head sets."""
    return HmOpts(memory, halloc_fraction, busy_fraction, roomy_fraction, sparse_fraction,
                  sb_sz_sh, nheads, hash_mode, fixed_sets, hash_step_frac)


_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        u32, u64, i32, i64, vp = (ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int32,
                                  ctypes.c_int64, ctypes.c_void_p)
        P = ctypes.POINTER
        for name, res, args in [
            ("hm_size_id", u32, [u32]),
            ("hm_chunk_val", u32, [u32]),
            ("hm_chunk_mul", u32, [u32, u32]),
            ("hm_chunk_div", u32, [u32, u32]),
            ("hm_max_prime_below", u32, [u32, u32]),
            ("hm_ichunk_init", u32, [u32, u32, u32, i32]),
            ("hm_probe_next", u32, [u32, u32, u32, u32, i32]),
            ("hm_size_table", ctypes.c_int, [P(HmOpts), vp]),
            ("hm_create", vp, [P(HmOpts)]),
            ("hm_destroy", None, [vp]),
            ("hm_nsbs", u32, [vp]),
            ("hm_crowded", u64, [vp]),
            ("hm_replay", ctypes.c_int, [vp, vp, vp, ctypes.c_int, ctypes.c_int, vp, vp]),
            ("hm_replay_batch", ctypes.c_int, [vp, vp, vp, ctypes.c_int, ctypes.c_int, vp, vp]),
            ("hm_wave", ctypes.c_int, [vp, vp, vp, ctypes.c_int, vp, vp]),
            ("hm_wave_raw", ctypes.c_int, [vp, u32, vp, ctypes.c_int, vp]),
            ("hm_dump_words", u64, [vp]),
            ("hm_dump", ctypes.c_int, [vp, vp]),
            ("hm_check_dump", i64, [P(HmOpts), vp, vp]),
            ("hm_malloc_threads", ctypes.c_int, [vp, ctypes.c_int, vp, ctypes.c_int, vp]),
            ("hm_free_threads", ctypes.c_int, [vp, ctypes.c_int, vp, ctypes.c_int]),
            ("hm_bench_plain", ctypes.c_double, [P(HmOpts), ctypes.c_int, ctypes.c_int, ctypes.c_int, P(u64)]),
            ("hm_bench_glibc", ctypes.c_double, [ctypes.c_int, ctypes.c_int, ctypes.c_int, P(u64)]),
        ]:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def size_table(opts: HmOpts) -> np.ndarray:
    out = np.zeros((NSIZES, len(SIZE_TABLE_FIELDS)), dtype=np.uint32)
    if lib().hm_size_table(ctypes.byref(opts), out.ctypes.data) != 0:
        raise ValueError("size table cannot be built for these options")
    return out


class ModelHeap:
    """One host model heap; replay() drives it one wave call at a time."""

    def __init__(self, opts: HmOpts):
        self.opts = opts
        self.h = lib().hm_create(ctypes.byref(opts))
        if not self.h:
            raise ValueError("invalid model options")
        self.nsbs = lib().hm_nsbs(self.h)

    def close(self):
        if self.h:
            lib().hm_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def replay(self, ops: np.ndarray, iheads: np.ndarray | None = None):
        """ops: (nsteps, lanes, 2) uint32 [kind, arg]; returns (recs (nsteps, lanes, 4), ptrs)."""
        ops = np.ascontiguousarray(ops, dtype=np.uint32)
        nsteps, lanes = ops.shape[0], ops.shape[1]
        recs = np.zeros((nsteps, lanes, 4), dtype=np.uint32)
        ptrs = np.zeros((nsteps, lanes), dtype=np.uint64)
        ih = None
        if iheads is not None:
            ih = np.ascontiguousarray(iheads, dtype=np.uint32)
        rc = lib().hm_replay(self.h, ops.ctypes.data, ih.ctypes.data if ih is not None else None,
                             nsteps, lanes, recs.ctypes.data, ptrs.ctypes.data)
        if rc != 0:
            raise ValueError(f"hm_replay failed: {rc}")
        return recs, ptrs

    def replay_batch(self, ops: np.ndarray, iheads: np.ndarray | None = None):
        """ops: (nsteps, lanes, 4) uint32 [kind, arg, count, 0] (kind 1 hamalloc_batch, 2 hafree_batch of
        op arg's blocks, 3 the same keeping the pointers); returns recs (nsteps, lanes, MAXB, 4)."""
        ops = np.ascontiguousarray(ops, dtype=np.uint32)
        nsteps, lanes = ops.shape[0], ops.shape[1]
        recs = np.zeros((nsteps, lanes, MAXB, 4), dtype=np.uint32)
        ptrs = np.zeros((nsteps, lanes, MAXB), dtype=np.uint64)
        ih = None if iheads is None else np.ascontiguousarray(iheads, dtype=np.uint32)
        rc = lib().hm_replay_batch(self.h, ops.ctypes.data, None if ih is None else ih.ctypes.data,
                                   nsteps, lanes, recs.ctypes.data, ptrs.ctypes.data)
        if rc != 0:
            raise ValueError(f"hm_replay_batch failed: {rc}")
        return recs

    def wave(self, kinds, args):
        """One wave step with explicit pointers: lanes of kind 1 malloc args[l]
        bytes, then lanes of kind 2 free the pointer args[l] (a value returned
        here earlier; 0 is NULL).  Returns (recs (lanes, 4), ptrs (lanes,))."""
        k = np.ascontiguousarray(kinds, dtype=np.uint32)
        a = np.ascontiguousarray(args, dtype=np.uint64)
        lanes = len(k)
        recs = np.zeros((lanes, 4), dtype=np.uint32)
        ptrs = np.zeros(lanes, dtype=np.uint64)
        if lib().hm_wave(self.h, k.ctypes.data, a.ctypes.data, lanes, recs.ctypes.data, ptrs.ctypes.data) != 0:
            raise ValueError("hm_wave failed")
        return recs, ptrs

    def crowded(self) -> int:
        """Probe sequences so far that marked their head crowded."""
        return int(lib().hm_crowded(self.h))

    def wave_raw(self, pair: int, sizes):
        """One wave of mallocs on head-set pair `pair` unfolded (a wave with a
        stale spread); returns arena offset + 1 per lane (0 = NULL)."""
        sz = np.ascontiguousarray(sizes, dtype=np.uint32)
        ptrs = np.zeros(len(sz), dtype=np.uint64)
        if lib().hm_wave_raw(self.h, pair, sz.ctypes.data, len(sz), ptrs.ctypes.data) != 0:
            raise ValueError("hm_wave_raw failed")
        return ptrs

    def malloc_threads(self, sizes, nthreads: int = 8) -> np.ndarray:
        """Mallocs of `sizes` in 64-lane waves raced over nthreads threads;
        returns arena offset + 1 per request (0 = NULL)."""
        sz = np.ascontiguousarray(sizes, dtype=np.uint32)
        ptrs = np.zeros(len(sz), dtype=np.uint64)
        if lib().hm_malloc_threads(self.h, nthreads, sz.ctypes.data, len(sz), ptrs.ctypes.data) != 0:
            raise ValueError("hm_malloc_threads failed")
        return ptrs

    def free_threads(self, ptrs, nthreads: int = 8) -> None:
        p = np.ascontiguousarray(ptrs, dtype=np.uint64)
        if lib().hm_free_threads(self.h, nthreads, p.ctypes.data, len(p)) != 0:
            raise ValueError("hm_free_threads failed")

    def dump(self) -> np.ndarray:
        n = lib().hm_dump_words(self.h)
        out = np.zeros(n, dtype=np.uint64)
        lib().hm_dump(self.h, out.ctypes.data)
        return out


def check_dump(opts: HmOpts, dump: np.ndarray):
    """Invariants over a canonical dump; returns (total, per-invariant counts)."""
    d = np.ascontiguousarray(dump, dtype=np.uint64)
    detail = np.zeros(7, dtype=np.int64)
    tot = lib().hm_check_dump(ctypes.byref(opts), d.ctypes.data, detail.ctypes.data)
    return int(tot), detail


def bench_plain(opts: HmOpts, nthreads: int, nlanes: int, iters: int):
    v = ctypes.c_uint64(0)
    r = lib().hm_bench_plain(ctypes.byref(opts), nthreads, nlanes, iters, ctypes.byref(v))
    return r, int(v.value)


def bench_glibc(nthreads: int, nlanes: int, iters: int):
    v = ctypes.c_uint64(0)
    r = lib().hm_bench_glibc(nthreads, nlanes, iters, ctypes.byref(v))
    return r, int(v.value)
