"""Host-side Python mirror of the MI355X halloc allocator's C ABI.

The allocator itself is device code (hamalloc / hafree in include/halloc.h,
implemented in halloc_amd/csrc/halloc.hip).  This module loads the C-ABI
library ``libhalloc_amd.so`` with ctypes and mirrors the reference's allocator
plug-in interface for host code (tst/common/halloc-wrapper.h:9-37:
``init`` / ``extfrag`` / ``shutdown``), plus the workload drivers its tests
and benchmarks run (the counterparts of tst/common/common.h:361-412 and the
tst/perf kernels).  There is no CPU fallback: if the library is missing,
loading fails with an error naming the build command.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

import numpy as np

__all__ = ["HallocOpts", "Halloc", "HallocError", "lib", "LIB_PATH"]

PKG = os.path.dirname(os.path.abspath(__file__))
# HALLOC_AMD_LIB selects another build of the library (libhalloc_amd_ool.so:
# the out-of-line entry points)
LIB_PATH = os.environ.get("HALLOC_AMD_LIB") or os.path.join(PKG, "libhalloc_amd.so")
NSIZES = 16
SIZE_TABLE_FIELDS = ("nchunks", "chunk_sz", "chunk_id", "ncib", "sparse", "hash_step",
                     "roomy", "busy", "block_sz", "nblocks", "tmod", "chunk_bytes")
# every symbol include/halloc_amd.h declares
EXPORTS = ("ha_default_opts", "ha_init_c", "ha_shutdown_c", "ha_extfrag_c", "ha_free_bytes_c",
           "ha_last_error", "ha_geometry_c", "ha_replay", "ha_dump_words", "ha_dump",
           "ha_run_plain", "ha_malloc_sizes", "ha_free_ptrs", "ha_read_words", "ha_probe_atomics",
           "ha_run_lists", "ha_run_phase", "ha_replay_batch", "ha_run_graphs", "ha_run_serve",
           "ha_probe_copy", "ha_run_test", "ha_run_sample", "ha_run_corr_test0", "ha_probe_chain",
           "ha_free_bytes_noalloc_c", "ha_run_touch")


class HallocError(RuntimeError):
    pass


class _Opts(ctypes.Structure):
    _fields_ = [("memory", ctypes.c_uint64),
                ("halloc_fraction", ctypes.c_double),
                ("busy_fraction", ctypes.c_double),
                ("roomy_fraction", ctypes.c_double),
                ("sparse_fraction", ctypes.c_double),
                ("sb_sz_sh", ctypes.c_int32),
                ("nheads", ctypes.c_int32),
                ("hash_step_frac", ctypes.c_double),
                ("cached_arena", ctypes.c_int32),
                ("fixed_sets", ctypes.c_int32)]


class _Geometry(ctypes.Structure):
    _fields_ = [("nsbs", ctypes.c_uint32), ("sb_sz", ctypes.c_uint32), ("sb_sz_sh", ctypes.c_uint32),
                ("nheads", ctypes.c_uint32), ("nsw", ctypes.c_uint32), ("bit_words", ctypes.c_uint32),
                ("nib_words", ctypes.c_uint32), ("pad", ctypes.c_uint32),
                ("arena_bytes", ctypes.c_uint64), ("large_bytes", ctypes.c_uint64),
                ("size_table", ctypes.c_uint32 * (16 * 12)),
                ("arena_base", ctypes.c_uint64), ("large_units", ctypes.c_uint32),
                ("large_classes", ctypes.c_uint32)]


class _PlainResult(ctypes.Structure):
    _fields_ = [("errors", ctypes.c_uint64), ("nulls", ctypes.c_uint64), ("total_ms", ctypes.c_double),
                ("wall_ms", ctypes.c_double)]


class _Corr0Result(ctypes.Structure):
    _fields_ = [("blocks", ctypes.c_uint64), ("nulls", ctypes.c_uint64), ("bad_contents", ctypes.c_uint64),
                ("bad_gaps", ctypes.c_uint64), ("free_ms", ctypes.c_double)]


class _ListResult(ctypes.Structure):
    _fields_ = [("requests", ctypes.c_uint64), ("nulls", ctypes.c_uint64), ("errors", ctypes.c_uint64),
                ("malloc_ms", ctypes.c_double), ("free_ms", ctypes.c_double), ("wall_ms", ctypes.c_double)]


class _PhaseArgs(ctypes.Structure):
    _fields_ = [("nthreads", ctypes.c_int32), ("bs", ctypes.c_int32), ("slots", ctypes.c_int32),
                ("nphases", ctypes.c_int32), ("max_steps", ctypes.c_int32), ("min_size", ctypes.c_int32),
                ("max_size", ctypes.c_int32), ("seed", ctypes.c_uint32), ("p_hi", ctypes.c_float),
                ("p_lo", ctypes.c_float), ("fill_hi", ctypes.c_double), ("fill_lo", ctypes.c_double)]


class _PhaseResult(ctypes.Structure):
    _fields_ = [("mallocs", ctypes.c_uint64), ("frees", ctypes.c_uint64), ("nulls", ctypes.c_uint64),
                ("errors", ctypes.c_uint64), ("kernel_ms", ctypes.c_double), ("wall_ms", ctypes.c_double),
                ("steps", ctypes.c_int32), ("trace_len", ctypes.c_int32), ("phase_steps", ctypes.c_int32 * 8)]


class _TestOpts(ctypes.Structure):
    _fields_ = [("test", ctypes.c_int32), ("nthreads", ctypes.c_int32), ("bs", ctypes.c_int32),
                ("ntries", ctypes.c_int32), ("nallocs", ctypes.c_int32), ("niters", ctypes.c_int32),
                ("alloc_sz", ctypes.c_uint32), ("max_alloc_sz", ctypes.c_uint32), ("max_alloc_sh", ctypes.c_uint32),
                ("distr", ctypes.c_uint32), ("period_mask", ctypes.c_uint32), ("group_sh", ctypes.c_uint32),
                ("seed", ctypes.c_uint32), ("pad", ctypes.c_uint32), ("probabs", (ctypes.c_float * 2) * 2),
                ("extfrag", ctypes.POINTER(ctypes.c_double))]


class _TestResult(ctypes.Structure):
    _fields_ = [("mallocs", ctypes.c_uint64), ("frees", ctypes.c_uint64), ("nulls", ctypes.c_uint64),
                ("checked", ctypes.c_uint64), ("checks_failed", ctypes.c_uint64), ("malloc_ms", ctypes.c_double),
                ("free_ms", ctypes.c_double), ("tries_run", ctypes.c_int32), ("failed_try", ctypes.c_int32),
                ("lat_n", ctypes.c_uint64 * 2), ("lat_avg", ctypes.c_double * 2), ("lat_min", ctypes.c_double * 2),
                ("lat_max", ctypes.c_double * 2), ("malloc_wall_ms", ctypes.c_double),
                ("free_wall_ms", ctypes.c_double)]


class _SampleArgs(ctypes.Structure):
    _fields_ = [("sample", ctypes.c_int32), ("n", ctypes.c_int32), ("bs", ctypes.c_int32), ("ncells", ctypes.c_int32),
                ("min_len", ctypes.c_int32), ("max_len", ctypes.c_int32), ("seed", ctypes.c_uint32),
                ("pad", ctypes.c_uint32)]


class _SampleResult(ctypes.Structure):
    _fields_ = [("nulls", ctypes.c_uint64), ("errors", ctypes.c_uint64), ("alloc_ms", ctypes.c_double),
                ("work_ms", ctypes.c_double), ("free_ms", ctypes.c_double)]


class _GraphArgs(ctypes.Structure):
    _fields_ = [("njobs", ctypes.c_int32), ("bs", ctypes.c_int32), ("batch", ctypes.c_int32),
                ("waves", ctypes.c_int32), ("plain_chunk_nodes", ctypes.c_int32), ("repeats", ctypes.c_int32),
                ("run_plain", ctypes.c_int32), ("run_batch", ctypes.c_int32), ("seed", ctypes.c_uint32),
                ("cross", ctypes.c_uint32)]


class _GraphResult(ctypes.Structure):
    _fields_ = [("nodes", ctypes.c_uint64), ("nulls", ctypes.c_uint64), ("errors", ctypes.c_uint64),
                ("plain_ms", ctypes.c_double), ("batch_ms", ctypes.c_double),
                ("plain_launches", ctypes.c_int32), ("pad", ctypes.c_int32)]


class _ServeIO(ctypes.Structure):
    _fields_ = [("arrival_ns", ctypes.c_void_p), ("length", ctypes.c_void_p), ("tag", ctypes.c_void_p),
                ("checksum", ctypes.c_void_p), ("produced", ctypes.c_void_p), ("first_ns", ctypes.c_void_p),
                ("done_ns", ctypes.c_void_p), ("lat_ns", ctypes.c_void_p), ("bad", ctypes.c_void_p),
                ("bad_first", ctypes.c_void_p), ("nreq", ctypes.c_int32),
                ("slots", ctypes.c_int32), ("sample_every", ctypes.c_int32), ("lat_cap", ctypes.c_int32),
                ("watermark", ctypes.c_uint64)]


class _ServeResult(ctypes.Structure):
    _fields_ = [("kernel_ms", ctypes.c_double), ("pages", ctypes.c_uint64), ("nulls", ctypes.c_uint64),
                ("stalls", ctypes.c_uint64), ("samples", ctypes.c_uint64), ("completed", ctypes.c_uint32),
                ("admitted", ctypes.c_uint32), ("clock_khz", ctypes.c_uint32), ("pad", ctypes.c_uint32)]


@dataclass
class HallocOpts:
    """Fields and defaults of halloc_opts_t (src/halloc.h:15-34), plus nheads
    (0: sized to the heap for live blocks of many size classes),
    hash_step_frac (0 keeps the reference's probe stride), cached_arena
    (0: uncached arena, coherent across XCDs; 1: L2-cached arena with a
    release in every hafree) and fixed_sets (0: the sets in use shrink as
    size classes appear; 1: every wave keeps its own set)."""
    memory: int = 512 << 20
    halloc_fraction: float = 0.75
    busy_fraction: float = 0.835
    roomy_fraction: float = 0.6
    sparse_fraction: float = 0.012
    sb_sz_sh: int = 22
    nheads: int = 32
    hash_step_frac: float = 0.0
    cached_arena: int = 0
    fixed_sets: int = 0

    def _c(self) -> _Opts:
        return _Opts(self.memory, self.halloc_fraction, self.busy_fraction, self.roomy_fraction,
                     self.sparse_fraction, self.sb_sz_sh, self.nheads, self.hash_step_frac,
                     self.cached_arena, self.fixed_sets)


_lib = None


def lib() -> ctypes.CDLL:
    """Loads libhalloc_amd.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise HallocError(f"{LIB_PATH} is missing: build it with `python -m halloc_amd.build`")
    L = ctypes.CDLL(LIB_PATH)
    P = ctypes.POINTER
    i32, u64, vp = ctypes.c_int, ctypes.c_uint64, ctypes.c_void_p
    sigs = {
        "ha_default_opts": (None, [P(_Opts)]),
        "ha_init_c": (i32, [P(_Opts)]),
        "ha_shutdown_c": (i32, []),
        "ha_extfrag_c": (i32, [i32, P(ctypes.c_double)]),
        "ha_free_bytes_c": (i32, [P(u64)]),
        "ha_last_error": (ctypes.c_char_p, []),
        "ha_geometry_c": (i32, [P(_Geometry)]),
        "ha_replay": (i32, [vp, i32, i32, vp, vp, vp]),
        "ha_dump_words": (ctypes.c_int64, []),
        "ha_dump": (i32, [vp, ctypes.c_int64]),
        "ha_run_plain": (i32, [i32, i32, i32, i32, i32, i32, vp, P(_PlainResult)]),
        "ha_malloc_sizes": (i32, [vp, i32, i32, vp]),
        "ha_run_corr_test0": (i32, [i32, i32, i32, P(_Corr0Result)]),
        "ha_probe_chain": (i32, [vp, i32, i32, vp]),
        "ha_free_bytes_noalloc_c": (i32, [P(ctypes.c_uint64)]),
        "ha_free_ptrs": (i32, [vp, i32, i32]),
        "ha_read_words": (i32, [vp, i32, vp]),
        "ha_run_touch": (i32, [i32, i32, i32, i32, vp, vp, vp]),
        "ha_probe_atomics": (i32, [vp]),
        "ha_run_lists": (i32, [vp, i32, i32, vp, i32, vp, P(_ListResult)]),
        "ha_run_phase": (i32, [P(_PhaseArgs), P(_PhaseResult), vp, i32]),
        "ha_replay_batch": (i32, [vp, i32, i32, vp, vp]),
        "ha_run_graphs": (i32, [P(_GraphArgs), vp, P(_GraphResult)]),
        "ha_run_serve": (i32, [P(_ServeIO), P(_ServeResult)]),
        "ha_probe_copy": (i32, [u64, i32, P(ctypes.c_double)]),
        "ha_run_test": (i32, [P(_TestOpts), P(_TestResult)]),
        "ha_run_sample": (i32, [P(_SampleArgs), P(_SampleResult)]),
    }
    for name, (res, args) in sigs.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def _check(rc: int) -> int:
    if rc < 0:
        raise HallocError(f"halloc error {rc}: {lib().ha_last_error().decode()}")
    return rc


class Halloc:
    """The allocator on the current device (mirrors tst/common/halloc-wrapper.h:9-37)."""

    @staticmethod
    def init(opts: HallocOpts | None = None) -> None:
        o = (opts or HallocOpts())._c()
        _check(lib().ha_init_c(ctypes.byref(o)))

    @staticmethod
    def shutdown() -> None:
        _check(lib().ha_shutdown_c())

    @staticmethod
    def extfrag(ideal: bool) -> float:
        v = ctypes.c_double(0)
        _check(lib().ha_extfrag_c(1 if ideal else 0, ctypes.byref(v)))
        return v.value

    @staticmethod
    def free_bytes() -> int:
        v = ctypes.c_uint64(0)
        _check(lib().ha_free_bytes_c(ctypes.byref(v)))
        return int(v.value)

    @staticmethod
    def free_bytes_noalloc() -> int:
        """free_bytes() reduced in the library's device scratch array: no allocation."""
        v = ctypes.c_uint64(0)
        _check(lib().ha_free_bytes_noalloc_c(ctypes.byref(v)))
        return int(v.value)

    @staticmethod
    def geometry() -> dict:
        g = _Geometry()
        _check(lib().ha_geometry_c(ctypes.byref(g)))
        table = np.frombuffer(bytes(g.size_table), dtype=np.uint32).reshape(NSIZES, 12).copy()
        return {"nsbs": g.nsbs, "sb_sz": g.sb_sz, "sb_sz_sh": g.sb_sz_sh, "nheads": g.nheads,
                "nsw": g.nsw, "bit_words": g.bit_words, "nib_words": g.nib_words,
                "arena_bytes": g.arena_bytes, "large_bytes": g.large_bytes, "size_table": table,
                "arena_base": g.arena_base, "large_units": g.large_units, "large_classes": g.large_classes}

    # ---------------------------------------------------------- workloads
    @staticmethod
    def replay(ops: np.ndarray, iheads=None):
        """ops: (nsteps, lanes, 2) uint32; iheads: optional head set per step.
        Returns (recs (nsteps, lanes, 4), head set used per step)."""
        ops = np.ascontiguousarray(ops, dtype=np.uint32)
        nsteps, lanes = ops.shape[0], ops.shape[1]
        recs = np.zeros((nsteps, lanes, 4), dtype=np.uint32)
        xcc = np.zeros(nsteps, dtype=np.uint32)
        ih = None if iheads is None else np.ascontiguousarray(iheads, dtype=np.uint32)
        _check(lib().ha_replay(ops.ctypes.data, nsteps, lanes, None if ih is None else ih.ctypes.data,
                               recs.ctypes.data, xcc.ctypes.data))
        return recs, xcc

    @staticmethod
    def replay_batch(ops: np.ndarray):
        """ops: (nsteps, lanes, 4) uint32 [kind, arg, count, 0] (kind 1 hamalloc_batch of count blocks of arg
        bytes; 2 hafree_batch of op arg's blocks; 3 the same, keeping the pointers for a later free);
        returns (recs (nsteps, lanes, 16, 4), xcc)."""
        ops = np.ascontiguousarray(ops, dtype=np.uint32)
        nsteps, lanes = ops.shape[0], ops.shape[1]
        recs = np.zeros((nsteps, lanes, 16, 4), dtype=np.uint32)
        xcc = np.zeros(nsteps, dtype=np.uint32)
        _check(lib().ha_replay_batch(ops.ctypes.data, nsteps, lanes, recs.ctypes.data, xcc.ctypes.data))
        return recs, xcc

    @staticmethod
    def dump() -> np.ndarray:
        n = _check(lib().ha_dump_words())
        out = np.zeros(n, dtype=np.uint64)
        _check(lib().ha_dump(out.ctypes.data, n))
        return out

    @staticmethod
    def run_plain(nthreads=1 << 20, bs=256, size=16, iters=1, warmup=1, launches=10):
        """Plain hamalloc/hafree kernel: `launches` timed launches back to back
        (total_ms between events at both ends, wall_ms by the host clock), then
        as many again, each between its own event pair (ms)."""
        ms = np.zeros(launches, dtype=np.float32)
        r = _PlainResult()
        _check(lib().ha_run_plain(nthreads, bs, size, iters, warmup, launches, ms.ctypes.data,
                                  ctypes.byref(r)))
        return {"errors": int(r.errors), "nulls": int(r.nulls), "total_ms": r.total_ms,
                "wall_ms": r.wall_ms, "ms": ms}

    @staticmethod
    def malloc_sizes(sizes, bs=256) -> np.ndarray:
        s = np.ascontiguousarray(sizes, dtype=np.uint32)
        ptrs = np.zeros(len(s), dtype=np.uint64)
        _check(lib().ha_malloc_sizes(s.ctypes.data, len(s), bs, ptrs.ctypes.data))
        return ptrs

    @staticmethod
    def probe_chain(rounds=(0, 1, 2, 4, 7, 8, 16), nthreads=1 << 20) -> dict:
        """Launch time (ms, best of 5) of a grid whose waves each make r
        dependent atomic round trips, for every r in rounds."""
        rr = np.ascontiguousarray(rounds, dtype=np.int32)
        ms = np.zeros(len(rr), dtype=np.float64)
        _check(lib().ha_probe_chain(rr.ctypes.data, len(rr), nthreads, ms.ctypes.data))
        return {int(r): float(m) for r, m in zip(rr, ms)}

    @staticmethod
    def run_corr_test0(nthreads=(2 << 20) // 8, nmallocs=8, bs=256) -> dict:
        """tst/corr/test/test.cu test 0: nthreads x nmallocs 16 B blocks, each
        written with its own address and read back, sorted gaps >= 16 B, freed."""
        r = _Corr0Result()
        _check(lib().ha_run_corr_test0(nthreads, nmallocs, bs, ctypes.byref(r)))
        return {"blocks": r.blocks, "nulls": r.nulls, "bad_contents": r.bad_contents, "bad_gaps": r.bad_gaps,
                "free_ms": r.free_ms}

    @staticmethod
    def free_ptrs(ptrs, bs=256) -> None:
        p = np.ascontiguousarray(ptrs, dtype=np.uint64)
        _check(lib().ha_free_ptrs(p.ctypes.data, len(p), bs))

    @staticmethod
    def read_words(ptrs) -> np.ndarray:
        p = np.ascontiguousarray(ptrs, dtype=np.uint64)
        out = np.zeros(len(p), dtype=np.uint32)
        _check(lib().ha_read_words(p.ctypes.data, len(p), out.ctypes.data))
        return out

    @staticmethod
    def run_touch(n=1 << 18, words=8, iters=64, reps=5) -> dict:
        """A caller's kernel over n blocks of 8*words bytes (ha_run_touch):
        mean launch ms and word sums of the store, load and update modes."""
        ms = np.zeros(3, dtype=np.float64)
        sums = np.zeros(3, dtype=np.uint64)
        nulls = np.zeros(1, dtype=np.uint64)
        _check(lib().ha_run_touch(n, words, iters, reps, ms.ctypes.data, sums.ctypes.data, nulls.ctypes.data))
        return {"store_ms": float(ms[2]), "load_ms": float(ms[0]), "update_ms": float(ms[1]),
                "store_sum": int(sums[2]), "load_sum": int(sums[0]), "update_sum": int(sums[1]),
                "nulls": int(nulls[0])}

    @staticmethod
    def probe_atomics() -> np.ndarray:
        out = np.zeros(3, dtype=np.float64)
        _check(lib().ha_probe_atomics(out.ctypes.data))
        return out

    @staticmethod
    def probe_copy(nbytes=1 << 30, reps=5) -> float:
        """GB/s (read + write) of a streaming device-to-device copy of nbytes."""
        v = ctypes.c_double(0)
        _check(lib().ha_probe_copy(nbytes, reps, ctypes.byref(v)))
        return v.value

    @staticmethod
    def run_lists(sizes: np.ndarray, order=None, bs=256, want_ptrs=False):
        """sizes: (per, n) uint32, 0 = no request; order: optional permutation of
        the per*n slots for the frees.  Returns timings, counts and (optionally)
        the addresses the mallocs returned."""
        sz = np.ascontiguousarray(sizes, dtype=np.uint32)
        per, n = sz.shape
        od = None if order is None else np.ascontiguousarray(order, dtype=np.uint32)
        ptrs = np.zeros(per * n, dtype=np.uint64) if want_ptrs else None
        r = _ListResult()
        _check(lib().ha_run_lists(sz.ctypes.data, n, per, None if od is None else od.ctypes.data, bs,
                                  None if ptrs is None else ptrs.ctypes.data, ctypes.byref(r)))
        out = {"requests": int(r.requests), "nulls": int(r.nulls), "errors": int(r.errors),
               "malloc_ms": r.malloc_ms, "free_ms": r.free_ms, "wall_ms": r.wall_ms}
        if ptrs is not None:
            out["ptrs"] = ptrs.reshape(per, n)
        return out

    @staticmethod
    def run_phase(nthreads=1 << 20, bs=256, slots=4, nphases=5, max_steps=400, min_size=16, max_size=256,
                  seed=0, p_hi=0.95, p_lo=0.05, fill_hi=0.91, fill_lo=0.09):
        a = _PhaseArgs(nthreads, bs, slots, nphases, max_steps, min_size, max_size, seed, p_hi, p_lo,
                       fill_hi, fill_lo)
        r = _PhaseResult()
        cap = nphases * max_steps
        trace = np.zeros((cap, 3), dtype=np.float64)
        _check(lib().ha_run_phase(ctypes.byref(a), ctypes.byref(r), trace.ctypes.data, cap))
        return {"mallocs": int(r.mallocs), "frees": int(r.frees), "nulls": int(r.nulls), "errors": int(r.errors),
                "kernel_ms": r.kernel_ms, "wall_ms": r.wall_ms, "steps": int(r.steps),
                "phase_steps": list(r.phase_steps)[:nphases], "trace": trace[:r.trace_len]}

    @staticmethod
    def run_test(test, nthreads, bs, ntries, nallocs, niters, alloc_sz, max_alloc_sz, max_alloc_sh, distr,
                 period_mask, group_sh, probabs, seed=0):
        """One of the reference test programs' kernels (test 0 checkptr, 1
        prob-checkptr, 2 throughput, 3 priv-throughput, 4 phase-throughput, 5 latency,
        6 phase-latency, 7 phase-alloc-write, 8 phase-extfrag) over a CommonOpts
        option set; see halloc_amd.harness for the command line.  Test 8 adds
        "extfrag": per try, [ha_extfrag(false), ha_extfrag(true)]."""
        pr = ((ctypes.c_float * 2) * 2)(*[(ctypes.c_float * 2)(*row) for row in probabs])
        ef = (ctypes.c_double * (2 * max(ntries, 1)))()
        a = _TestOpts(test, nthreads, bs, ntries, nallocs, niters, alloc_sz, max_alloc_sz, max_alloc_sh, distr,
                      period_mask, group_sh, seed, 0, pr, ef)
        r = _TestResult()
        _check(lib().ha_run_test(ctypes.byref(a), ctypes.byref(r)))
        out = {k: (list(getattr(r, k)) if k.startswith("lat_") else getattr(r, k)) for k, _ in _TestResult._fields_}
        if test == 8:
            out["extfrag"] = [(ef[2 * t], ef[2 * t + 1]) for t in range(r.tries_run)]
        return out

    @staticmethod
    def run_sample(sample, n, bs=128, ncells=32, min_len=31, max_len=31, seed=0):
        """A reference sample as an application check: "grid-points",
        "add-strings", "random-graph" (max_len: the largest degree) or "objects"
        (typed new (halloc) / hadelete; max_len: the largest array in words); see
        include/halloc_amd.h ha_run_sample."""
        kind = {"grid-points": 0, "add-strings": 1, "random-graph": 2, "objects": 3}[sample]
        a = _SampleArgs(kind, n, bs, ncells, min_len, max_len, seed, 0)
        r = _SampleResult()
        _check(lib().ha_run_sample(ctypes.byref(a), ctypes.byref(r)))
        return {k: getattr(r, k) for k, _ in _SampleResult._fields_}

    @staticmethod
    def run_graphs(nnodes, batch=8, waves=1024, bs=256, plain_chunk_nodes=1 << 23, repeats=1,
                   run_plain=True, run_batch=True, seed=0, cross=False):
        """Random-graph builds, plain per-call form and batched persistent form.
        cross: instead of timing, build with each form and check and free with
        the other (errors counts nodes that fail the other form's check)."""
        nn = np.ascontiguousarray(nnodes, dtype=np.uint32)
        a = _GraphArgs(len(nn), bs, batch, waves, plain_chunk_nodes, repeats, int(run_plain), int(run_batch),
                       seed, int(cross))
        r = _GraphResult()
        _check(lib().ha_run_graphs(ctypes.byref(a), nn.ctypes.data, ctypes.byref(r)))
        return {"nodes": int(r.nodes), "nulls": int(r.nulls), "errors": int(r.errors),
                "plain_ms": r.plain_ms, "batch_ms": r.batch_ms, "plain_launches": int(r.plain_launches)}

    @staticmethod
    def run_serve(arrival_ns, length, tag, slots=65536, watermark=64 << 20, sample_every=8, lat_cap=1 << 22):
        """Serving loop with continuous batching (one persistent launch).
        Returns per-request checksum, produced tokens, first-token and
        completion times (ns after the start), and sampled page-malloc
        latencies (ns)."""
        arr = np.ascontiguousarray(arrival_ns, dtype=np.uint64)
        ln = np.ascontiguousarray(length, dtype=np.uint32)
        tg = np.ascontiguousarray(tag, dtype=np.uint32)
        n = len(arr)
        if len(ln) != n or len(tg) != n:
            raise ValueError("arrival_ns, length and tag must have one entry per request")
        cs = np.zeros(n, dtype=np.uint32)
        prod = np.zeros(n, dtype=np.uint32)
        first = np.zeros(n, dtype=np.uint64)
        done = np.zeros(n, dtype=np.uint64)
        lat = np.zeros(max(lat_cap, 1), dtype=np.uint32)
        bad = np.zeros(n, dtype=np.uint32)
        badf = np.zeros(n, dtype=np.uint64)
        io = _ServeIO(arr.ctypes.data, ln.ctypes.data, tg.ctypes.data, cs.ctypes.data, prod.ctypes.data,
                      first.ctypes.data, done.ctypes.data, lat.ctypes.data, bad.ctypes.data, badf.ctypes.data,
                      n, slots, sample_every, lat_cap, watermark)
        r = _ServeResult()
        _check(lib().ha_run_serve(ctypes.byref(io), ctypes.byref(r)))
        return {"kernel_ms": r.kernel_ms, "pages": int(r.pages), "nulls": int(r.nulls), "stalls": int(r.stalls),
                "completed": int(r.completed), "admitted": int(r.admitted), "clock_khz": int(r.clock_khz),
                "checksum": cs, "produced": prod, "first_ns": first, "done_ns": done, "bad": bad,
                "bad_first": badf,
                "lat_ns": lat[:int(r.samples)].copy()}
