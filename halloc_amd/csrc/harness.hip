// harness.hip -- workload drivers of libhalloc_amd.so: the counterparts of the
// reference's test harness kernels (tst/common/common.h:361-412,
// tst/perf/priv-throughput/priv-throughput.cu:14-27, tst/corr/test/test.cu:38-95).
// Each driver copies host inputs to device memory once, launches on a stream
// it owns, times its launches with HIP events on that stream, and copies
// results back.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#define HALLOC_CPP  // the typed forms, for the objects sample
#include "../../include/halloc.h"
#include "../../include/halloc_amd.h"
#include "ha_internal.h"
#include "ha_workload.h"

__device__ void* ha_malloc_on_head(size_t nbytes, unsigned full);
__device__ int ha_malloc_batch_on_head(size_t nbytes, int cnt, void** out, unsigned full);
__device__ unsigned ha_head_set_full();
__device__ int hamalloc_batch(size_t nbytes, int cnt, void** out);
__device__ void hafree_batch(void* const* ptrs, int cnt);

namespace ha {
const Params& host_params();
int fail(int code, const std::string& msg);
}

using ha::u32;
using u16 = uint16_t;
using u8 = uint8_t;
using ha::u64;

namespace {

int hfail(int code, const std::string& m) { return ha::fail(code, m); }

#define HH(call)                                                                                 \
  do {                                                                                           \
    hipError_t e_ = (call);                                                                      \
    if (e_ != hipSuccess) return hfail(HA_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)


// Device buffer that frees itself.
struct DBuf {
  void* p = nullptr;
  ~DBuf() {
    if (p) (void)hipFree(p);
  }
  hipError_t alloc(size_t n) { return hipMalloc(&p, n ? n : 1); }
};

struct Stream {
  hipStream_t s = nullptr;
  ~Stream() {
    if (s) (void)hipStreamDestroy(s);
  }
};

struct Events {
  std::vector<hipEvent_t> e;
  ~Events() {
    for (hipEvent_t x : e) (void)hipEventDestroy(x);
  }
};

// ------------------------------------------------------------------ replay
// Record of a malloc of `nbytes` that returned p: a slab block is (size class,
// slab, chunk); a large block is (16 + its class, LARGE_SB, unit).
__device__ void malloc_rec(const ha::Params& P, void* p, u64 nbytes, u32& r0, u32& r1, u32& r2) {
  const u64 off = (u64)((char*)p - P.arena);
  if (off < P.arena_bytes) {
    const u32 sid = ha::size_id_from_nbytes((u32)nbytes);
    r0 = sid;
    r1 = (u32)(off >> P.sb_sz_sh);
    r2 = ha::chunk_div((u32)(off & (P.sb_sz - 1)), P.si[sid].chunk_sz);
  } else {
    const u64 units = (nbytes >> ha::LARGE_UNIT_SH) + ((nbytes & ((1ull << ha::LARGE_UNIT_SH) - 1)) != 0);
    r0 = ha::NSIZES + (units <= 1 ? 0u : 64u - (u32)__builtin_clzll(units - 1));
    r1 = ha::LARGE_SB;
    r2 = (u32)((off - P.arena_bytes) >> ha::LARGE_UNIT_SH);
  }
}
// Record of a free of p, read before the free: a slab block gives (chunks
// freed, slab, chunk) from its nibble, a large block (units freed, LARGE_SB,
// unit) from its start word; (0, none, none) when the free changes nothing.
__device__ void free_rec(const ha::Params& P, void* p, u32& r0, u32& r1, u32& r2) {
  const u64 off = (u64)((char*)p - P.arena);
  if (off < P.arena_bytes) {
    const u32 sb = (u32)(off >> P.sb_sz_sh);
    const u32 cs = __hip_atomic_load(&P.desc[sb].chunk_sz, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (cs) {
      const u32 ich = ha::chunk_div((u32)(off & (P.sb_sz - 1)), cs);
      const u64 w = __hip_atomic_load(P.nibs + (u64)sb * P.nib_words + ich / 16, __ATOMIC_RELAXED,
                                      __HIP_MEMORY_SCOPE_AGENT);
      const u32 nch = (u32)((w >> (4 * (ich % 16))) & 0xF);
      if (nch) { r0 = nch; r1 = sb; r2 = ich; }
    }
  } else if (off - P.arena_bytes < P.large_bytes && !((off - P.arena_bytes) & ((1ull << ha::LARGE_UNIT_SH) - 1))) {
    const u32 unit = (u32)((off - P.arena_bytes) >> ha::LARGE_UNIT_SH);
    const u32 s = __hip_atomic_load(P.lstart + unit, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (s) { r0 = 1u << (s - 1); r1 = ha::LARGE_SB; r2 = unit; }
  }
}

// One wavefront replays ops step by step: malloc lanes, then free lanes.
__global__ void __launch_bounds__(64) replay_k(const u32* ops, int nsteps, int lanes, const u32* iheads,
                                               u32* recs, u32* xcc, void** ptrs) {
  const u32 l = threadIdx.x;
  const ha::Params& P = ha::ha_p;
  for (int st = 0; st < nsteps; ++st) {
    const size_t i = (size_t)st * lanes + l;
    const u32 kind = ops[2 * i], arg = ops[2 * i + 1];
    // the full-spread head set of the step (given, or the wave's own); the
    // allocator folds it onto the sets in use, read coherently, so the
    // model can follow the fold step by step
    const u32 ih = iheads ? iheads[st] : ha_head_set_full();
    if (l == 0) xcc[st] = ih;
    if (kind == 1) {
      void* p = ha_malloc_on_head(arg, ih);
      ptrs[i] = p;
      u32 r0 = 0xffffffffu, r1 = 0xffffffffu, r2 = 0xffffffffu;
      if (p) malloc_rec(P, p, arg, r0, r1, r2);
      recs[4 * i] = r0; recs[4 * i + 1] = r1; recs[4 * i + 2] = r2; recs[4 * i + 3] = 0;
    }
    __syncthreads();
    if (kind == 2 || kind == 3) {
      // record the block's length from its nibble or start word, then free
      // it; kind 3 frees the pointer one large unit (4 KiB) into the block
      void* p = ptrs[arg];
      if (kind == 3 && p) p = (char*)p + (1ull << ha::LARGE_UNIT_SH);
      u32 r0 = 0, r1 = 0xffffffffu, r2 = 0xffffffffu;
      if (p) free_rec(P, p, r0, r1, r2);
      hafree(p);
      recs[4 * i] = r0; recs[4 * i + 1] = r1; recs[4 * i + 2] = r2; recs[4 * i + 3] = 0;
      ptrs[i] = nullptr;
    }
    __syncthreads();
  }
}

// ------------------------------------------------------------ batch replay
// One wavefront: per step the kind-1 lanes call hamalloc_batch together, then
// the kind-2 and kind-3 lanes call hafree_batch on an earlier op's blocks;
// kind 2 forgets the pointers, kind 3 keeps them, so that a later op can free
// the same blocks again.
__global__ void __launch_bounds__(64) replay_batch_k(const u32* ops, int nsteps, int lanes, u32* recs, u32* xcc,
                                                     void** ptrs) {
  const u32 l = threadIdx.x;
  const ha::Params& P = ha::ha_p;
  for (int st = 0; st < nsteps; ++st) {
    const size_t i = (size_t)st * lanes + l;
    const u32 kind = ops[4 * i], arg = ops[4 * i + 1], cnt = ops[4 * i + 2];
    const u32 ih = ha_head_set_full();
    if (l == 0) xcc[st] = ih;
    if (kind == 1) {
      void** out = ptrs + i * ha::MAXB;
      ha_malloc_batch_on_head(arg, (int)cnt, out, ih);
      const u32 c = cnt > ha::MAXB ? ha::MAXB : cnt;
      for (u32 j = 0; j < c; ++j) {
        void* p = out[j];
        u32* r = recs + 4 * (i * ha::MAXB + j);
        u32 r0 = 0xffffffffu, r1 = 0xffffffffu, r2 = 0xffffffffu;
        if (p) malloc_rec(P, p, arg, r0, r1, r2);
        r[0] = r0; r[1] = r1; r[2] = r2; r[3] = 0;
      }
    }
    __syncthreads();
    if (kind == 2 || kind == 3) {
      void** fp = ptrs + (size_t)arg * ha::MAXB;
      const u32 c = ops[4 * (size_t)arg + 2] > ha::MAXB ? ha::MAXB : ops[4 * (size_t)arg + 2];
      for (u32 j = 0; j < c; ++j) {
        void* p = fp[j];
        if (!p) continue;
        u32* r = recs + 4 * (i * ha::MAXB + j);
        r[0] = 0; r[1] = 0xffffffffu; r[2] = 0xffffffffu; r[3] = 0;
        free_rec(P, p, r[0], r[1], r[2]);
      }
      hafree_batch(fp, (int)c);
      if (kind == 2)
        for (u32 j = 0; j < c; ++j) fp[j] = nullptr;
    }
    __syncthreads();
  }
}

// ------------------------------------------------------------------- dump
__global__ void dump_k(u64* out, u64 total) {
  const ha::Params& P = ha::ha_p;
  const u64 n1 = P.nsbs, n2 = 2ull * P.nsbs, n3 = (u64)ha::NSETS * P.nsw, nh = (u64)P.nheads * ha::NSIZES;
  const u64 n5 = (u64)P.nsbs * P.bit_words, n6 = (u64)P.nsbs * P.nib_words;
  for (u64 i = blockIdx.x * (u64)blockDim.x + threadIdx.x; i < total; i += (u64)gridDim.x * blockDim.x) {
    u64 j = i, v = 0;
    if (j < n1) { v = P.ctr[j * ha::CTR_STRIDE]; }
    else if ((j -= n1) < n2) {
      const ha::Desc& d = P.desc[j / 2];
      v = (j & 1) ? ((u64)d.chunk_sz | (u64)d.is_head << 32) : ((u64)d.size_id | (u64)d.chunk_id << 32);
    } else if ((j -= n2) < n3) { v = P.sets[j]; }
    else if ((j -= n3) < nh) { v = P.head[j * ha::HOT_STRIDE]; }
    else if ((j -= nh) < nh) { v = P.cached[j * ha::HOT_STRIDE]; }
    else if ((j -= nh) < nh) { v = P.ticket[j * ha::HOT_STRIDE]; }
    else if ((j -= nh) < n5) { v = P.bits[j]; }
    else if ((j -= n5) < n6) { v = P.nibs[j]; }
    else if ((j -= n6) < 2) { v = j == 0 ? *P.spread : *P.class_seen; }
    else if ((j -= 2) < ha::LARGE_NCLASSES) { v = P.lticket[j * ha::HOT_STRIDE]; }
    else if ((j -= ha::LARGE_NCLASSES) < P.lwords) { v = P.lbits[j]; }
    else { j -= P.lwords; v = P.lstart[j]; }
    out[i] = v;
  }
}

u64 dump_words_of(const ha::Params& P) {
  return (u64)P.nsbs * 3 + (u64)ha::NSETS * P.nsw + 3ull * P.nheads * ha::NSIZES +
         (u64)P.nsbs * P.bit_words + (u64)P.nsbs * P.nib_words + 2 + ha::LARGE_NCLASSES + P.lwords + P.lunits;
}

// ---------------------------------------------------------------- plain
// adds the number of active lanes with `hit` to *c, one atomic per wave
__device__ __forceinline__ void wave_count(unsigned long long* c, bool hit) {
  const unsigned long long m = __ballot(hit);
  if (m && (u32)__lane_id() == (u32)__builtin_ctzll(__ballot(1))) atomicAdd(c, (unsigned long long)__popcll(m));
}
// Built for 8 waves per SIMD (64 VGPRs): the kernel waits on dependent
// atomics, so the eighth resident wave per SIMD pays.  NULLs and damaged
// patterns are counted where they happen, so that no per-lane counter stays
// live through the inlined allocator: with two such counters the 8-wave
// build spilled 12 VGPRs and ran about 7% slower.
__global__ void __launch_bounds__(256, 8) plain_k(int nthreads, int size, int iters, unsigned long long* errs,
                        unsigned long long* nulls) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nthreads) return;
  const u32 t = (u32)i;
  for (int it = 0; it < iters; ++it) {
    uint2* p = (uint2*)hamalloc((size_t)size);
    wave_count(nulls, !p);
    if (!p) continue;
    const uint2 a = make_uint2(t, ~t), b = make_uint2(t ^ 0x9E3779B9u, (u32)it);
    p[0] = a;
    p[1] = b;
    __asm__ __volatile__("" ::: "memory");
    const uint2 ra = p[0], rb = p[1];
    wave_count(errs, ra.x != a.x || ra.y != a.y || rb.x != b.x || rb.y != b.y);
    hafree(p);
  }
}

// -------------------------------------------------------- arrays of blocks
__global__ void malloc_sizes_k(const u32* sizes, int n, void** ptrs) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  void* p = hamalloc(sizes[i]);
  ptrs[i] = p;
  if (p) *(u32*)p = sizes[i];
}
__global__ void free_ptrs_k(void* const* ptrs, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  hafree(ptrs[i]);
}
__global__ void read_words_k(void* const* ptrs, int n, u32* out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  out[i] = ptrs[i] ? *(const u32*)ptrs[i] : 0u;
}

// ------------------------------------------------- a caller's own accesses
// What the arena's memory type costs a kernel that uses its blocks: lane i
// owns the `words` 64-bit words of block ptrs[i] and makes `iters` passes
// over them.  Mode 0 loads and sums, 1 loads, updates and stores, 2 stores.
// The empty asm between passes keeps every pass's accesses in the code.
constexpr u64 TOUCH_MUL = 6364136223846793005ull;
__global__ void __launch_bounds__(256) touch_alloc_k(u64** ptrs, int n, int words, unsigned long long* nulls) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  u64* p = (u64*)hamalloc((size_t)words * 8);
  ptrs[i] = p;
  if (!p) atomicAdd(nulls, 1ull);
}
__global__ void __launch_bounds__(256) touch_k(u64* const* ptrs, int n, int words, int iters, int mode,
                                               u64* sink) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  u64* b = ptrs[i];
  if (!b) return;
  u64 acc = 0;
  for (int it = 0; it < iters; ++it) {
    for (int w = 0; w < words; ++w) {
      if (mode == 0) acc += b[w];
      else if (mode == 1) b[w] = b[w] * TOUCH_MUL + (u64)it;
      else b[w] = (u64)it * (u64)words + (u64)w + (u64)i;
    }
    asm volatile("" ::: "memory");
  }
  if (mode == 0) sink[i] = acc;
}
// lane i: the sum of its block's words (mode 0: what touch_k stored in sink)
__global__ void __launch_bounds__(256) touch_sum_k(u64* const* ptrs, int n, int words, int mode, u64* sink) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || mode == 0) return;
  u64 s = 0;
  if (const u64* b = ptrs[i])
    for (int w = 0; w < words; ++w) s += b[w];
  sink[i] = s;
}
__global__ void __launch_bounds__(256) touch_free_k(u64* const* ptrs, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) hafree(ptrs[i]);
}

// ------------------------------------------ tst/corr/test/test.cu test 0
// malloc_k (:47-55): thread t of nthreads allocates ntimes 16 B blocks at
// ptrs[t], ptrs[t + nthreads], ... (a grid-stride loop)
__global__ void t0_malloc_k(void** ptrs, int ntimes, unsigned long long* nulls) {
  const u64 tid = blockIdx.x * (u64)blockDim.x + threadIdx.x, nt = (u64)blockDim.x * gridDim.x;
  u32 nul = 0;
  for (u64 i = tid; i < (u64)ntimes * nt; i += nt) {
    void* p = hamalloc(16);
    ptrs[i] = p;
    nul += p == nullptr;
  }
  if (nul) atomicAdd(nulls, (unsigned long long)nul);
}
// check_ptrs_k (:84-95), first half: two u64 holding the block's own address
__global__ void t0_write_k(void* const* ptrs, u64 n) {
  const u64 i = blockIdx.x * (u64)blockDim.x + threadIdx.x;
  if (i >= n || !ptrs[i]) return;
  u64* p = (u64*)ptrs[i];
  p[0] = (u64)p;
  p[1] = (u64)p;
}
// read back after the whole grid has written: a block that overlaps another
// holds the other's address
__global__ void t0_read_k(void* const* ptrs, u64 n, unsigned long long* bad) {
  const u64 i = blockIdx.x * (u64)blockDim.x + threadIdx.x;
  if (i >= n || !ptrs[i]) return;
  const u64* p = (const u64*)ptrs[i];
  if (p[0] != (u64)p || p[1] != (u64)p) atomicAdd(bad, 1ull);
}
// free_k (:57-63)
__global__ void t0_free_k(void* const* ptrs, int ntimes) {
  const u64 tid = blockIdx.x * (u64)blockDim.x + threadIdx.x, nt = (u64)blockDim.x * gridDim.x;
  for (u64 i = tid; i < (u64)ntimes * nt; i += nt) hafree(ptrs[i]);
}

// ------------------------------------------------------------------ probe
__device__ __forceinline__ u32 mix32(u32 x) {
  x ^= x >> 16; x *= 0x7feb352dU; x ^= x >> 15; x *= 0x846ca68bU; x ^= x >> 16;
  return x;
}
__global__ void probe_k(u64* buf, u64 mask, int iters, int shape, u64* sink) {
  const u32 tid = blockIdx.x * blockDim.x + threadIdx.x;
  const u32 lane = threadIdx.x & 63, wave = tid >> 6;
  u64 acc = 0;
  for (int i = 0; i < iters; ++i) {
    u64 w;
    if (shape == 0) w = mix32(tid * 131u + (u32)i * 7919u + 1u);
    else if (shape == 1) w = (u64)mix32(wave * 977u + (u32)i * 31u + 3u) * 64u + lane;
    else w = 0;
    w &= mask;
    acc ^= __hip_atomic_fetch_or(buf + w, 1ull << (lane & 31), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  sink[tid] = acc;
}

// Dependent-chain probe: every wave makes `rounds` atomic round trips one
// after another, as an allocator call does (head/ticket, probe, counter on
// malloc; nibble, bits, counter on free).  Lane 0 issues each atomic on a
// cache line picked from the previous result, so no round can start before
// the one before it returns; the result reaches the wave through
// readfirstlane.  The launch time over the number of rounds is the round-trip
// latency with the whole chip waiting on memory.
__global__ void __launch_bounds__(256) chain_k(u64* buf, u64 mask, int rounds, u64* sink) {
  const u32 tid = blockIdx.x * blockDim.x + threadIdx.x;
  const u32 lane = threadIdx.x & 63, wave = tid >> 6;
  u64 prev = 0;
  for (int r = 0; r < rounds; ++r) {
    const u64 w = ((u64)mix32(wave * 977u + (u32)r * 31u + (u32)(prev & 1u)) * 16u) & mask;
    u64 v = 0;
    if (lane == 0) v = __hip_atomic_fetch_or(buf + w, 1ull << (r & 31), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    prev = (u64)__builtin_amdgcn_readfirstlane((int)(u32)v);
  }
  if (lane == 0) sink[wave] = prev;
}

// Streaming copy of n16 16-byte words, grid-stride: the known-size kernel
// against which measured bandwidth and HBM byte counters are calibrated.
__global__ void __launch_bounds__(256) copy_k(const uint4* __restrict__ src, uint4* __restrict__ dst, u64 n16) {
  const u64 stride = (u64)gridDim.x * blockDim.x;
  for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += stride) dst[i] = src[i];
}

// --------------------------------------------------------------- lists
// Lane i allocates sizes[j*n + i] for j < per (0 = no request) and writes the
// size into the block's first word (tst/common/common.h:375-388).
__global__ void list_malloc_k(const u32* sizes, int n, int per, void** ptrs,
                              unsigned long long* nulls) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  u32 nul = 0;
  for (int j = 0; j < per; ++j) {
    const size_t k = (size_t)j * n + i;
    const u32 sz = sizes[k];
    void* p = nullptr;
    if (sz) {
      p = hamalloc(sz);
      if (p) *(u32*)p = sz;
      else nul++;
    }
    ptrs[k] = p;
  }
  if (nul) atomicAdd(nulls, (unsigned long long)nul);
}
// Thread t frees ptrs[order[j*n + t]] (or ptrs[j*n + t]) after checking the
// first word (tst/common/common.h:391-399).
__global__ void list_free_k(void** ptrs, const u32* sizes, const u32* order, int n, int per,
                            unsigned long long* errs) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  u32 bad = 0;
  for (int j = 0; j < per; ++j) {
    const size_t k = order ? (size_t)order[(size_t)j * n + t] : (size_t)j * n + t;
    void* p = ptrs[k];
    if (p) {
      if (*(const u32*)p != sizes[k]) bad++;
      hafree(p);
    }
  }
  if (bad) atomicAdd(errs, (unsigned long long)bad);
}

// ---------------------------------------------------------------- phases
__device__ __forceinline__ u32 hash3(u32 a, u32 b, u32 c) {
  u32 x = a * 0x9E3779B1u ^ (b + 0x7F4A7C15u) * 0x85EBCA77u ^ (c + 0x165667B1u) * 0xC2B2AE3Du;
  x ^= x >> 16; x *= 0x7feb352dU; x ^= x >> 15; x *= 0x846ca68bU; x ^= x >> 16;
  return x;
}
// One step of the phase workload (tst/perf/phase-throughput/phase-throughput.cu:18-50):
// each lane takes one action per step: with probability p_alloc it
// allocates a block of 8*U{lo8..hi8} bytes into its next empty slot and
// stamps it, otherwise it checks the stamp of its next live slot and frees
// it.  stats: mallocs, frees, nulls, errors, bytes allocated, bytes freed,
// summed per wave before one lane adds them.
__device__ __forceinline__ unsigned long long wave_sum64(unsigned long long v) {
  for (int o = 32; o; o >>= 1) v += __shfl_xor(v, o);
  return v;
}
__global__ void __launch_bounds__(256) phase_step_k(void** ptrs, u32* sizes, int n, int slots, u32 seed, u32 step,
                                                    float p_alloc, u32 lo8, u32 hi8, unsigned long long* stats) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  u32 nm = 0, nf = 0, nn = 0, ne = 0;
  unsigned long long ba = 0, bf = 0;
  if (i < n) {
    const u32 r = hash3(seed, (u32)i, step);
    const float u = (float)(r >> 8) * (1.0f / 16777216.0f);
    const int s0 = (int)(hash3(r, seed, 11u) % (u32)slots);
    if (u < p_alloc) {
      for (int t = 0; t < slots; ++t) {
        const int s = (s0 + t) % slots;
        const size_t k = (size_t)s * n + i;
        if (ptrs[k]) continue;
        const u32 sz = 8u * (lo8 + hash3(r, seed, 7u) % (hi8 - lo8 + 1));
        void* q = hamalloc(sz);
        if (q) {
          *(u32*)q = (u32)i ^ (sz << 20);
          ptrs[k] = q;
          sizes[k] = sz;
          nm++;
          ba += sz;
        } else {
          nn++;
        }
        break;
      }
    } else {
      for (int t = 0; t < slots; ++t) {
        const int s = (s0 + t) % slots;
        const size_t k = (size_t)s * n + i;
        void* p = ptrs[k];
        if (!p) continue;
        const u32 sz = sizes[k];
        if (*(const u32*)p != ((u32)i ^ (sz << 20))) ne++;
        hafree(p);
        ptrs[k] = nullptr;
        nf++;
        bf += sz;
        break;
      }
    }
  }
  const unsigned long long v[6] = {wave_sum64(nm), wave_sum64(nf), wave_sum64(nn), wave_sum64(ne), wave_sum64(ba),
                                   wave_sum64(bf)};
  if ((threadIdx.x & 63) == 0)
    for (int c = 0; c < 6; ++c)
      if (v[c]) atomicAdd(&stats[c], v[c]);
}
// frees what is left (free_rest_k, tst/common/common.h:403-412)
__global__ void __launch_bounds__(256) phase_rest_k(void** ptrs, u32* sizes, int n, int slots,
                                                    unsigned long long* stats) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  u32 nf = 0, ne = 0;
  for (int s = 0; i < n && s < slots; ++s) {
    const size_t k = (size_t)s * n + i;
    void* p = ptrs[k];
    if (!p) continue;
    if (*(const u32*)p != ((u32)i ^ (sizes[k] << 20))) ne++;
    hafree(p);
    ptrs[k] = nullptr;
    nf++;
  }
  const unsigned long long wf = wave_sum64(nf), we = wave_sum64(ne);
  if ((threadIdx.x & 63) == 0) {
    if (wf) atomicAdd(&stats[1], wf);
    if (we) atomicAdd(&stats[3], we);
  }
}

// ---------------------------------------------------------------- graphs
// Random-graph builds (the shape of samples/random-graph/random-graph.cu:79-84,
// 167-181, with frees): node v of job j has a 16 B header {v, degree, adjacency}
// and an int32 adjacency array of Poisson(8) targets in [0, n_j); degree and
// targets come from a counter-based hash, so every form builds the same graphs.
struct NodeHdr {
  u32 node, deg;
  u32* adj;
};
__device__ __forceinline__ u32 gh(u32 seed, u32 job, u32 node, u32 e) {
  return hash3(seed ^ (job * 0x9E3779B9u), node, e);
}
// Poisson(8) by inversion (the thresholds: ha_workload.h): the degree is the
// number of thresholds at or below a 24-bit uniform from the hash.
// Branch-free: every lane does the same compares, where a product-of-uniforms
// loop would run as many hash rounds as the wave's largest degree.
__device__ __forceinline__ u32 poisson8(u32 seed, u32 job, u32 node) {
  const u32 u = gh(seed, job, node, 1000u) >> 8;
  u32 k = 0;
#pragma unroll
  for (u32 i = 0; i < ha::POISSON8_N; ++i) k += u >= ha::kPoisson8.t[i];
  return k;
}
__device__ __forceinline__ u32 adj_bytes(u32 deg) { return deg ? 4u * deg : 16u; }
// The adjacency array is written and read as 8-byte pairs: blocks are 8-byte
// aligned and their class size is a multiple of 8 that is at least 4*deg, so
// the pair holding the last entry of an odd degree stays inside the block.
// Targets are hash mod n; m = mod_magic(n) replaces the division.
__device__ __forceinline__ void fill_node(NodeHdr* h, u32* a, u32 seed, u32 job, u32 node, u32 deg, u32 n, u32 m) {
  for (u32 e = 0; e < deg; e += 2) {
    const u32 v0 = ha::mod_by(gh(seed, job, node, e), n, m);
    if (e + 1 < deg) *(uint2*)(a + e) = make_uint2(v0, ha::mod_by(gh(seed, job, node, e + 1), n, m));
    else a[e] = v0;
  }
  h->node = node;
  h->deg = deg;
  h->adj = a;
}
__device__ __forceinline__ u32 check_node(const NodeHdr* h, u32 seed, u32 job, u32 node, u32 n, u32 deg) {
  if (h->node != node || h->deg != deg || !h->adj) return 1;
  const u32 m = ha::mod_magic(n);
  u32 bad = 0;
  for (u32 e = 0; e < deg; e += 2) {
    const uint2 v = *(const uint2*)(h->adj + e);
    bad += v.x != ha::mod_by(gh(seed, job, node, e), n, m);
    if (e + 1 < deg) bad += v.y != ha::mod_by(gh(seed, job, node, e + 1), n, m);
  }
  return bad ? 1 : 0;
}
__device__ __forceinline__ u32 job_of(const u64* offs, u32 njobs, u64 gi) {
  u32 lo = 0, hi = njobs;  // offs[lo] <= gi < offs[hi]
  while (hi - lo > 1) {
    u32 mid = (lo + hi) / 2;
    if (offs[mid] <= gi) lo = mid; else hi = mid;
  }
  return lo;
}

// plain per-call form: one thread per node of the job range [j0, j1)
__global__ void graph_plain_build_k(const u64* offs, const u32* nnodes, u32 njobs, u32 seed, u64 g0, u64 g1,
                                    NodeHdr** table, unsigned long long* stats) {
  const u64 gi = g0 + blockIdx.x * (u64)blockDim.x + threadIdx.x;
  if (gi >= g1) return;
  const u32 job = job_of(offs, njobs, gi), node = (u32)(gi - offs[job]), n = nnodes[job];
  const u32 deg = poisson8(seed, job, node);
  NodeHdr* h = (NodeHdr*)hamalloc(sizeof(NodeHdr));
  u32* a = (u32*)hamalloc(adj_bytes(deg));
  if (h && a) {
    fill_node(h, a, seed, job, node, deg, n, ha::mod_magic(n));
  } else {
    atomicAdd(&stats[0], 1ull);
    if (a) hafree(a);
    if (h) hafree(h);
    h = nullptr;
  }
  table[gi] = h;
}
__global__ void graph_plain_free_k(const u64* offs, const u32* nnodes, u32 njobs, u32 seed, u64 g0, u64 g1,
                                   NodeHdr** table, unsigned long long* stats) {
  const u64 gi = g0 + blockIdx.x * (u64)blockDim.x + threadIdx.x;
  if (gi >= g1) return;
  NodeHdr* h = table[gi];
  if (!h) return;
  const u32 job = job_of(offs, njobs, gi), node = (u32)(gi - offs[job]);
  if (check_node(h, seed, job, node, nnodes[job], poisson8(seed, job, node))) {
    atomicAdd(&stats[1], 1ull);  // a damaged header's adjacency pointer is not trusted
  } else {
    hafree(h->adj);
  }
  hafree(h);
}

// Batched persistent form.  Work items are chunks of 64*B nodes of one job
// (node = chunk*64*B + j*64 + lane, j < B).  The host lays the items out as
// one sequence: every build item in job order, and each job's free items (check
// and free one chunk) a fixed lag after the job's last build item.  Waves take
// GRAPH_GRAB items at a time from one counter; a free item waits until its
// job's build items, all taken earlier by running waves, are done.  Inside a
// chunk, allocation is class-major: all of the wave's adjacency arrays of one
// size class come from ONE hamalloc_batch call in which every lane takes an
// equal share, handed to the nodes through LDS; the frees use the same layout,
// so each round frees a contiguous run.
constexpr u32 GRAPH_GRAB = 2;

__global__ void graph_reset_k(u32* next, u32* built, u32 njobs) {
  const u32 i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0) *next = 0;
  for (u32 k = i; k < njobs; k += gridDim.x * blockDim.x) built[k] = 0;
}

__device__ __forceinline__ u32 rfl32(u32 v) { return (u32)__builtin_amdgcn_readfirstlane((int)v); }

// LDS of one wave is written and read back by other lanes of the same wave
__device__ __forceinline__ void wave_lds_sync() { __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront"); }

// A node's degree, adjacency size class and class-major slot in one word:
// deg | cls << 8 | slot << 16 (deg < 64, cls < 32, slot < 64*B <= 2^16)
__device__ __forceinline__ u32 nd_deg(u32 x) { return x & 0xFFu; }
__device__ __forceinline__ u32 nd_cls(u32 x) { return (x >> 8) & 0xFFu; }
__device__ __forceinline__ u32 nd_slot(u32 x) { return x >> 16; }
// in the free path the slot bits hold the owner lane (6 bits) and list round (5 bits)
__device__ __forceinline__ u32 nd_owner(u32 x) { return (x >> 21) & 63u; }
__device__ __forceinline__ u32 nd_round(u32 x) { return (x >> 16) & 31u; }

template <int B>
__device__ __forceinline__ u32 classes_present(const u32* nd, u32 cnt) {
  u32 mine = 0;
#pragma unroll
  for (int j = 0; j < B; ++j)
    if ((u32)j < cnt) mine |= 1u << nd_cls(nd[j]);
  u32 all = 0;
  for (u32 c = 0; c < ha::NSIZES; ++c)
    if (__ballot((mine >> c) & 1u)) all |= 1u << c;
  return all;
}

// A node in the rank table: j (4 bits), its lane (6 bits), its degree (6 bits)
__device__ __forceinline__ u16 node_rec(u32 j, u32 lane, u32 deg) { return (u16)(j | lane << 4 | deg << 10); }
static_assert(ha::POISSON8_N < 64, "degrees must fit the node record");

// Slots the wave's nodes of class c class-major from `at` on (nodes in (j,
// lane) order), records each slot's node in the rank table and returns how
// many there are.
template <int B>
__device__ __forceinline__ u32 slot_class(u32* nd, u32 cnt, u32 c, u32 at, u64 lt, u16* rk, u32 me) {
  u32 T = 0;
#pragma unroll
  for (int j = 0; j < B; ++j) {
    const bool mine = (u32)j < cnt && nd_cls(nd[j]) == c;
    const u64 m = __ballot(mine);
    if (mine) {
      const u32 slot = at + T + (u32)__builtin_popcountll(m & lt);
      nd[j] |= slot << 16;
      rk[slot] = node_rec((u32)j, me, nd_deg(nd[j]));
    }
    T += (u32)__builtin_popcountll(m);
  }
  return T;
}

// The adjacency contents of a chunk, class-major.  Class c's T blocks come
// from one batched allocation in which lane l took ranks [l*q, l*q + q),
// round i of the allocation holding rank l*q + i for every lane: one slab's
// contiguous run.  The wave goes through the runs 8 bytes (two entries) per
// lane at a time, so that consecutive lanes touch consecutive pieces of the
// run instead of 64 scattered blocks.  slots[at + r] is rank r's block (NULL:
// skip), rk[at + r] its node.  WRITE fills the blocks (entries past the
// degree are 0); otherwise each piece is checked and flags[j*64 + lane] set
// for a node whose contents differ.
template <int B, bool WRITE>
__device__ void adjacency_runs(const u32* nd, u32 cnt, u32 job, u32 base, u32 n, u32 seed, void* const* slots,
                               const u16* rk, u8* flags, u32 me) {
  const u32 m = ha::mod_magic(n);
  u32 at = 0;
  for (u32 cs = classes_present<B>(nd, cnt); cs; cs &= cs - 1) {
    const u32 c = (u32)__builtin_ctz(cs);
    u32 T = 0;
#pragma unroll
    for (int j = 0; j < B; ++j) T += (u32)__builtin_popcountll(__ballot((u32)j < cnt && nd_cls(nd[j]) == c));
    const u32 q = (T + 63) / 64;
    const u32 ppb = ha::ha_p.si[c].block_sz / 8u, pm = ha::mod_magic(ppb);  // 8 B pieces per block
    for (u32 i = 0; i < q; ++i) {
      for (u32 x = me; x < 64u * ppb; x += 64u) {
        u32 b = ppb == 1 ? x : (u32)(((u64)x * pm) >> 32);  // x / ppb, or one less
        b += x - b * ppb >= ppb;
        const u32 r = b * q + i, e = 2u * (x - b * ppb);
        if (r >= T) continue;
        u32* a = (u32*)slots[at + r];
        if (!a) continue;
        const u32 rec = rk[at + r], j = rec & 15u, l = (rec >> 4) & 63u, d = rec >> 10;
        if (!WRITE && e >= d) continue;
        const u32 v = base + j * 64u + l;
        const u32 w0 = e < d ? ha::mod_by(gh(seed, job, v, e), n, m) : 0u;
        const u32 w1 = e + 1 < d ? ha::mod_by(gh(seed, job, v, e + 1), n, m) : 0u;
        if (WRITE) {
          *(uint2*)(a + e) = make_uint2(w0, w1);
        } else {
          const uint2 got = *(const uint2*)(a + e);
          if (got.x != w0 || (e + 1 < d && got.y != w1)) flags[j * 64u + l] = 1;
        }
      }
    }
    at += T;
  }
}

// LDS row of one lane: its own hamalloc_batch/hafree_batch pointer list of up
// to max(B, 16) entries.  The odd stride puts the 64 lanes' rows in different
// banks.
template <int B>
constexpr u32 row_stride() {
  return B + 1 > 17 ? B + 1 : 17;
}
template <int B>
__device__ __forceinline__ void** lane_row(void** lds, u32 me) {
  return lds + me * row_stride<B>();
}

template <int B>
__device__ void graph_chunk_build(u32 job, u32 chunk, u32 n, u32 seed, NodeHdr** tab, void** lds, u16* rk, u32 me,
                                  u64 lt, u32& nulls) {
  const u32 base = chunk * 64u * B;
  u32 nd[B];
  u32 hdr[B];  // header offsets in the arena, in 16 B units (~0u: none)
  u32 cnt = 0;  // this lane's nodes: a prefix of j
#pragma unroll
  for (int j = 0; j < B; ++j) {
    const u32 v = base + (u32)j * 64u + me;
    cnt += v < n;
    const u32 d = v < n ? poisson8(seed, job, v) : 0u;
    nd[j] = d | (ha::size_id_from_nbytes(adj_bytes(d)) << 8);
  }
  // headers: the lane's cnt headers into its LDS row, then into registers
  void** row = lane_row<B>(lds, me);
  if (cnt) hamalloc_batch(sizeof(NodeHdr), (int)cnt, row);
#pragma unroll
  for (int j = 0; j < B; ++j) hdr[j] = (u32)j < cnt && row[j] ? (u32)(((char*)row[j] - ha::ha_p.arena) >> 4) : ~0u;
  wave_lds_sync();
  // adjacency arrays, class-major: class c's blocks fill LDS slots
  // [at, at + T) from ONE hamalloc_batch in which every lane takes an equal
  // share
  u32 at = 0;
  for (u32 cs = classes_present<B>(nd, cnt); cs; cs &= cs - 1) {
    const u32 c = (u32)__builtin_ctz(cs);
    const u32 T = slot_class<B>(nd, cnt, c, at, lt, rk, me);
    const u32 q = (T + 63) / 64, lo = me * q;  // lane share: [lo, lo + k)
    const u32 k = lo >= T ? 0u : (T - lo < q ? T - lo : q);
    if (k) hamalloc_batch(ha::ha_p.si[c].block_sz, (int)k, lds + at + lo);  // straight into LDS
    at += T;
  }
  wave_lds_sync();
  // headers and table entries per node; a node that got one block but not
  // the other gives it back (its slot goes NULL before the contents pass)
#pragma unroll
  for (int j = 0; j < B; ++j) {
    if ((u32)j >= cnt) continue;
    const u32 v = base + (u32)j * 64u + me;
    NodeHdr* h = hdr[j] == ~0u ? nullptr : (NodeHdr*)(ha::ha_p.arena + ((u64)hdr[j] << 4));
    u32* a = (u32*)lds[nd_slot(nd[j])];
    if (h && a) {
      h->node = v;
      h->deg = nd_deg(nd[j]);
      h->adj = a;
    } else {
      nulls++;
      if (a) hafree(a);
      if (h) hafree(h);
      lds[nd_slot(nd[j])] = nullptr;
      h = nullptr;
    }
    tab[v] = h;
  }
  wave_lds_sync();
  adjacency_runs<B, true>(nd, cnt, job, base, n, seed, lds, rk, nullptr, me);
  wave_lds_sync();  // the LDS slots are read before the next item reuses them
}

template <int B>
__device__ void graph_chunk_free(u32 job, u32 chunk, u32 n, u32 seed, NodeHdr** tab, void** lds, u16* rk, u8* flags,
                                 u32 me, u64 lt, u32& bad) {
  const u32 base = chunk * 64u * B;
  u32 nd[B];
  void* hdr[B];
  void* adj[B];
  u32 cnt = 0;
  // the lane's nodes are checked together: table entries, then headers, then
  // the adjacency pairs, each set of loads issued before any is used
#pragma unroll
  for (int j = 0; j < B; ++j) {
    const u32 v = base + (u32)j * 64u + me;
    cnt += v < n;
    hdr[j] = v < n ? tab[v] : nullptr;
  }
  constexpr int H = B < 8 ? B : 8;  // headers in groups of 8 bound the registers they hold
#pragma unroll
  for (int j0 = 0; j0 < B; j0 += H) {
    uint4 hv[H];
#pragma unroll
    for (int t = 0; t < H; ++t) hv[t] = hdr[j0 + t] ? *(const uint4*)hdr[j0 + t] : make_uint4(0u, 0u, 0u, 0u);
#pragma unroll
    for (int t = 0; t < H; ++t) {
      const int j = j0 + t;
      const u32 v = base + (u32)j * 64u + me;
      const u32 d = (u32)j < cnt ? poisson8(seed, job, v) : 0u;
      nd[j] = d | (ha::size_id_from_nbytes(adj_bytes(d)) << 8);
      adj[j] = nullptr;
      if (hdr[j]) {
        const u32* a = (const u32*)(((u64)hv[t].w << 32) | hv[t].z);
        if (hv[t].x != v || hv[t].y != d || !a) bad++;  // a damaged node's adjacency is not trusted
        else adj[j] = (void*)a;
      }
    }
  }
  // the headers are read: free them (the lane's list through its LDS row)
  void** row = lane_row<B>(lds, me);
#pragma unroll
  for (int j = 0; j < B; ++j)
    if ((u32)j < cnt) row[j] = hdr[j];
  if (cnt) hafree_batch(row, (int)cnt);
  wave_lds_sync();
  // the adjacency contents, class-major in the layout the build allocated
  // them in (same nodes, same classes, same ranks)
#pragma unroll
  for (int j = 0; j < B; ++j) flags[j * 64 + me] = 0;
  u32 at = 0;
  for (u32 cs = classes_present<B>(nd, cnt); cs; cs &= cs - 1) at += slot_class<B>(nd, cnt, (u32)__builtin_ctz(cs), at, lt, rk, me);
#pragma unroll
  for (int j = 0; j < B; ++j)
    if ((u32)j < cnt) lds[nd_slot(nd[j])] = adj[j];
  wave_lds_sync();
  adjacency_runs<B, false>(nd, cnt, job, base, n, seed, lds, rk, flags, me);
  wave_lds_sync();
  u32 mism = 0;  // bit j: node j's adjacency differs from its hash
#pragma unroll
  for (int j = 0; j < B; ++j)
    if (flags[j * 64 + me]) mism |= 1u << j;
#pragma unroll
  for (int j = 0; j < B; ++j) nd[j] &= 0xFFFFu;  // the slot bits are reused below
  bad += (u32)__builtin_popcount(mism);
  // adjacency arrays: class c's T blocks are shared out as in their
  // allocation (lane l frees ranks [l*q, l*q + q), q = ceil(T/64), rank
  // class-major), and each lane lists its shares of every class one after
  // another: round off_c + i of the lists is round i of class c, one slab's
  // contiguous runs.  hafree_batch takes the lists 16 rounds at a time
  // through each lane's LDS row, so all classes go in one or two calls.  A
  // node's owner lane and list round replace its slot bits.
  u32 nr = 0;  // rounds of the lists
  for (u32 cs = classes_present<B>(nd, cnt); cs; cs &= cs - 1) {
    const u32 c = (u32)__builtin_ctz(cs);
    u32 T = 0;
#pragma unroll
    for (int j = 0; j < B; ++j) T += (u32)__builtin_popcountll(__ballot((u32)j < cnt && nd_cls(nd[j]) == c));
    const u32 q = (T + 63) / 64, qm = ha::mod_magic(q);
    u32 r = 0;
#pragma unroll
    for (int j = 0; j < B; ++j) {
      const bool mine = (u32)j < cnt && nd_cls(nd[j]) == c;
      const u64 mk = __ballot(mine);
      if (mine) {
        const u32 rank = r + (u32)__builtin_popcountll(mk & lt);
        u32 own = q == 1 ? rank : (u32)(((u64)rank * qm) >> 32);  // rank / q, or one less
        own += rank - own * q >= q;
        nd[j] |= (own << 5 | (nr + rank - own * q)) << 16;
      }
      r += (u32)__builtin_popcountll(mk);
    }
    nr += q;
  }
  for (u32 r0 = 0; r0 < nr; r0 += ha::MAXB) {
    const u32 nrr = nr - r0 < ha::MAXB ? nr - r0 : ha::MAXB;
    for (u32 i = 0; i < nrr; ++i) row[i] = nullptr;
    wave_lds_sync();
#pragma unroll
    for (int j = 0; j < B; ++j) {
      const u32 rd = nd_round(nd[j]);
      if ((u32)j < cnt && rd >= r0 && rd < r0 + nrr)
        lds[nd_owner(nd[j]) * row_stride<B>() + (rd - r0)] = (mism >> j) & 1u ? nullptr : adj[j];
    }
    wave_lds_sync();
    hafree_batch(row, (int)nrr);  // straight from LDS
    wave_lds_sync();
  }
}

// at least 3 waves per SIMD (<= 168 VGPRs): measured faster than the
// unconstrained 2 at B = 16, despite a few spills
template <int B>
__global__ void __launch_bounds__(256, 3) graph_batch_k(const u64* seq, u32 nseq, const u32* coffs, const u64* offs,
                                                     const u32* nnodes, u32 seed, NodeHdr** table, u32* next,
                                                     u32* built, unsigned long long* stats) {
  static_assert(B <= 16, "the node record holds j in 4 bits");
  __shared__ void* lds_all[4][64 * row_stride<B>()];
  __shared__ u16 rk_all[4][64 * B];   // rank table: the node of each class-major slot
  __shared__ u8 flags_all[4][64 * B];  // per node: its adjacency check failed
  void** lds = lds_all[threadIdx.x >> 6];
  u16* rk = rk_all[threadIdx.x >> 6];
  u8* flags = flags_all[threadIdx.x >> 6];
  const u32 me = threadIdx.x & 63;
  const u64 lt = me ? (~0ull >> (64 - me)) : 0ull;
  u32 nulls = 0, bad = 0;
  for (;;) {
    u32 g = 0;
    if (me == 0) g = atomicAdd(next, GRAPH_GRAB);
    g = rfl32(g);
    if (g >= nseq) break;
    const u32 gend = g + GRAPH_GRAB < nseq ? g + GRAPH_GRAB : nseq;
    for (u32 it = g; it < gend; ++it) {
      const u64 e = seq[it];
      const u32 job = (u32)(e >> 32) & 0x7fffffffu, chunk = (u32)e;
      const u32 nch = coffs[job + 1] - coffs[job];
      if (e >> 63) {
        while (rfl32(__hip_atomic_load(&built[job], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) < nch)
          __builtin_amdgcn_s_sleep(1);
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        graph_chunk_free<B>(job, chunk, nnodes[job], seed, table + offs[job], lds, rk, flags, me, lt, bad);
      } else {
        graph_chunk_build<B>(job, chunk, nnodes[job], seed, table + offs[job], lds, rk, me, lt, nulls);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");  // the chunk and its table entries before the count
        if (me == 0) atomicAdd(&built[job], 1u);
      }
    }
  }
  const unsigned long long wn = wave_sum64(nulls), wb = wave_sum64(bad);
  if (me == 0) {
    if (wn) atomicAdd(&stats[0], wn);
    if (wb) atomicAdd(&stats[1], wb);
  }
}

// ---------------------------------------------------------------- serving
// Device-resident request loop with continuous batching.  Each lane is one
// sequence slot; wave w serves the request stream r = w, w + W, w + 2W, ...
// (W waves), in arrival order: its idle lanes admit the stream's requests
// that have arrived (arrival[r] wall-clock ticks after the start) while the
// heap's free bytes stay >= the watermark.  Wave 0 keeps that figure current:
// every iteration it adds 64 slabs' live bytes from the slab counters (the
// per-slab sum of find_free_mem_k, src/statistics.cuh:34-48) and publishes a
// full pass.  Per iteration a live sequence either decodes one token -- 16
// bytes {tag, t, h, ~h}, h = tok_hash(tag, t), 16 tokens to a 256 B page from
// hamalloc -- or, once it has its length, drains one page: reads its tokens
// back, adds the hashes of the intact ones to the checksum and frees the
// page.  A page malloc that returns NULL ends the sequence early; `produced`
// says how many tokens it holds.  A wave ends when its stream is exhausted and
// its lanes are idle, or at the deadline.
struct SrvState {
  u64 t0;        // wall clock at the start
  u64 free_est;  // heap free bytes from the slab counters
  u32 admitted, done, pad0, pad1;
  unsigned long long nulls, samples, pages, stalls;
};

struct SrvArgs {
  const u64* arrival;  // ticks after the start, sorted
  const u32* length;   // tokens, 1..32768
  const u32* tag;
  u32* checksum;
  u32* produced;
  u64* first_tk;       // tick of the first token
  u64* done_tk;        // tick of completion
  u32* lat_tk;         // sampled page-malloc latencies
  u32* bad;            // tokens that did not read back intact
  u64* bad_first;      // words 0 and 1 of the first such token
  void** page_tab;     // max_pages page pointers per slot
  SrvState* st;
  u64 watermark, deadline_tk;
  u32 nreq, max_pages, sample_every, lat_cap;
};

constexpr u32 SRV_NONE = 0xffffffffu;
constexpr u32 SRV_TOK_BYTES = 16, SRV_PAGE_BYTES = 256, SRV_TOK_PER_PAGE = SRV_PAGE_BYTES / SRV_TOK_BYTES;

__device__ __forceinline__ u32 tok_hash(u32 tag, u32 t) { return hash3(tag, t, 0x5E7A11C3u); }

__global__ void srv_start_k(SrvState* st) { st->t0 = wall_clock64(); }

__device__ __forceinline__ u32 wave_sum(u32 v) {
  for (int o = 32; o; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

__global__ void __launch_bounds__(256) srv_loop_k(SrvArgs a) {
  const ha::Params& P = ha::ha_p;
  SrvState* st = a.st;
  const u32 me = threadIdx.x & 63;
  const u32 slot = blockIdx.x * blockDim.x + threadIdx.x;
  const u32 wave = slot >> 6, nwaves = (gridDim.x * blockDim.x) >> 6;
  const u64 lt = me ? (~0ull >> (64 - me)) : 0ull;  // lanes below me
  void** pages = a.page_tab + (u64)slot * a.max_pages;
  const u64 t0 = __hip_atomic_load(&st->t0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  u32 req = SRV_NONE, tag = 0, len = 0, tok = 0, cs = 0, np = 0, drain = 0, nbad = 0;
  u64 first_bad = 0;
  bool draining = false;
  char* cur = nullptr;
  u32 my_pages = 0, my_nulls = 0, my_stalls = 0, my_admits = 0;
  u32 next = wave;                    // the wave's next request
  unsigned long long used_acc = 0;   // wave 0: live bytes of the slabs scanned so far
  u32 scan = 0;
  for (;;) {
    const u64 now = wall_clock64() - t0;
    if (wave == 0) {
      const u32 b = scan + me;
      if (b < P.nsbs) {
        const u32 csz = __hip_atomic_load(&P.desc[b].chunk_sz, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const u64 c = __hip_atomic_load(P.ctr + (u64)b * ha::CTR_STRIDE, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (csz) used_acc += ha::chunk_mul(ha::ctr_count(c), csz);
      }
      scan += 64;
      if (scan >= P.nsbs) {
        const unsigned long long used = wave_sum64(used_acc);
        if (me == 0) {
          const u64 fe = used < P.arena_bytes ? P.arena_bytes - used : 0;
          __hip_atomic_store(&st->free_est, fe, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        used_acc = 0;
        scan = 0;
      }
    }
    // ---- admission from the wave's own stream
    const u64 I = __ballot(req == SRV_NONE);
    if (I == ~0ull && next >= a.nreq) break;
    if (now > a.deadline_tk) break;
    bool admitted = false;
    if (I && next < a.nreq && a.arrival[next] <= now) {
      const u64 fe0 = __hip_atomic_load(&st->free_est, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      const u64 fe = ((u64)rfl32((u32)(fe0 >> 32)) << 32) | rfl32((u32)fe0);
      if (fe >= a.watermark) {
        const bool idle = req == SRV_NONE;
        const u32 rank = (u32)__builtin_popcountll(I & lt);
        const u32 cand = next + rank * nwaves;
        const bool take = idle && cand < a.nreq && a.arrival[cand] <= now;  // a prefix of the idle lanes
        const u32 k = (u32)__builtin_popcountll(__ballot(take));
        if (take) {
          req = cand;
          tag = a.tag[req];
          len = a.length[req];
          tok = cs = np = drain = nbad = 0;
          first_bad = 0;
          draining = false;
          a.first_tk[req] = now;
          ++my_admits;
        }
        next += k * nwaves;
        admitted = k != 0;
      } else {
        ++my_stalls;
      }
    }
    // ---- one step per live sequence
    bool sampled = false, finished = false;
    u32 lat = 0;
    if (req != SRV_NONE && !draining) {
      if (tok % SRV_TOK_PER_PAGE == 0) {
        const u32 pi = tok / SRV_TOK_PER_PAGE;
        const bool timed = a.sample_every && (req + pi) % a.sample_every == 0;
        const u64 c0 = wall_clock64();
        char* pg = (char*)hamalloc(SRV_PAGE_BYTES);
        __asm__ volatile("" : : "v"(pg));  // the page pointer is ready before the second clock read
        const u64 c1 = wall_clock64();
        if (pg) {
          pages[pi] = pg;
          cur = pg;
          ++my_pages;
          if (timed) {
            sampled = true;
            lat = (u32)(c1 - c0);
          }
        } else {
          ++my_nulls;
          len = tok;  // end the sequence with what it has
        }
      }
      if (tok < len) {
        const u32 h = tok_hash(tag, tok);
        *(uint4*)(cur + (tok % SRV_TOK_PER_PAGE) * SRV_TOK_BYTES) = make_uint4(tag, tok, h, ~h);
        ++tok;
      }
      if (tok >= len) {
        draining = true;
        np = (tok + SRV_TOK_PER_PAGE - 1) / SRV_TOK_PER_PAGE;
        drain = 0;
      }
    } else if (req != SRV_NONE) {
      if (drain < np) {
        const uint4* pg = (const uint4*)pages[drain];
        const u32 t_lo = drain * SRV_TOK_PER_PAGE;
        const u32 n = tok - t_lo < SRV_TOK_PER_PAGE ? tok - t_lo : SRV_TOK_PER_PAGE;
        uint4 v[SRV_TOK_PER_PAGE];
#pragma unroll
        for (u32 j = 0; j < SRV_TOK_PER_PAGE; ++j)
          if (j < n) v[j] = pg[j];
#pragma unroll
        for (u32 j = 0; j < SRV_TOK_PER_PAGE; ++j) {
          if (j < n) {
            const u32 t = t_lo + j, h = tok_hash(tag, t);
            if (v[j].x == tag && v[j].y == t && v[j].z == h && v[j].w == ~h) {
              cs += h;
            } else {
              if (!nbad) first_bad = (u64)v[j].x | ((u64)v[j].y << 32);
              ++nbad;
            }
          }
        }
        hafree((void*)pg);
        ++drain;
      }
      if (drain >= np) {
        a.checksum[req] = cs;
        a.produced[req] = tok;
        a.bad[req] = nbad;
        a.bad_first[req] = first_bad;
        a.done_tk[req] = now;
        req = SRV_NONE;
        finished = true;
      }
    }
    // ---- wave-aggregated counters
    const u64 F = __ballot(finished);
    if (F && me == 0) atomicAdd(&st->done, (u32)__builtin_popcountll(F));
    const u64 S = __ballot(sampled);
    if (S) {
      unsigned long long base = 0;
      if (me == 0) base = atomicAdd(&st->samples, (unsigned long long)__builtin_popcountll(S));
      base = ((unsigned long long)rfl32((u32)(base >> 32)) << 32) | rfl32((u32)base);
      if (sampled) {
        const unsigned long long i = base + (u64)__builtin_popcountll(S & lt);
        if (i < a.lat_cap) a.lat_tk[i] = lat;
      }
    }
    if (!admitted && I == ~0ull) __builtin_amdgcn_s_sleep(4);
  }
  const u32 wp = wave_sum(my_pages), wn = wave_sum(my_nulls), wa = wave_sum(my_admits);
  if (me == 0) {
    if (wp) atomicAdd(&st->pages, (unsigned long long)wp);
    if (wn) atomicAdd(&st->nulls, (unsigned long long)wn);
    if (wa) atomicAdd(&st->admitted, wa);
    if (my_stalls) atomicAdd(&st->stalls, (unsigned long long)my_stalls);
  }
}

// ------------------------------------------------------------ test harness
// The reference test programs' device side (tst/common/common.h:213-416):
// active threads (tid < n and not (tid >> group_sh) & period_mask), sizes
// from the uniform / exponential-uniform / exponential-equal distributions,
// and the two-state alloc/free action chain.  The reference draws from a
// per-thread state array; here every draw is a hash of (seed, tid, draw
// index), so a run is reproducible.
struct TOpts {
  u32 n, nallocs, niters, alloc_sz, max_alloc_sz, max_alloc_sh, distr, period_mask, group_sh, seed;
  float p[2][2];  // p[phase][allocated]: the probability of acting
};
__device__ __forceinline__ bool t_inactive(const TOpts& o, u32 i) {
  return i >= o.n || ((i >> o.group_sh) & o.period_mask);
}
// drandom(a, b) of tst/common/common.h:96-99 on a hashed draw
__device__ __forceinline__ u32 t_between(u32 r, u32 a, u32 b) { return a + (r & 0x00ffffffu) % (b - a + 1); }
__device__ __forceinline__ u32 t_draw(const TOpts& o, u32 i, u32 k) { return hash3(o.seed, i, k); }
// next_alloc_sz (tst/common/common.h:236-276); k numbers the thread's draws
__device__ u32 t_size(const TOpts& o, u32 i, u32 k) {
  if (o.alloc_sz == o.max_alloc_sz) return o.alloc_sz;
  const u32 r1 = t_draw(o, i, 2 * k), r2 = t_draw(o, i, 2 * k + 1);
  u32 sh = 0;
  if (o.distr == 1) return t_between(r1, o.alloc_sz, o.max_alloc_sz);  // uniform
  if (o.distr == 2) sh = t_between(r1, 0, o.max_alloc_sh);             // exponential-uniform
  else sh = (u32)__builtin_ffs((int)t_between(r1, 1, 1u << (o.max_alloc_sh + 1))) - 1;  // exponential-equal
  const u32 sz = t_between(r2, o.alloc_sz << sh, (o.alloc_sz << (sh + 1)) - 1);
  return sz < o.max_alloc_sz ? sz : o.max_alloc_sz;
}
// drandomb(p) of tst/common/common.h:109-114 (drandomf: 20-bit steps)
__device__ __forceinline__ bool t_bernoulli(float p, u32 r) {
  if (0.0f < p && p < 1.0f) return (1.0f / (1024.0f * 1024.0f)) * (float)t_between(r, 0, 1024u * 1024u - 1) <= p;
  return p >= 1.0f;
}
// stats: mallocs, frees, nulls
__device__ __forceinline__ void t_stats(unsigned long long* st, u32 m, u32 f, u32 nn) {
  const unsigned long long a = wave_sum64(m), b = wave_sum64(f), c = wave_sum64(nn);
  if ((threadIdx.x & 63) == 0) {
    if (a) atomicAdd(&st[0], a);
    if (b) atomicAdd(&st[1], b);
    if (c) atomicAdd(&st[2], c);
  }
}
// malloc_k / malloc_corr_k (tst/common/common.h:347-372): stamp writes the size into the block
__global__ void t_malloc_k(TOpts o, void** ptrs, u32 try_id, int stamp, unsigned long long* st) {
  const u32 i = blockIdx.x * blockDim.x + threadIdx.x;
  u32 m = 0, nn = 0;
  if (!t_inactive(o, i)) {
    for (u32 a = 0; a < o.nallocs; ++a) {
      const u32 sz = t_size(o, i, try_id * o.nallocs + a);
      void* q = hamalloc(sz);
      ptrs[i + (size_t)o.n * a] = q;
      if (q && stamp) *(u32*)q = sz;
      m += q != nullptr;
      nn += q == nullptr;
    }
  }
  t_stats(st, m, 0, nn);
}
// free_k (tst/common/common.h:375-383)
__global__ void t_free_k(TOpts o, void** ptrs, unsigned long long* st) {
  const u32 i = blockIdx.x * blockDim.x + threadIdx.x;
  u32 f = 0;
  if (!t_inactive(o, i)) {
    for (u32 a = 0; a < o.nallocs; ++a) {
      void*& q = ptrs[i + (size_t)o.n * a];
      f += q != nullptr;
      hafree(q);
      q = nullptr;
    }
  }
  t_stats(st, 0, f, 0);
}
// throughput_malloc_free_k (tst/perf/priv-throughput/priv-throughput.cu:14-27)
__global__ void t_priv_k(TOpts o, void** ptrs, u32 try_id, unsigned long long* st) {
  const u32 i = blockIdx.x * blockDim.x + threadIdx.x;
  u32 m = 0, f = 0, nn = 0;
  if (!t_inactive(o, i)) {
    for (u32 it = 0; it < o.niters; ++it) {
      for (u32 a = 0; a < o.nallocs; ++a) {
        void* q = hamalloc(t_size(o, i, (try_id * o.niters + it) * o.nallocs + a));
        ptrs[i + (size_t)o.n * a] = q;
        m += q != nullptr;
        nn += q == nullptr;
      }
      for (u32 a = 0; a < o.nallocs; ++a) {
        void*& q = ptrs[i + (size_t)o.n * a];
        f += q != nullptr;
        hafree(q);
        q = nullptr;
      }
    }
  }
  t_stats(st, m, f, nn);
}
// prob_corr_k (tst/corr/prob-checkptr/prob-checkptr.cu:13-45): per iteration a
// thread in state allocated/free acts with p[phase][state], phase alternating
// with (try * niters + iter); ctrs[i] = blocks the thread holds.  stamp: 0
// none, 1 the size into the first word, 2 the fill of phase-alloc-write
__global__ void t_prob_k(TOpts o, void** ptrs, u32* ctrs, u32 try_id, int stamp, unsigned long long* st) {
  const u32 i = blockIdx.x * blockDim.x + threadIdx.x;
  u32 m = 0, f = 0, nn = 0;
  if (!t_inactive(o, i)) {
    u32 ctr = ctrs[i];
    for (u32 it = 0; it < o.niters; ++it) {
      const u32 step = try_id * o.niters + it, phase = step % 2;
      const u32 r = t_draw(o, i, 0x80000000u | step);
      if (!t_bernoulli(o.p[phase][ctr > 0], r)) continue;
      if (ctr == 0) {
        for (u32 a = 0; a < o.nallocs; ++a) {
          const u32 sz = t_size(o, i, step * o.nallocs + a);
          void* q = hamalloc(sz);
          ptrs[i + (size_t)o.n * a] = q;
          if (q && stamp == 1) *(u32*)q = sz;
          if (q && stamp == 2)  // phase-alloc-write.cu:33-37: 123 into every whole 8 B word
            for (u32 w = 0; w < sz / 8; ++w) ((unsigned long long*)q)[w] = 123ull;
          m += q != nullptr;
          nn += q == nullptr;
        }
        ctr = o.nallocs;
      } else {
        for (u32 a = 0; a < o.nallocs; ++a) {
          void*& q = ptrs[i + (size_t)o.n * a];
          f += q != nullptr;
          hafree(q);
          q = nullptr;
        }
        ctr = 0;
      }
    }
    ctrs[i] = ctr;
  }
  t_stats(st, m, f, nn);
}
// free_rest_k (tst/common/common.h:387-398)
__global__ void t_free_rest_k(TOpts o, void** ptrs, u32* ctrs, unsigned long long* st) {
  const u32 i = blockIdx.x * blockDim.x + threadIdx.x;
  u32 f = 0;
  if (!t_inactive(o, i)) {
    for (u32 a = 0; a < ctrs[i]; ++a) {
      void*& q = ptrs[i + (size_t)o.n * a];
      f += q != nullptr;
      hafree(q);
      q = nullptr;
    }
    ctrs[i] = 0;
  }
  t_stats(st, 0, f, 0);
}
// Latencies in shader-clock cycles (tst/perf/latency/latency.cu:14-42,
// tst/perf/phase-latency/phase-latency.cu:18-63), summed per wave before one
// lane adds them: L = {count, sum, min, max} for mallocs, then for frees.
__device__ __forceinline__ unsigned long long wave_min64(unsigned long long v) {
  for (int o = 32; o; o >>= 1) {
    const unsigned long long w = __shfl_xor(v, o);
    v = w < v ? w : v;
  }
  return v;
}
__device__ __forceinline__ unsigned long long wave_max64(unsigned long long v) {
  for (int o = 32; o; o >>= 1) {
    const unsigned long long w = __shfl_xor(v, o);
    v = w > v ? w : v;
  }
  return v;
}
// every lane of the wave calls it; has: this lane timed an operation of kind k
__device__ __forceinline__ void lat_add(unsigned long long* L, u32 k, bool has, unsigned long long cyc) {
  const unsigned long long c = wave_sum64(has), sm = wave_sum64(has ? cyc : 0ull);
  const unsigned long long mn = wave_min64(has ? cyc : ~0ull), mx = wave_max64(has ? cyc : 0ull);
  if ((threadIdx.x & 63) == 0 && c) {
    atomicAdd(&L[4 * k], c);
    atomicAdd(&L[4 * k + 1], sm);
    atomicMin(&L[4 * k + 2], mn);
    atomicMax(&L[4 * k + 3], mx);
  }
}
// latency test: every active thread's mallocs, then (second launch) its frees
__global__ void t_lat_k(TOpts o, void** ptrs, u32 try_id, int do_free, unsigned long long* L,
                        unsigned long long* st) {
  const u32 i = blockIdx.x * blockDim.x + threadIdx.x;
  const bool act = !t_inactive(o, i);
  u32 m = 0, f = 0, nn = 0;
  for (u32 a = 0; a < o.nallocs; ++a) {
    unsigned long long cyc = 0;
    if (act) {
      void*& q = ptrs[i + (size_t)o.n * a];
      if (do_free) {
        f += q != nullptr;
        const unsigned long long t0 = clock64();
        hafree(q);
        cyc = clock64() - t0;
        q = nullptr;
      } else {
        const u32 sz = t_size(o, i, try_id * o.nallocs + a);
        const unsigned long long t0 = clock64();
        q = hamalloc(sz);
        cyc = clock64() - t0;
        m += q != nullptr;
        nn += q == nullptr;
      }
    }
    lat_add(L, do_free ? 1u : 0u, act, cyc);
  }
  t_stats(st, m, f, nn);
}
// phase-latency: the chain of t_prob_k with every malloc and free timed
__global__ void t_phase_lat_k(TOpts o, void** ptrs, u32* ctrs, u32 try_id, unsigned long long* L,
                              unsigned long long* st) {
  const u32 i = blockIdx.x * blockDim.x + threadIdx.x;
  const bool act = !t_inactive(o, i);
  u32 ctr = act ? ctrs[i] : 0u, m = 0, f = 0, nn = 0;
  for (u32 it = 0; it < o.niters; ++it) {
    const u32 step = try_id * o.niters + it, phase = step % 2;
    const bool go = act && t_bernoulli(o.p[phase][ctr > 0], t_draw(o, i, 0x80000000u | step));
    const bool alloc = go && ctr == 0, release = go && ctr > 0;  // both imply an active thread
    for (u32 a = 0; a < o.nallocs; ++a) {
      unsigned long long cyc = 0;
      if (alloc) {
        const u32 sz = t_size(o, i, step * o.nallocs + a);
        const unsigned long long t0 = clock64();
        void* q = hamalloc(sz);
        cyc = clock64() - t0;
        ptrs[i + (size_t)o.n * a] = q;
        m += q != nullptr;
        nn += q == nullptr;
      } else if (release) {
        void*& q = ptrs[i + (size_t)o.n * a];
        f += q != nullptr;
        const unsigned long long t0 = clock64();
        hafree(q);
        cyc = clock64() - t0;
        q = nullptr;
      }
      lat_add(L, 0u, alloc, cyc);
      lat_add(L, 1u, release, cyc);
    }
    if (alloc) ctr = o.nallocs;
    if (release) ctr = 0;
  }
  if (act) ctrs[i] = ctr;
  t_stats(st, m, f, nn);
}

// the stamped size of every non-NULL pointer (0 for NULL)
// check_alloc (tst/common/common.cu:412-529) on the device.  Every pointer an
// active thread holds must be non-NULL and inside the pool, and its
// [p, p + size) claims the 8-byte granules it covers in a bitmap over the
// pool (size: the fixed request, or the size stamped in the block's first
// word); a granule found claimed already is an overlap.  The reference sorts
// the pointers and compares neighbours: both flag exactly the sets of held
// blocks in which two overlap.  Block starts are 8-byte aligned (24 B chunks),
// so blocks that do not overlap never share a granule.  st[0] counts the
// blocks checked, st[1] the violations.
__global__ void t_check_k(TOpts o, void* const* ptrs, const u32* ctrs, int with_ctrs, int stamped, u64* gran,
                          const char* base, u64 pool_bytes, unsigned long long* st) {
  const u64 idx = blockIdx.x * (u64)blockDim.x + threadIdx.x;
  const u32 i = (u32)(idx % o.n), k = (u32)(idx / o.n);
  const bool held = idx < (u64)o.n * o.nallocs && !t_inactive(o, i) && k < (with_ctrs ? ctrs[i] : o.nallocs);
  bool bad = false;
  if (held) {
    const char* p = (const char*)ptrs[idx];
    if (!p) {
      bad = true;
    } else {
      const u64 sz = stamped ? *(const u32*)p : o.alloc_sz;
      const u64 off = (u64)(p - base);
      if (off >= pool_bytes || off + sz > pool_bytes) {
        bad = true;
      } else {
        for (u64 g = off / 8, g1 = (off + sz + 7) / 8; g < g1;) {
          const u64 w = g / 64, lo = g % 64, n = g1 - g < 64 - lo ? g1 - g : 64 - lo;
          const u64 m = (n >= 64 ? ~0ull : ((1ull << n) - 1)) << lo;
          if (atomicOr((unsigned long long*)(gran + w), (unsigned long long)m) & m) bad = true;
          g += n;
        }
      }
    }
  }
  // one counter update per wave
  const u64 hm = __ballot(held), bm = __ballot(bad);
  if ((threadIdx.x & 63) == 0) {
    if (hm) atomicAdd(&st[0], (unsigned long long)__builtin_popcountll(hm));
    if (bm) atomicAdd(&st[1], (unsigned long long)__builtin_popcountll(bm));
  }
}



// ---------------------------------------------------------------- samples
// samples/grid-points/grid-points.cu:61-133: every point gets a 16 B list
// node from hamalloc, pushed onto its grid cell's list with an atomic
// exchange; the lists are then walked and freed.  Points are uniform in
// [0, 1)^3 in steps of 2^-16 (the sample's random() % 65536 / 65536), from a
// hash.
struct PointNode {
  u32 ip, pad;
  PointNode* next;
};
__device__ __forceinline__ u32 point_coord(u32 seed, u32 ip, u32 axis) { return hash3(seed, ip, axis) & 0xFFFFu; }
__device__ __forceinline__ u32 point_cell(u32 seed, u32 ip, u32 nc) {
  u32 c[3];
  for (u32 a = 0; a < 3; ++a) {
    const u32 k = (u32)floorf((float)point_coord(seed, ip, a) * (1.0f / 65536.0f) * (float)nc);
    c[a] = k < nc ? k : nc - 1;
  }
  return c[0] + nc * (c[1] + nc * c[2]);
}
__global__ void grid_insert_k(PointNode** grid, u32 nc, u32 n, u32 seed, unsigned long long* st) {
  const u32 ip = blockIdx.x * blockDim.x + threadIdx.x;
  if (ip >= n) return;
  PointNode* q = (PointNode*)hamalloc(sizeof(PointNode));
  if (!q) {
    atomicAdd(&st[0], 1ull);
    return;
  }
  q->ip = ip;
  q->next = (PointNode*)atomicExch((unsigned long long*)&grid[point_cell(seed, ip, nc)], (unsigned long long)q);
}
// a node pointer is followed only inside the arena (at most n steps per list)
__device__ __forceinline__ bool in_arena(const void* q) {
  return (u64)((const char*)q - ha::ha_p.arena) + sizeof(PointNode) <= ha::ha_p.arena_bytes;
}
// every point is in the list of its own cell, exactly once
__global__ void grid_check_k(PointNode* const* grid, u32 nc, u32 n, u32 seed, u32* seen, unsigned long long* st) {
  const u32 c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nc * nc * nc) return;
  u32 cnt = 0, bad = 0;
  for (const PointNode* q = grid[c]; q && cnt <= n; ++cnt) {
    if (!in_arena(q)) {
      ++bad;
      break;
    }
    const u32 ip = q->ip;
    if (ip >= n || point_cell(seed, ip, nc) != c || atomicAdd(&seen[ip], 1u) != 0) ++bad;
    q = q->next;
  }
  atomicAdd(&st[1], (unsigned long long)bad);
  atomicAdd(&st[2], (unsigned long long)cnt);
}
__global__ void grid_free_k(PointNode** grid, u32 nc, u32 n) {
  const u32 c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nc * nc * nc) return;
  PointNode* q = grid[c];
  for (u32 k = 0; q && k <= n && in_arena(q); ++k) {
    PointNode* nx = q->next;
    hafree(q);
    q = nx;
  }
  grid[c] = nullptr;
}

// samples/random-graph/random-graph.cu:56-181: vertex v gets U{1..max_degree}
// edges, each a 16 B list node {target, next} from hamalloc pushed onto its
// own list, targets uniform over the vertices; the lists are then checked
// (length = degree, targets in range) and freed (the sample leaves them).
struct EdgeNode {
  const void* target;
  EdgeNode* next;
};
struct Vertex {
  u32 id, nedges;
  EdgeNode* edges;
};
__global__ void rgraph_build_k(Vertex* vs, u32 nv, u32 max_degree, u32 seed, unsigned long long* st) {
  const u32 v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= nv) return;
  const u32 ne = 1 + hash3(seed, v, 0xE0u) % max_degree;
  Vertex x{v, 0u, nullptr};
  for (u32 e = 0; e < ne; ++e) {
    EdgeNode* q = (EdgeNode*)hamalloc(sizeof(EdgeNode));
    if (!q) {
      atomicAdd(&st[0], 1ull);
      continue;
    }
    q->target = vs + hash3(seed, v, e) % nv;
    q->next = x.edges;
    x.edges = q;
    x.nedges++;
  }
  vs[v] = x;
}
__global__ void rgraph_check_free_k(Vertex* vs, u32 nv, u32 max_degree, u32 seed, int check, unsigned long long* st) {
  const u32 v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= nv) return;
  const u32 ne = 1 + hash3(seed, v, 0xE0u) % max_degree;
  u32 cnt = 0, bad = 0;
  EdgeNode* q = vs[v].edges;
  // pushed last to first: node k of the walk is edge (nedges - 1 - k); a
  // node pointer is followed only inside the arena
  for (; q && cnt < ne; ++cnt) {
    if ((u64)((char*)q - ha::ha_p.arena) + sizeof(EdgeNode) > ha::ha_p.arena_bytes) {
      ++bad;
      break;
    }
    EdgeNode* nx = q->next;
    if (check && q->target != vs + hash3(seed, v, vs[v].nedges - 1 - cnt) % nv) ++bad;
    if (!check) hafree(q);
    q = nx;
  }
  if (check && (q || cnt != vs[v].nedges || vs[v].id != v)) ++bad;
  if (bad) atomicAdd(&st[1], 1ull);
}

// samples/add-strings/add-strings.cu:71-248: two arrays of n strings are
// allocated and filled, concatenated elementwise into a third (its length
// measured, its block from hamalloc), and all three freed.  Lengths are
// uniform in [min_len, max_len] (31 in the sample); the characters follow
// the string and position, so the concatenation can be checked.
__device__ __forceinline__ u32 str_len(u32 seed, u32 which, u32 i, u32 lo, u32 hi) {
  return lo + hash3(seed ^ which, i, 77u) % (hi - lo + 1);
}
__device__ __forceinline__ char str_char(u32 which, u32 i, u32 j) { return (char)('a' + (i * 7u + j + which * 13u) % 26u); }
__global__ void str_alloc_k(char** strs, u32 which, u32 n, u32 lo, u32 hi, u32 seed, unsigned long long* st) {
  const u32 i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const u32 l = str_len(seed, which, i, lo, hi);
  u64* w = (u64*)hamalloc(l + 1);
  strs[i] = (char*)w;
  if (!w) {
    atomicAdd(&st[0], 1ull);
    return;
  }
  // 8 bytes at a time (blocks are 8-byte aligned and a multiple of 8 long)
  for (u32 j0 = 0; j0 <= l; j0 += 8) {
    u64 v = 0;
    for (u32 b = 0; b < 8; ++b) {
      const u32 j = j0 + b;
      const u64 ch = j < l ? (u64)(unsigned char)str_char(which, i, j) : 0ull;
      v |= ch << (8 * b);
    }
    w[j0 / 8] = v;
  }
}
__global__ void str_add_k(char** c, char* const* a, char* const* b, u32 n, u32 max_len, unsigned long long* st) {
  const u32 i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || !a[i] || !b[i]) {
    if (i < n) c[i] = nullptr;
    return;
  }
  u32 la = 0, lb = 0;
  while (la < max_len && a[i][la]) ++la;
  while (lb < max_len && b[i][lb]) ++lb;
  char* s = (char*)hamalloc(la + lb + 1);
  c[i] = s;
  if (!s) {
    atomicAdd(&st[0], 1ull);
    return;
  }
  for (u32 j = 0; j < la; ++j) s[j] = a[i][j];
  for (u32 j = 0; j < lb; ++j) s[la + j] = b[i][j];
  s[la + lb] = 0;
}
__global__ void str_check_k(char* const* c, u32 n, u32 lo, u32 hi, u32 seed, unsigned long long* st) {
  const u32 i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || !c[i]) return;
  const u32 la = str_len(seed, 0, i, lo, hi), lb = str_len(seed, 1, i, lo, hi);
  u32 bad = 0;
  for (u32 j = 0; j < la; ++j) bad += c[i][j] != str_char(0, i, j);
  for (u32 j = 0; j < lb; ++j) bad += c[i][la + j] != str_char(1, i, j);
  bad += c[i][la + lb] != 0;
  if (bad) atomicAdd(&st[1], 1ull);
}
__global__ void str_free_k(char** s, u32 n) {
  const u32 i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  hafree(s[i]);
  s[i] = nullptr;
}

// Typed objects through new (halloc) / hadelete: an object whose constructor
// and destructor count themselves, and an array of words, per thread.
struct CountedObj {
  u32 id;
  unsigned long long key;
  unsigned long long* census;  // [0] constructed, [1] destroyed
  __device__ CountedObj(u32 i, u32 seed, unsigned long long* c) : id(i), key(0), census(c) {
    key = ((unsigned long long)hash3(seed, i, 7) << 32) | hash3(seed, i, 8);
    atomicAdd(&census[0], 1ull);
  }
  __device__ ~CountedObj() { atomicAdd(&census[1], 1ull); }
};
__device__ __forceinline__ u32 obj_words(u32 seed, u32 i, u32 hi) { return 1 + hash3(seed, i, 9) % hi; }
__global__ void obj_new_k(CountedObj** objs, u32** arrs, u32 n, u32 hi, u32 seed, unsigned long long* st) {
  const u32 i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  CountedObj* o = new (halloc) CountedObj(i, seed, st + 2);
  const u32 k = obj_words(seed, i, hi);
  u32* w = new (halloc) u32[k];
  objs[i] = o;
  arrs[i] = w;
  if (!o) atomicAdd(&st[0], 1ull);
  if (!w) {
    atomicAdd(&st[0], 1ull);
    return;
  }
  for (u32 j = 0; j < k; ++j) w[j] = i * 131u + j;
}
__global__ void obj_check_delete_k(CountedObj** objs, u32** arrs, u32 n, u32 hi, u32 seed, unsigned long long* st) {
  const u32 i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  CountedObj* o = objs[i];
  u32* w = arrs[i];
  const u32 k = obj_words(seed, i, hi);
  u32 bad = 0;
  if (o) {
    bad += o->id != i || o->key != (((unsigned long long)hash3(seed, i, 7) << 32) | hash3(seed, i, 8)) ||
           o->census != st + 2 || ((uintptr_t)o & 7);
  }
  if (w)
    for (u32 j = 0; j < k; ++j) bad += w[j] != i * 131u + j;
  if (o && w && (uintptr_t)o < (uintptr_t)(w + k) && (uintptr_t)w < (uintptr_t)(o + 1)) ++bad;
  if (bad) atomicAdd(&st[1], 1ull);
  hadelete(o);
  hadelete_array(w);
  objs[i] = nullptr;
  arrs[i] = nullptr;
}

}  // namespace

extern "C" {

int ha_replay(const uint32_t* ops, int nsteps, int lanes, const uint32_t* iheads, uint32_t* recs,
              uint32_t* xcc) {
  if (!ha::host_params().initialized) return hfail(HA_E_STATE, "allocator not initialized");
  if (!ops || !recs || !xcc || nsteps < 0 || lanes < 1 || lanes > 64) return hfail(HA_E_INVALID, "bad replay arguments");
  const size_t nops = (size_t)nsteps * lanes;
  for (size_t i = 0; i < nops; ++i) {
    if (ops[2 * i] > 3) return hfail(HA_E_INVALID, "bad op kind");
    if (ops[2 * i] >= 2 && (ops[2 * i + 1] >= nops || ops[2 * i + 1] / lanes >= i / lanes))
      return hfail(HA_E_INVALID, "a free must name a malloc of an earlier step");
  }
  DBuf dops, drecs, dxcc, dptrs, dih;
  if (iheads) {
    HH(dih.alloc((size_t)nsteps * 4));
    HH(hipMemcpy(dih.p, iheads, (size_t)nsteps * 4, hipMemcpyHostToDevice));
  }
  HH(dops.alloc(nops * 8));
  HH(drecs.alloc(nops * 16));
  HH(dxcc.alloc((size_t)nsteps * 4));
  HH(dptrs.alloc(nops * 8));
  HH(hipMemcpy(dops.p, ops, nops * 8, hipMemcpyHostToDevice));
  HH(hipMemset(dptrs.p, 0, nops * 8));
  HH(hipMemset(drecs.p, 0, nops * 16));
  replay_k<<<1, lanes>>>((const u32*)dops.p, nsteps, lanes, (const u32*)dih.p, (u32*)drecs.p, (u32*)dxcc.p,
                         (void**)dptrs.p);
  HH(hipGetLastError());
  HH(hipDeviceSynchronize());
  HH(hipMemcpy(recs, drecs.p, nops * 16, hipMemcpyDeviceToHost));
  HH(hipMemcpy(xcc, dxcc.p, (size_t)nsteps * 4, hipMemcpyDeviceToHost));
  return HA_OK;
}

int64_t ha_dump_words(void) {
  const ha::Params& P = ha::host_params();
  if (!P.initialized) return hfail(HA_E_STATE, "allocator not initialized");
  return (int64_t)dump_words_of(P);
}

int ha_dump(uint64_t* out, int64_t nwords) {
  const ha::Params& P = ha::host_params();
  if (!P.initialized) return hfail(HA_E_STATE, "allocator not initialized");
  const u64 total = dump_words_of(P);
  if (!out || nwords < (int64_t)total) return hfail(HA_E_INVALID, "dump buffer too small");
  DBuf d;
  HH(d.alloc(total * 8));
  HH(hipDeviceSynchronize());
  dump_k<<<1024, 256>>>((u64*)d.p, total);
  HH(hipGetLastError());
  HH(hipMemcpy(out, d.p, total * 8, hipMemcpyDeviceToHost));
  return HA_OK;
}

int ha_run_plain(int nthreads, int bs, int size, int iters, int warmup, int launches, float* ms,
                 ha_plain_result* res) {
  if (!ha::host_params().initialized) return hfail(HA_E_STATE, "allocator not initialized");
  if (nthreads < 1 || bs < 1 || bs > 256 || size < 16 || size > 3072 || iters < 1 ||
      warmup < 0 || launches < 1 || !res)
    return hfail(HA_E_INVALID, "bad plain-run arguments");
  DBuf dc;
  HH(dc.alloc(16));
  HH(hipMemset(dc.p, 0, 16));
  Stream s;
  HH(hipStreamCreate(&s.s));
  // ev[0], ev[1]: the timed region; ev[2 + 2l], ev[3 + 2l]: launch l of the
  // per-launch pass
  Events ev;
  for (int i = 0; i < 2 + (ms ? 2 * launches : 0); ++i) {
    hipEvent_t e;
    HH(hipEventCreate(&e));
    ev.e.push_back(e);
  }
  const int grid = (nthreads + bs - 1) / bs;
  unsigned long long* c = (unsigned long long*)dc.p;
  for (int w = 0; w < warmup; ++w) plain_k<<<grid, bs, 0, s.s>>>(nthreads, size, iters, c, c + 1);
  HH(hipGetLastError());
  HH(hipStreamSynchronize(s.s));
  HH(hipMemsetAsync(dc.p, 0, 16, s.s));
  HH(hipStreamSynchronize(s.s));
  // the timed launches are queued back to back, as a caller's stream of
  // launches runs: no host synchronization and no event between them (an
  // event pair around every launch lengthened each step by about 6 us)
  const auto w0 = std::chrono::steady_clock::now();
  HH(hipEventRecord(ev.e[0], s.s));
  for (int l = 0; l < launches; ++l) plain_k<<<grid, bs, 0, s.s>>>(nthreads, size, iters, c, c + 1);
  HH(hipEventRecord(ev.e[1], s.s));
  HH(hipStreamSynchronize(s.s));
  const auto w1 = std::chrono::steady_clock::now();
  HH(hipGetLastError());
  float total = 0;
  HH(hipEventElapsedTime(&total, ev.e[0], ev.e[1]));
  if (ms) {
    // the per-launch pass, after the timed region
    for (int l = 0; l < launches; ++l) {
      HH(hipEventRecord(ev.e[2 + 2 * l], s.s));
      plain_k<<<grid, bs, 0, s.s>>>(nthreads, size, iters, c, c + 1);
      HH(hipEventRecord(ev.e[3 + 2 * l], s.s));
    }
    HH(hipStreamSynchronize(s.s));
    HH(hipGetLastError());
    for (int l = 0; l < launches; ++l) HH(hipEventElapsedTime(&ms[l], ev.e[2 + 2 * l], ev.e[3 + 2 * l]));
  }
  unsigned long long h[2];
  HH(hipMemcpy(h, dc.p, 16, hipMemcpyDeviceToHost));
  res->errors = h[0];
  res->nulls = h[1];
  res->total_ms = total;
  res->wall_ms = std::chrono::duration<double, std::milli>(w1 - w0).count();
  return HA_OK;
}

int ha_malloc_sizes(const uint32_t* sizes, int n, int bs, uint64_t* ptrs) {
  if (!ha::host_params().initialized) return hfail(HA_E_STATE, "allocator not initialized");
  if (!sizes || !ptrs || n < 0 || bs < 1 || bs > 1024) return hfail(HA_E_INVALID, "bad arguments");
  if (n == 0) return HA_OK;
  DBuf ds, dp;
  HH(ds.alloc((size_t)n * 4));
  HH(dp.alloc((size_t)n * 8));
  HH(hipMemcpy(ds.p, sizes, (size_t)n * 4, hipMemcpyHostToDevice));
  malloc_sizes_k<<<(n + bs - 1) / bs, bs>>>((const u32*)ds.p, n, (void**)dp.p);
  HH(hipGetLastError());
  HH(hipDeviceSynchronize());
  HH(hipMemcpy(ptrs, dp.p, (size_t)n * 8, hipMemcpyDeviceToHost));
  return HA_OK;
}

int ha_run_corr_test0(int nthreads, int nmallocs, int bs, ha_corr0_result* r) {
  if (!ha::host_params().initialized) return hfail(HA_E_STATE, "allocator not initialized");
  if (!r || nthreads < 1 || nmallocs < 1 || bs < 64 || bs > 1024 || nthreads % bs ||
      (u64)nthreads * (u64)nmallocs > (1ull << 28))
    return hfail(HA_E_INVALID, "bad arguments");
  std::memset(r, 0, sizeof *r);
  const u64 n = (u64)nthreads * (u64)nmallocs;
  DBuf dp, dc;
  HH(dp.alloc(n * 8));
  HH(dc.alloc(16));
  HH(hipMemset(dp.p, 0, n * 8));
  HH(hipMemset(dc.p, 0, 16));
  unsigned long long* cnt = (unsigned long long*)dc.p;
  hipEvent_t e1, e2;
  HH(hipEventCreate(&e1));
  HH(hipEventCreate(&e2));
  t0_malloc_k<<<nthreads / bs, bs>>>((void**)dp.p, nmallocs, cnt);
  HH(hipGetLastError());
  const unsigned g = (unsigned)((n + 255) / 256);
  t0_write_k<<<g, 256>>>((void* const*)dp.p, n);
  HH(hipGetLastError());
  t0_read_k<<<g, 256>>>((void* const*)dp.p, n, cnt + 1);
  HH(hipGetLastError());
  HH(hipDeviceSynchronize());
  // sorted, consecutive blocks at least 16 B apart (test.cu:103-121)
  std::vector<uint64_t> hp(n);
  HH(hipMemcpy(hp.data(), dp.p, n * 8, hipMemcpyDeviceToHost));
  std::sort(hp.begin(), hp.end());
  for (u64 i = 0; i + 1 < n; ++i)
    if (hp[i] && hp[i + 1] - hp[i] < 16) ++r->bad_gaps;
  HH(hipEventRecord(e1, nullptr));
  t0_free_k<<<nthreads / bs, bs>>>((void* const*)dp.p, nmallocs);
  HH(hipGetLastError());
  HH(hipEventRecord(e2, nullptr));
  HH(hipDeviceSynchronize());
  float ms = 0;
  HH(hipEventElapsedTime(&ms, e1, e2));
  r->free_ms = ms;
  unsigned long long h[2];
  HH(hipMemcpy(h, dc.p, 16, hipMemcpyDeviceToHost));
  r->blocks = n;
  r->nulls = h[0];
  r->bad_contents = h[1];
  (void)hipEventDestroy(e1);
  (void)hipEventDestroy(e2);
  return HA_OK;
}

int ha_free_ptrs(const uint64_t* ptrs, int n, int bs) {
  if (!ha::host_params().initialized) return hfail(HA_E_STATE, "allocator not initialized");
  if (!ptrs || n < 0 || bs < 1 || bs > 1024) return hfail(HA_E_INVALID, "bad arguments");
  if (n == 0) return HA_OK;
  DBuf dp;
  HH(dp.alloc((size_t)n * 8));
  HH(hipMemcpy(dp.p, ptrs, (size_t)n * 8, hipMemcpyHostToDevice));
  free_ptrs_k<<<(n + bs - 1) / bs, bs>>>((void* const*)dp.p, n);
  HH(hipGetLastError());
  HH(hipDeviceSynchronize());
  return HA_OK;
}

int ha_read_words(const uint64_t* ptrs, int n, uint32_t* out) {
  if (!ha::host_params().initialized) return hfail(HA_E_STATE, "allocator not initialized");
  if (!ptrs || !out || n < 0) return hfail(HA_E_INVALID, "bad arguments");
  if (n == 0) return HA_OK;
  DBuf dp, dout;
  HH(dp.alloc((size_t)n * 8));
  HH(dout.alloc((size_t)n * 4));
  HH(hipMemcpy(dp.p, ptrs, (size_t)n * 8, hipMemcpyHostToDevice));
  read_words_k<<<(n + 255) / 256, 256>>>((void* const*)dp.p, n, (u32*)dout.p);
  HH(hipGetLastError());
  HH(hipMemcpy(out, dout.p, (size_t)n * 4, hipMemcpyDeviceToHost));
  return HA_OK;
}

int ha_run_touch(int n, int words, int iters, int reps, double* ms, uint64_t* sums, uint64_t* nulls) {
  if (!ha::host_params().initialized) return hfail(HA_E_STATE, "allocator not initialized");
  if (n < 1 || words < 1 || words > 384 || iters < 1 || reps < 1 || !ms || !sums || !nulls)
    return hfail(HA_E_INVALID, "bad touch arguments");
  const int bs = 256, grid = (n + bs - 1) / bs;
  DBuf dp, ds, dn;
  HH(dp.alloc((size_t)n * 8));
  HH(ds.alloc((size_t)n * 8));
  HH(dn.alloc(8));
  HH(hipMemset(dn.p, 0, 8));
  Stream s;
  HH(hipStreamCreate(&s.s));
  hipEvent_t e0, e1;
  HH(hipEventCreate(&e0));
  HH(hipEventCreate(&e1));
  u64** ptrs = (u64**)dp.p;
  u64* sink = (u64*)ds.p;
  touch_alloc_k<<<grid, bs, 0, s.s>>>(ptrs, n, words, (unsigned long long*)dn.p);
  HH(hipGetLastError());
  std::vector<u64> h((size_t)n);
  // stores first, so that the loads and updates start from known values
  const int order[3] = {2, 0, 1};
  for (int m : order) {
    touch_k<<<grid, bs, 0, s.s>>>(ptrs, n, words, iters, m, sink);  // untimed first launch
    HH(hipEventRecord(e0, s.s));
    for (int r = 0; r < reps; ++r) touch_k<<<grid, bs, 0, s.s>>>(ptrs, n, words, iters, m, sink);
    HH(hipEventRecord(e1, s.s));
    HH(hipEventSynchronize(e1));
    float t = 0;
    HH(hipEventElapsedTime(&t, e0, e1));
    ms[m] = (double)t / reps;
    touch_sum_k<<<grid, bs, 0, s.s>>>(ptrs, n, words, m, sink);
    HH(hipMemcpyAsync(h.data(), sink, (size_t)n * 8, hipMemcpyDeviceToHost, s.s));
    HH(hipStreamSynchronize(s.s));
    u64 sum = 0;
    for (u64 v : h) sum += v;
    sums[m] = sum;
  }
  touch_free_k<<<grid, bs, 0, s.s>>>(ptrs, n);
  HH(hipGetLastError());
  HH(hipMemcpyAsync(nulls, dn.p, 8, hipMemcpyDeviceToHost, s.s));
  HH(hipStreamSynchronize(s.s));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return HA_OK;
}

int ha_probe_atomics(double* out) {
  if (!out) return hfail(HA_E_INVALID, "null output");
  const size_t BYTES = 1ull << 30;
  const int NT = 1 << 20, BS = 256;
  DBuf buf, sink;
  HH(buf.alloc(BYTES));
  HH(sink.alloc((size_t)NT * 8));
  HH(hipMemset(buf.p, 0, BYTES));
  hipEvent_t e0, e1;
  HH(hipEventCreate(&e0));
  HH(hipEventCreate(&e1));
  const u64 mask = BYTES / 8 - 1;
  for (int shape = 0; shape < 3; ++shape) {
    const int iters = shape == 2 ? 1 : 16;
    float best = 1e30f;
    for (int rep = 0; rep < 3; ++rep) {
      HH(hipEventRecord(e0, 0));
      probe_k<<<NT / BS, BS>>>((u64*)buf.p, mask, iters, shape, (u64*)sink.p);
      HH(hipEventRecord(e1, 0));
      HH(hipEventSynchronize(e1));
      float t = 0;
      HH(hipEventElapsedTime(&t, e0, e1));
      best = t < best ? t : best;
    }
    out[shape] = (double)NT * iters / (best * 1e-3);
  }
  HH(hipGetLastError());
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return HA_OK;
}


int ha_probe_chain(const int* rounds, int n, int nthreads, double* ms) {
  if (!rounds || !ms || n < 1 || n > 64 || nthreads < 256 || nthreads > (1 << 24) || nthreads % 256)
    return hfail(HA_E_INVALID, "bad arguments");
  for (int i = 0; i < n; ++i)
    if (rounds[i] < 0 || rounds[i] > 4096) return hfail(HA_E_INVALID, "rounds must be in 0..4096");
  const size_t BYTES = 1ull << 30;
  DBuf buf, sink;
  HH(buf.alloc(BYTES));
  HH(sink.alloc((size_t)nthreads / 64 * 8));
  HH(hipMemset(buf.p, 0, BYTES));
  hipEvent_t e0, e1;
  HH(hipEventCreate(&e0));
  HH(hipEventCreate(&e1));
  for (int i = 0; i < n; ++i) {
    float best = 1e30f;
    for (int rep = 0; rep < 5; ++rep) {
      HH(hipEventRecord(e0, 0));
      chain_k<<<nthreads / 256, 256>>>((u64*)buf.p, BYTES / 8 - 1, rounds[i], (u64*)sink.p);
      HH(hipEventRecord(e1, 0));
      HH(hipEventSynchronize(e1));
      float t = 0;
      HH(hipEventElapsedTime(&t, e0, e1));
      best = t < best ? t : best;
    }
    ms[i] = best;
  }
  HH(hipGetLastError());
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return HA_OK;
}

int ha_run_test(const ha_test_opts* a, ha_test_result* r) {
  if (!ha::host_params().initialized) return hfail(HA_E_STATE, "allocator not initialized");
  if (!a || !r || a->test < 0 || a->test > 8 || a->nthreads < 1 || a->nthreads > (1 << 26) || a->bs < 1 ||
      a->bs > 1024 || a->ntries < 1 || a->nallocs < 1 || a->nallocs > 64 || a->niters < 1 ||
      a->max_alloc_sz < a->alloc_sz || a->distr < 1 || a->distr > 3 || a->group_sh > 31 || a->max_alloc_sh > 24 ||
      ((uint64_t)a->alloc_sz << (a->max_alloc_sh + 1)) > 0xFFFFFFFFull ||
      (uint64_t)a->nthreads * (uint64_t)a->nallocs > (1ull << 28))
    return hfail(HA_E_INVALID, "bad test options");
  const u32 n = (u32)a->nthreads;
  const size_t nptrs = (size_t)n * (u32)a->nallocs;
  TOpts o{};
  o.n = n; o.nallocs = (u32)a->nallocs; o.niters = (u32)a->niters;
  o.alloc_sz = a->alloc_sz; o.max_alloc_sz = a->max_alloc_sz; o.max_alloc_sh = a->max_alloc_sh;
  o.distr = a->distr; o.period_mask = a->period_mask; o.group_sh = a->group_sh; o.seed = a->seed;
  for (int x = 0; x < 2; ++x)
    for (int y = 0; y < 2; ++y) o.p[x][y] = a->probabs[x][y];
  DBuf dp, dc, ds, dl;
  HH(dl.alloc(8 * 8));
  {
    const unsigned long long l0[8] = {0, 0, ~0ull, 0, 0, 0, ~0ull, 0};
    HH(hipMemcpy(dl.p, l0, sizeof l0, hipMemcpyHostToDevice));
  }
  unsigned long long* L = (unsigned long long*)dl.p;
  HH(dp.alloc(nptrs * 8));
  HH(dc.alloc((size_t)n * 4));
  HH(ds.alloc(3 * 8));
  HH(hipMemset(dp.p, 0, nptrs * 8));
  HH(hipMemset(dc.p, 0, (size_t)n * 4));
  HH(hipMemset(ds.p, 0, 3 * 8));
  const bool stamped = a->test == 0 || a->test == 1;
  Stream s;
  HH(hipStreamCreate(&s.s));
  hipEvent_t e0, e1;
  HH(hipEventCreate(&e0));
  HH(hipEventCreate(&e1));
  // acc: kernel time from events; the wall-clock time around launch and
  // synchronization (what the reference's tests report, omp_get_wtime) goes
  // to the matching *_wall_ms field
  auto timed = [&](auto&& launch, double& acc) -> int {
    const auto w0 = std::chrono::steady_clock::now();
    HH(hipEventRecord(e0, s.s));
    launch();
    HH(hipGetLastError());
    HH(hipEventRecord(e1, s.s));
    HH(hipStreamSynchronize(s.s));
    const auto w1 = std::chrono::steady_clock::now();
    float t = 0;
    HH(hipEventElapsedTime(&t, e0, e1));
    acc += t;
    (&acc == &r->free_ms ? r->free_wall_ms : r->malloc_wall_ms) +=
        std::chrono::duration<double, std::milli>(w1 - w0).count();
    return HA_OK;
  };
  std::memset(r, 0, sizeof *r);
  r->failed_try = -1;
  std::vector<void*> hp(nptrs);
  std::vector<u32> hc(n);
  // check_alloc on the device (t_check_k): a granule bitmap over the pool,
  // cleared before each check; only the two counts come back
  const ha::Params& HP = ha::host_params();
  const u64 pool_bytes = HP.arena_bytes + HP.large_bytes;
  DBuf dgran, dchk;
  if (stamped) {
    HH(dgran.alloc((pool_bytes / 8 + 63) / 64 * 8));
    HH(dchk.alloc(16));
  }
  auto check = [&](bool with_ctrs) -> int {
    HH(hipMemsetAsync(dgran.p, 0, (pool_bytes / 8 + 63) / 64 * 8, s.s));
    HH(hipMemsetAsync(dchk.p, 0, 16, s.s));
    t_check_k<<<(unsigned)((nptrs + 255) / 256), 256, 0, s.s>>>(o, (void* const*)dp.p, (const u32*)dc.p,
                                                              with_ctrs ? 1 : 0, o.alloc_sz != o.max_alloc_sz,
                                                              (u64*)dgran.p, HP.arena, pool_bytes,
                                                              (unsigned long long*)dchk.p);
    HH(hipGetLastError());
    unsigned long long c[2];
    HH(hipMemcpyAsync(c, dchk.p, 16, hipMemcpyDeviceToHost, s.s));
    HH(hipStreamSynchronize(s.s));
    r->checked += c[0];
    r->checks_failed += c[1];
    return HA_OK;
  };
  const unsigned grid = (n + (u32)a->bs - 1) / (u32)a->bs;
  unsigned long long* st = (unsigned long long*)ds.p;
  for (int t = 0; t < a->ntries; ++t) {
    const uint64_t before = r->checks_failed;
    switch (a->test) {
      case 0:  // checkptr
        if (int rc = timed([&] { t_malloc_k<<<grid, a->bs, 0, s.s>>>(o, (void**)dp.p, (u32)t, 1, st); }, r->malloc_ms))
          return rc;
        if (int rc = check(false)) return rc;
        if (int rc = timed([&] { t_free_k<<<grid, a->bs, 0, s.s>>>(o, (void**)dp.p, st); }, r->free_ms)) return rc;
        break;
      case 1:  // prob-checkptr
        if (int rc = timed([&] { t_prob_k<<<grid, a->bs, 0, s.s>>>(o, (void**)dp.p, (u32*)dc.p, (u32)t, 1, st); },
                           r->malloc_ms))
          return rc;
        if (int rc = check(true)) return rc;
        break;
      case 4:    // phase-throughput: even tries count as malloc time, odd as free time; check_nz only
      case 7:    // phase-alloc-write: the same, each block filled
      case 8: {  // phase-extfrag: the same, ha_extfrag after each try
        const int fill = a->test == 7 ? 2 : 0;
        if (int rc = timed([&] { t_prob_k<<<grid, a->bs, 0, s.s>>>(o, (void**)dp.p, (u32*)dc.p, (u32)t, fill, st); },
                           t % 2 ? r->free_ms : r->malloc_ms))
          return rc;
        if (a->test == 8 && a->extfrag) {
          if (int rc = ha_extfrag_c(0, &a->extfrag[2 * t])) return rc;
          if (int rc = ha_extfrag_c(1, &a->extfrag[2 * t + 1])) return rc;
        }
        HH(hipMemcpy(hp.data(), dp.p, nptrs * 8, hipMemcpyDeviceToHost));
        HH(hipMemcpy(hc.data(), dc.p, (size_t)n * 4, hipMemcpyDeviceToHost));
        for (u32 i = 0; i < n; ++i) {
          if ((i >> o.group_sh) & o.period_mask) continue;
          for (u32 k = 0; k < hc[i]; ++k) r->checks_failed += hp[i + (size_t)n * k] == nullptr;
        }
        break;
      }
      case 2: {  // throughput: check_nz only
        if (int rc = timed([&] { t_malloc_k<<<grid, a->bs, 0, s.s>>>(o, (void**)dp.p, (u32)t, 0, st); }, r->malloc_ms))
          return rc;
        HH(hipMemcpy(hp.data(), dp.p, nptrs * 8, hipMemcpyDeviceToHost));
        for (u32 i = 0; i < n; ++i) {
          if ((i >> o.group_sh) & o.period_mask) continue;
          for (u32 k = 0; k < o.nallocs; ++k) r->checks_failed += hp[i + (size_t)n * k] == nullptr;
        }
        if (int rc = timed([&] { t_free_k<<<grid, a->bs, 0, s.s>>>(o, (void**)dp.p, st); }, r->free_ms)) return rc;
        break;
      }
      case 5:  // latency: mallocs timed, NULL check, frees timed
        if (int rc = timed([&] { t_lat_k<<<grid, a->bs, 0, s.s>>>(o, (void**)dp.p, (u32)t, 0, L, st); }, r->malloc_ms))
          return rc;
        HH(hipMemcpy(hp.data(), dp.p, nptrs * 8, hipMemcpyDeviceToHost));
        for (u32 i = 0; i < n; ++i) {
          if ((i >> o.group_sh) & o.period_mask) continue;
          for (u32 k = 0; k < o.nallocs; ++k) r->checks_failed += hp[i + (size_t)n * k] == nullptr;
        }
        if (int rc = timed([&] { t_lat_k<<<grid, a->bs, 0, s.s>>>(o, (void**)dp.p, (u32)t, 1, L, st); }, r->free_ms))
          return rc;
        break;
      case 6:  // phase-latency, check_nz of the blocks held
        if (int rc = timed([&] { t_phase_lat_k<<<grid, a->bs, 0, s.s>>>(o, (void**)dp.p, (u32*)dc.p, (u32)t, L, st); },
                           r->malloc_ms))
          return rc;
        HH(hipMemcpy(hp.data(), dp.p, nptrs * 8, hipMemcpyDeviceToHost));
        HH(hipMemcpy(hc.data(), dc.p, (size_t)n * 4, hipMemcpyDeviceToHost));
        for (u32 i = 0; i < n; ++i) {
          if ((i >> o.group_sh) & o.period_mask) continue;
          for (u32 k = 0; k < hc[i]; ++k) r->checks_failed += hp[i + (size_t)n * k] == nullptr;
        }
        break;
      default:  // priv-throughput
        if (int rc = timed([&] { t_priv_k<<<grid, a->bs, 0, s.s>>>(o, (void**)dp.p, (u32)t, st); }, r->malloc_ms))
          return rc;
        break;
    }
    r->tries_run = t + 1;
    if (r->checks_failed != before && r->failed_try < 0) r->failed_try = t;
  }
  if (a->test == 1 || a->test == 6) {
    t_free_rest_k<<<grid, a->bs, 0, s.s>>>(o, (void**)dp.p, (u32*)dc.p, st);
    HH(hipGetLastError());
  } else if (a->test == 4 || a->test == 7 || a->test == 8) {  // the rest counts as free time
    if (int rc = timed([&] { t_free_rest_k<<<grid, a->bs, 0, s.s>>>(o, (void**)dp.p, (u32*)dc.p, st); }, r->free_ms))
      return rc;
  }
  HH(hipStreamSynchronize(s.s));
  unsigned long long h[3];
  HH(hipMemcpy(h, ds.p, sizeof h, hipMemcpyDeviceToHost));
  r->mallocs = h[0];
  r->frees = h[1];
  r->nulls = h[2];
  unsigned long long l[8];
  HH(hipMemcpy(l, dl.p, sizeof l, hipMemcpyDeviceToHost));
  for (int k = 0; k < 2; ++k) {
    r->lat_n[k] = l[4 * k];
    r->lat_avg[k] = l[4 * k] ? (double)l[4 * k + 1] / (double)l[4 * k] : 0.0;
    r->lat_min[k] = l[4 * k] ? (double)l[4 * k + 2] : 0.0;
    r->lat_max[k] = (double)l[4 * k + 3];
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return HA_OK;
}


int ha_run_sample(const ha_sample_args* a, ha_sample_result* r) {
  if (!ha::host_params().initialized) return hfail(HA_E_STATE, "allocator not initialized");
  if (!a || !r || a->sample < 0 || a->sample > 3 || a->n < 1 || a->n > (1 << 26) || a->bs < 64 || a->bs > 1024 ||
      (a->sample == 0 && (a->ncells < 1 || a->ncells > 512)) ||
      (a->sample == 1 && (a->min_len < 0 || a->max_len < a->min_len || a->max_len > 4096)) ||
      (a->sample == 2 && (a->max_len < 1 || a->max_len > 64)) || (a->sample == 3 && (a->max_len < 1 || a->max_len > 768)))
    return hfail(HA_E_INVALID, "bad sample arguments");
  const u32 n = (u32)a->n;
  Stream s;
  HH(hipStreamCreate(&s.s));
  hipEvent_t e[5];  // alloc, work, free: e0-e1, e1-e2, e3-e4
  for (auto& x : e) HH(hipEventCreate(&x));
  DBuf dst;
  HH(dst.alloc(4 * 8));
  HH(hipMemset(dst.p, 0, 4 * 8));
  unsigned long long* st = (unsigned long long*)dst.p;
  const unsigned grid = (n + (u32)a->bs - 1) / (u32)a->bs;
  std::memset(r, 0, sizeof *r);
  float t[3] = {0, 0, 0};
  if (a->sample == 0) {
    const u32 nc = (u32)a->ncells, ncells3 = nc * nc * nc;
    DBuf dg, dseen;
    HH(dg.alloc((size_t)ncells3 * 8));
    HH(dseen.alloc((size_t)n * 4));
    HH(hipMemset(dg.p, 0, (size_t)ncells3 * 8));
    HH(hipMemset(dseen.p, 0, (size_t)n * 4));
    const unsigned cgrid = (ncells3 + (u32)a->bs - 1) / (u32)a->bs;
    HH(hipEventRecord(e[0], s.s));
    grid_insert_k<<<grid, a->bs, 0, s.s>>>((PointNode**)dg.p, nc, n, a->seed, st);
    HH(hipEventRecord(e[1], s.s));
    grid_check_k<<<cgrid, a->bs, 0, s.s>>>((PointNode* const*)dg.p, nc, n, a->seed, (u32*)dseen.p, st);
    HH(hipEventRecord(e[2], s.s));
    HH(hipEventRecord(e[3], s.s));
    grid_free_k<<<cgrid, a->bs, 0, s.s>>>((PointNode**)dg.p, nc, n);
    HH(hipEventRecord(e[4], s.s));
    HH(hipGetLastError());
    HH(hipStreamSynchronize(s.s));
  } else if (a->sample == 3) {  // objects: max_len is the largest array in words
    DBuf dobj, darr;
    HH(dobj.alloc((size_t)n * 8));
    HH(darr.alloc((size_t)n * 8));
    HH(hipEventRecord(e[0], s.s));
    obj_new_k<<<grid, a->bs, 0, s.s>>>((CountedObj**)dobj.p, (u32**)darr.p, n, (u32)a->max_len, a->seed, st);
    HH(hipEventRecord(e[1], s.s));
    HH(hipEventRecord(e[2], s.s));
    HH(hipEventRecord(e[3], s.s));
    obj_check_delete_k<<<grid, a->bs, 0, s.s>>>((CountedObj**)dobj.p, (u32**)darr.p, n, (u32)a->max_len, a->seed,
                                                st);
    HH(hipEventRecord(e[4], s.s));
    HH(hipGetLastError());
    HH(hipStreamSynchronize(s.s));
  } else if (a->sample == 2) {  // random-graph: max_len is the largest degree
    DBuf dv;
    HH(dv.alloc((size_t)n * sizeof(Vertex)));
    HH(hipEventRecord(e[0], s.s));
    rgraph_build_k<<<grid, a->bs, 0, s.s>>>((Vertex*)dv.p, n, (u32)a->max_len, a->seed, st);
    HH(hipEventRecord(e[1], s.s));
    rgraph_check_free_k<<<grid, a->bs, 0, s.s>>>((Vertex*)dv.p, n, (u32)a->max_len, a->seed, 1, st);
    HH(hipEventRecord(e[2], s.s));
    HH(hipEventRecord(e[3], s.s));
    rgraph_check_free_k<<<grid, a->bs, 0, s.s>>>((Vertex*)dv.p, n, (u32)a->max_len, a->seed, 0, st);
    HH(hipEventRecord(e[4], s.s));
    HH(hipGetLastError());
    HH(hipStreamSynchronize(s.s));
  } else {
    DBuf da, db, dc;
    HH(da.alloc((size_t)n * 8));
    HH(db.alloc((size_t)n * 8));
    HH(dc.alloc((size_t)n * 8));
    const u32 lo = (u32)a->min_len, hi = (u32)a->max_len;
    HH(hipEventRecord(e[0], s.s));
    str_alloc_k<<<grid, a->bs, 0, s.s>>>((char**)da.p, 0, n, lo, hi, a->seed, st);
    str_alloc_k<<<grid, a->bs, 0, s.s>>>((char**)db.p, 1, n, lo, hi, a->seed, st);
    HH(hipEventRecord(e[1], s.s));
    str_add_k<<<grid, a->bs, 0, s.s>>>((char**)dc.p, (char* const*)da.p, (char* const*)db.p, n, hi, st);
    HH(hipEventRecord(e[2], s.s));
    str_check_k<<<grid, a->bs, 0, s.s>>>((char* const*)dc.p, n, lo, hi, a->seed, st);  // not timed
    HH(hipEventRecord(e[3], s.s));
    str_free_k<<<grid, a->bs, 0, s.s>>>((char**)da.p, n);
    str_free_k<<<grid, a->bs, 0, s.s>>>((char**)db.p, n);
    str_free_k<<<grid, a->bs, 0, s.s>>>((char**)dc.p, n);
    HH(hipEventRecord(e[4], s.s));
    HH(hipGetLastError());
    HH(hipStreamSynchronize(s.s));
  }
  HH(hipEventElapsedTime(&t[0], e[0], e[1]));
  HH(hipEventElapsedTime(&t[1], e[1], e[2]));
  HH(hipEventElapsedTime(&t[2], e[3], e[4]));
  unsigned long long h[4];
  HH(hipMemcpy(h, dst.p, sizeof h, hipMemcpyDeviceToHost));
  r->nulls = h[0];
  r->errors = h[1] + (a->sample == 0 && h[2] + h[0] != n ? 1 : 0);  // grid: every point listed once
  // objects: every object constructed was destroyed once
  if (a->sample == 3 && h[2] != h[3]) r->errors += 1;
  r->alloc_ms = t[0];
  r->work_ms = t[1];
  r->free_ms = t[2];
  for (auto& x : e) (void)hipEventDestroy(x);
  return HA_OK;
}

int ha_probe_copy(uint64_t bytes, int reps, double* gbps) {
  if (!gbps || reps < 1 || bytes < (1u << 20) || bytes % 16 || bytes > (64ull << 30))
    return hfail(HA_E_INVALID, "bad copy probe arguments");
  DBuf a, b;
  HH(a.alloc(bytes));
  HH(b.alloc(bytes));
  HH(hipMemset(a.p, 1, bytes));
  HH(hipMemset(b.p, 0, bytes));
  int dev = 0, ncu = 0;
  HH(hipGetDevice(&dev));
  HH(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  const u64 n16 = bytes / 16;
  const int grid = ncu * 8;
  Stream s;
  HH(hipStreamCreate(&s.s));
  hipEvent_t e0, e1;
  HH(hipEventCreate(&e0));
  HH(hipEventCreate(&e1));
  float best = 1e30f;
  for (int r = 0; r < reps; ++r) {
    HH(hipEventRecord(e0, s.s));
    copy_k<<<grid, 256, 0, s.s>>>((const uint4*)a.p, (uint4*)b.p, n16);
    HH(hipEventRecord(e1, s.s));
    HH(hipEventSynchronize(e1));
    float t = 0;
    HH(hipEventElapsedTime(&t, e0, e1));
    best = t < best ? t : best;
  }
  HH(hipGetLastError());
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  *gbps = 2.0 * (double)bytes / (best * 1e-3) / 1e9;
  return HA_OK;
}

int ha_run_lists(const uint32_t* sizes, int n, int per, const uint32_t* order, int bs, uint64_t* ptrs_out,
                 ha_list_result* res) {
  if (!ha::host_params().initialized) return hfail(HA_E_STATE, "allocator not initialized");
  if (!sizes || n < 1 || per < 1 || bs < 1 || bs > 1024 || !res) return hfail(HA_E_INVALID, "bad arguments");
  const size_t total = (size_t)n * per;
  if (order) {
    std::vector<uint8_t> seen(total, 0);
    for (size_t k = 0; k < total; ++k) {
      if (order[k] >= total || seen[order[k]]) return hfail(HA_E_INVALID, "order is not a permutation");
      seen[order[k]] = 1;
    }
  }
  DBuf ds, dp, dord, dc;
  HH(ds.alloc(total * 4));
  HH(dp.alloc(total * 8));
  HH(dc.alloc(16));
  HH(hipMemcpy(ds.p, sizes, total * 4, hipMemcpyHostToDevice));
  if (order) {
    HH(dord.alloc(total * 4));
    HH(hipMemcpy(dord.p, order, total * 4, hipMemcpyHostToDevice));
  }
  HH(hipMemset(dc.p, 0, 16));
  Stream s;
  HH(hipStreamCreate(&s.s));
  hipEvent_t e0, e1, e2;
  HH(hipEventCreate(&e0));
  HH(hipEventCreate(&e1));
  HH(hipEventCreate(&e2));
  unsigned long long* c = (unsigned long long*)dc.p;
  const int grid = (n + bs - 1) / bs;
  HH(hipStreamSynchronize(s.s));
  const auto w0 = std::chrono::steady_clock::now();
  HH(hipEventRecord(e0, s.s));
  list_malloc_k<<<grid, bs, 0, s.s>>>((const u32*)ds.p, n, per, (void**)dp.p, c);
  HH(hipEventRecord(e1, s.s));
  HH(hipStreamSynchronize(s.s));
  const auto w1 = std::chrono::steady_clock::now();
  if (ptrs_out) HH(hipMemcpy(ptrs_out, dp.p, total * 8, hipMemcpyDeviceToHost));
  HH(hipStreamSynchronize(s.s));
  const auto w2 = std::chrono::steady_clock::now();
  list_free_k<<<grid, bs, 0, s.s>>>((void**)dp.p, (const u32*)ds.p, (const u32*)dord.p, n, per, c + 1);
  HH(hipEventRecord(e2, s.s));
  HH(hipStreamSynchronize(s.s));
  const auto w3 = std::chrono::steady_clock::now();
  HH(hipGetLastError());
  float tm = 0, tf = 0;
  HH(hipEventElapsedTime(&tm, e0, e1));
  HH(hipEventElapsedTime(&tf, e1, e2));
  unsigned long long h[2];
  HH(hipMemcpy(h, dc.p, 16, hipMemcpyDeviceToHost));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipEventDestroy(e2);
  res->nulls = h[0];
  res->errors = h[1];
  res->malloc_ms = tm;
  res->free_ms = tf;
  res->wall_ms = std::chrono::duration<double, std::milli>((w1 - w0) + (w3 - w2)).count();
  uint64_t nreq = 0;
  for (size_t k = 0; k < total; ++k) nreq += sizes[k] != 0;
  res->requests = nreq;
  return HA_OK;
}

int ha_run_phase(const ha_phase_args* a, ha_phase_result* res, double* trace, int trace_cap) {
  const ha::Params& P = ha::host_params();
  if (!P.initialized) return hfail(HA_E_STATE, "allocator not initialized");
  if (!a || !res || a->nthreads < 1 || a->bs < 64 || a->bs > 1024 || a->bs % 64 || a->slots < 1 || a->nphases < 1 ||
      a->min_size < 8 || a->max_size > 3072 || a->min_size > a->max_size || a->max_steps < 1)
    return hfail(HA_E_INVALID, "bad phase arguments");
  const int n = a->nthreads, slots = a->slots;
  const size_t total = (size_t)n * slots;
  DBuf dp, ds, dc;
  HH(dp.alloc(total * 8));
  HH(ds.alloc(total * 4));
  HH(dc.alloc(8 * 8));
  HH(hipMemset(dp.p, 0, total * 8));
  HH(hipMemset(dc.p, 0, 8 * 8));
  Stream s;
  HH(hipStreamCreate(&s.s));
  hipEvent_t e0, e1;
  HH(hipEventCreate(&e0));
  HH(hipEventCreate(&e1));
  unsigned long long* c = (unsigned long long*)dc.p;
  const int grid = (n + a->bs - 1) / a->bs;
  const double heap = (double)P.arena_bytes;
  std::memset(res, 0, sizeof *res);
  unsigned long long h[6] = {0, 0, 0, 0, 0, 0}, prev[6] = {0, 0, 0, 0, 0, 0};
  double kern = 0, wall = 0;
  int ntrace = 0;
  u32 step = 0;
  for (int ph = 0; ph < a->nphases; ++ph) {
    const bool alloc_phase = (ph % 2) == 0;
    int st = 0;
    for (; st < a->max_steps; ++st, ++step) {
      HH(hipStreamSynchronize(s.s));
      const auto w0 = std::chrono::steady_clock::now();
      HH(hipEventRecord(e0, s.s));
      phase_step_k<<<grid, a->bs, 0, s.s>>>((void**)dp.p, (u32*)ds.p, n, slots, a->seed, step,
                                              alloc_phase ? a->p_hi : a->p_lo, a->min_size / 8, a->max_size / 8, c);
      HH(hipEventRecord(e1, s.s));
      HH(hipStreamSynchronize(s.s));
      const auto w1 = std::chrono::steady_clock::now();
      float t = 0;
      HH(hipEventElapsedTime(&t, e0, e1));
      kern += t;
      wall += std::chrono::duration<double, std::milli>(w1 - w0).count();
      HH(hipMemcpy(h, dc.p, sizeof h, hipMemcpyDeviceToHost));
      const double fill = ((double)h[4] - (double)h[5]) / heap;
      const unsigned long long dm = h[0] - prev[0], dn = h[2] - prev[2];
      std::memcpy(prev, h, sizeof h);
      if (trace && ntrace < trace_cap) {
        trace[3 * ntrace] = ph;
        trace[3 * ntrace + 1] = fill;
        trace[3 * ntrace + 2] = (double)dn;
        ntrace++;
      }
      if (alloc_phase && (fill >= a->fill_hi || dn * 100 > dm + dn)) { ++st; ++step; break; }
      if (!alloc_phase && fill <= a->fill_lo) { ++st; ++step; break; }
    }
    res->phase_steps[ph < 8 ? ph : 7] += st;
  }
  // free what is left
  HH(hipStreamSynchronize(s.s));
  const auto w0 = std::chrono::steady_clock::now();
  HH(hipEventRecord(e0, s.s));
  phase_rest_k<<<grid, a->bs, 0, s.s>>>((void**)dp.p, (u32*)ds.p, n, slots, c);
  HH(hipEventRecord(e1, s.s));
  HH(hipStreamSynchronize(s.s));
  const auto w1 = std::chrono::steady_clock::now();
  float t = 0;
  HH(hipEventElapsedTime(&t, e0, e1));
  kern += t;
  wall += std::chrono::duration<double, std::milli>(w1 - w0).count();
  HH(hipGetLastError());
  HH(hipMemcpy(h, dc.p, sizeof h, hipMemcpyDeviceToHost));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  res->mallocs = h[0];
  res->frees = h[1];
  res->nulls = h[2];
  res->errors = h[3];
  res->kernel_ms = kern;
  res->wall_ms = wall;
  res->steps = (int)step;
  res->trace_len = ntrace;
  return HA_OK;
}

int ha_replay_batch(const uint32_t* ops, int nsteps, int lanes, uint32_t* recs, uint32_t* xcc) {
  if (!ha::host_params().initialized) return hfail(HA_E_STATE, "allocator not initialized");
  if (!ops || !recs || !xcc || nsteps < 0 || lanes < 1 || lanes > 64) return hfail(HA_E_INVALID, "bad replay arguments");
  const size_t nops = (size_t)nsteps * lanes;
  for (size_t i = 0; i < nops; ++i) {
    const uint32_t kind = ops[4 * i], arg = ops[4 * i + 1];
    if (kind > 3) return hfail(HA_E_INVALID, "bad op kind");
    if ((kind == 2 || kind == 3) && (arg >= nops || arg / lanes >= i / lanes || ops[4 * (size_t)arg] != 1))
      return hfail(HA_E_INVALID, "a batch free must name a batch malloc of an earlier step");
  }
  DBuf dops, drecs, dxcc, dptrs;
  HH(dops.alloc(nops * 16));
  HH(drecs.alloc(nops * ha::MAXB * 16));
  HH(dxcc.alloc((size_t)nsteps * 4));
  HH(dptrs.alloc(nops * ha::MAXB * 8));
  HH(hipMemcpy(dops.p, ops, nops * 16, hipMemcpyHostToDevice));
  HH(hipMemset(dptrs.p, 0, nops * ha::MAXB * 8));
  HH(hipMemset(drecs.p, 0, nops * ha::MAXB * 16));
  replay_batch_k<<<1, lanes>>>((const u32*)dops.p, nsteps, lanes, (u32*)drecs.p, (u32*)dxcc.p, (void**)dptrs.p);
  HH(hipGetLastError());
  HH(hipDeviceSynchronize());
  HH(hipMemcpy(recs, drecs.p, nops * ha::MAXB * 16, hipMemcpyDeviceToHost));
  HH(hipMemcpy(xcc, dxcc.p, (size_t)nsteps * 4, hipMemcpyDeviceToHost));
  return HA_OK;
}

int ha_run_graphs(const ha_graph_args* a, const uint32_t* nnodes, ha_graph_result* res) {
  if (!ha::host_params().initialized) return hfail(HA_E_STATE, "allocator not initialized");
  if (!a || !nnodes || !res || a->njobs < 1 || a->njobs > (1 << 30) || a->bs < 64 || a->bs > 1024 || a->bs % 64 ||
      (a->batch != 4 && a->batch != 8 && a->batch != 16) || a->waves < 1 || a->plain_chunk_nodes < 1 ||
      a->repeats < 1)
    return hfail(HA_E_INVALID, "bad graph arguments");
  // the batched kernel keeps header offsets in 32 bits of 16 B units
  if (a->run_batch && ha::host_params().arena_bytes > (64ull << 30))
    return hfail(HA_E_INVALID, "the batched graph form supports arenas up to 64 GiB");
  const u32 njobs = (u32)a->njobs;
  std::vector<u64> offs(njobs + 1, 0);
  for (u32 j = 0; j < njobs; ++j) {
    if (nnodes[j] < 1) return hfail(HA_E_INVALID, "a job needs at least one node");
    offs[j + 1] = offs[j] + nnodes[j];
  }
  const u64 total = offs[njobs];
  DBuf doffs, dn, dtab, dstats;
  HH(doffs.alloc((njobs + 1) * 8));
  HH(dn.alloc(njobs * 4));
  HH(dtab.alloc(total * 8));
  HH(dstats.alloc(16));
  HH(hipMemcpy(doffs.p, offs.data(), (njobs + 1) * 8, hipMemcpyHostToDevice));
  HH(hipMemcpy(dn.p, nnodes, njobs * 4, hipMemcpyHostToDevice));
  HH(hipMemset(dstats.p, 0, 16));
  Stream s;
  HH(hipStreamCreate(&s.s));
  hipEvent_t e0, e1;
  HH(hipEventCreate(&e0));
  HH(hipEventCreate(&e1));
  unsigned long long* st = (unsigned long long*)dstats.p;
  std::memset(res, 0, sizeof *res);
  res->nodes = total;
  // plain per-call form: chunks of jobs whose nodes fit plain_chunk_nodes
  std::vector<std::pair<u64, u64>> chunks;
  for (u32 j = 0; j < njobs;) {
    u32 k = j;
    while (k < njobs && (k == j || offs[k + 1] - offs[j] <= (u64)a->plain_chunk_nodes)) ++k;
    chunks.push_back({offs[j], offs[k]});
    j = k;
  }
  auto run_plain = [&]() -> int {
    for (auto& c : chunks) {
      const u64 nn = c.second - c.first;
      const unsigned grid = (unsigned)((nn + a->bs - 1) / a->bs);
      graph_plain_build_k<<<grid, a->bs, 0, s.s>>>((const u64*)doffs.p, (const u32*)dn.p, njobs, a->seed, c.first,
                                                    c.second, (NodeHdr**)dtab.p, st);
      graph_plain_free_k<<<grid, a->bs, 0, s.s>>>((const u64*)doffs.p, (const u32*)dn.p, njobs, a->seed, c.first,
                                                   c.second, (NodeHdr**)dtab.p, st);
    }
    return (int)hipGetLastError();
  };
  // batched persistent form, captured in a graph with its counters reset
  const u32 per_chunk = 64u * (u32)a->batch;
  std::vector<u32> coffs(njobs + 1, 0);
  for (u32 j = 0; j < njobs; ++j) coffs[j + 1] = coffs[j] + (nnodes[j] + per_chunk - 1) / per_chunk;
  const u32 nitems = coffs[njobs];
  // the item sequence: builds in job order, each job's frees `lag` items after its last build
  const u64 lag = (u64)a->waves * GRAPH_GRAB;
  std::vector<u64> seq;
  seq.reserve(2 * (size_t)nitems);
  u32 jf = 0;  // next job whose frees are due
  for (u64 b = 0; b < (u64)nitems + lag; ++b) {
    if (b < nitems) {
      const u32 j = (u32)(std::upper_bound(coffs.begin(), coffs.end(), (u32)b) - coffs.begin()) - 1;
      seq.push_back(((u64)j << 32) | (b - coffs[j]));
    }
    while (jf < njobs && (u64)coffs[jf + 1] - 1 + lag <= b) {
      for (u32 c = 0; c < coffs[jf + 1] - coffs[jf]; ++c) seq.push_back((1ull << 63) | ((u64)jf << 32) | c);
      ++jf;
    }
  }
  while (jf < njobs) {
    for (u32 c = 0; c < coffs[jf + 1] - coffs[jf]; ++c) seq.push_back((1ull << 63) | ((u64)jf << 32) | c);
    ++jf;
  }
  const u32 nseq = (u32)seq.size();
  DBuf dcoffs, dseq, dnext, dbuilt;
  HH(dcoffs.alloc((njobs + 1) * 4));
  HH(dseq.alloc((size_t)nseq * 8));
  HH(dnext.alloc(4));
  HH(dbuilt.alloc(njobs * 4));
  HH(hipMemcpy(dcoffs.p, coffs.data(), (njobs + 1) * 4, hipMemcpyHostToDevice));
  HH(hipMemcpy(dseq.p, seq.data(), (size_t)nseq * 8, hipMemcpyHostToDevice));
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  const unsigned bgrid = (unsigned)((a->waves * 64 + 255) / 256);
  HH(hipStreamBeginCapture(s.s, hipStreamCaptureModeThreadLocal));
  graph_reset_k<<<16, 256, 0, s.s>>>((u32*)dnext.p, (u32*)dbuilt.p, njobs);
#define HA_GRAPH_BATCH(BB)                                                                                  \
  graph_batch_k<BB><<<bgrid, 256, 0, s.s>>>((const u64*)dseq.p, nseq, (const u32*)dcoffs.p, (const u64*)doffs.p, \
                                            (const u32*)dn.p, a->seed, (NodeHdr**)dtab.p, (u32*)dnext.p,        \
                                            (u32*)dbuilt.p, st)
  if (a->batch == 4)
    HA_GRAPH_BATCH(4);
  else if (a->batch == 8)
    HA_GRAPH_BATCH(8);
  else
    HA_GRAPH_BATCH(16);
#undef HA_GRAPH_BATCH
  HH(hipStreamEndCapture(s.s, &graph));
  HH(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
  if (a->cross) {
    // each form's graphs checked by the other form's code: the batched kernel
    // builds (build items only) and the plain kernels check and free; then
    // the plain kernels build and the batched kernel checks and frees (check
    // items only, every job's build count preset to complete)
    std::vector<u64> sb, sf;
    std::vector<u32> nch(njobs);
    for (u32 j = 0; j < njobs; ++j) {
      nch[j] = coffs[j + 1] - coffs[j];
      for (u32 c = 0; c < nch[j]; ++c) {
        sb.push_back(((u64)j << 32) | c);
        sf.push_back((1ull << 63) | ((u64)j << 32) | c);
      }
    }
    // the batched kernel over `items`; preset: every job's build count starts complete
    auto launch_items = [&](const std::vector<u64>& items, bool preset) -> int {
      graph_reset_k<<<16, 256, 0, s.s>>>((u32*)dnext.p, (u32*)dbuilt.p, njobs);
      if (preset) HH(hipMemcpyAsync(dbuilt.p, nch.data(), (size_t)njobs * 4, hipMemcpyHostToDevice, s.s));
      HH(hipMemcpyAsync(dseq.p, items.data(), items.size() * 8, hipMemcpyHostToDevice, s.s));
#define HA_GRAPH_ITEMS(BB)                                                                                   \
  graph_batch_k<BB><<<bgrid, 256, 0, s.s>>>((const u64*)dseq.p, (u32)items.size(), (const u32*)dcoffs.p,     \
                                            (const u64*)doffs.p, (const u32*)dn.p, a->seed, (NodeHdr**)dtab.p, \
                                            (u32*)dnext.p, (u32*)dbuilt.p, st)
      if (a->batch == 4)
        HA_GRAPH_ITEMS(4);
      else if (a->batch == 8)
        HA_GRAPH_ITEMS(8);
      else
        HA_GRAPH_ITEMS(16);
#undef HA_GRAPH_ITEMS
      HH(hipGetLastError());
      return HA_OK;
    };
    if (int rc = launch_items(sb, false)) return rc;
    for (auto& c : chunks) {
      const unsigned grid = (unsigned)((c.second - c.first + a->bs - 1) / a->bs);
      graph_plain_free_k<<<grid, a->bs, 0, s.s>>>((const u64*)doffs.p, (const u32*)dn.p, njobs, a->seed, c.first,
                                                   c.second, (NodeHdr**)dtab.p, st);
    }
    for (auto& c : chunks) {
      const unsigned grid = (unsigned)((c.second - c.first + a->bs - 1) / a->bs);
      graph_plain_build_k<<<grid, a->bs, 0, s.s>>>((const u64*)doffs.p, (const u32*)dn.p, njobs, a->seed, c.first,
                                                    c.second, (NodeHdr**)dtab.p, st);
    }
    if (int rc = launch_items(sf, true)) return rc;
    HH(hipStreamSynchronize(s.s));
    HH(hipGetLastError());
  }
  double best_plain = 1e30, best_batch = 1e30;
  for (int r = 0; r < (a->cross ? 0 : a->repeats); ++r) {
    if (a->run_plain) {
      HH(hipStreamSynchronize(s.s));
      HH(hipEventRecord(e0, s.s));
      if (run_plain()) return hfail(HA_E_HIP, "plain graph launch failed");
      HH(hipEventRecord(e1, s.s));
      HH(hipStreamSynchronize(s.s));
      float t = 0;
      HH(hipEventElapsedTime(&t, e0, e1));
      best_plain = t < best_plain ? t : best_plain;
    }
    if (a->run_batch) {
      HH(hipStreamSynchronize(s.s));
      HH(hipEventRecord(e0, s.s));
      HH(hipGraphLaunch(exec, s.s));
      HH(hipEventRecord(e1, s.s));
      HH(hipStreamSynchronize(s.s));
      float t = 0;
      HH(hipEventElapsedTime(&t, e0, e1));
      best_batch = t < best_batch ? t : best_batch;
    }
  }
  HH(hipGetLastError());
  unsigned long long h[2];
  HH(hipMemcpy(h, dstats.p, 16, hipMemcpyDeviceToHost));
  (void)hipGraphExecDestroy(exec);
  (void)hipGraphDestroy(graph);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  res->nulls = h[0];
  res->errors = h[1];
  res->plain_ms = a->run_plain ? best_plain : 0;
  res->batch_ms = a->run_batch ? best_batch : 0;
  res->plain_launches = (int32_t)(2 * chunks.size());
  return HA_OK;
}

int ha_run_serve(const ha_serve_io* io, ha_serve_result* res) {
  const ha::Params& P = ha::host_params();
  if (!P.initialized) return hfail(HA_E_STATE, "allocator not initialized");
  if (!io || !res) return hfail(HA_E_INVALID, "null argument");
  const int nreq = io->nreq;
  if (!io->arrival_ns || !io->length || !io->tag || !io->checksum || !io->produced || nreq < 1 || nreq > (1 << 26) ||
      io->slots < 256 || io->slots > (1 << 20) || io->slots % 256 || io->lat_cap < 0 || io->sample_every < 0 ||
      (io->lat_cap && !io->lat_ns))
    return hfail(HA_E_INVALID, "bad serving arguments");
  if (io->watermark > P.arena_bytes) return hfail(HA_E_INVALID, "watermark above the heap size");
  uint32_t maxlen = 0;
  for (int i = 0; i < nreq; ++i) {
    if (io->length[i] < 1 || io->length[i] > 32768) return hfail(HA_E_INVALID, "request length must be in 1..32768");
    if (i && io->arrival_ns[i] < io->arrival_ns[i - 1]) return hfail(HA_E_INVALID, "arrivals must be sorted");
    maxlen = io->length[i] > maxlen ? io->length[i] : maxlen;
  }
  if (io->arrival_ns[nreq - 1] > 60000000000ull) return hfail(HA_E_INVALID, "arrivals beyond 60 s");
  int dev = 0, khz = 0;
  HH(hipGetDevice(&dev));
  HH(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev));
  if (khz <= 0) return hfail(HA_E_HIP, "wall clock rate unavailable");
  const double tk_per_ns = khz * 1e-6;
  std::vector<u64> arr_tk(nreq);
  for (int i = 0; i < nreq; ++i) arr_tk[i] = (u64)((double)io->arrival_ns[i] * tk_per_ns);
  const u32 max_pages = (maxlen + SRV_TOK_PER_PAGE - 1) / SRV_TOK_PER_PAGE;
  const size_t n = (size_t)nreq;
  DBuf darr, dlen, dtag, dcs, dprod, dfirst, ddone, dlat, dtab, dst, dbad, dbadf;
  HH(dbad.alloc(n * 4));
  HH(dbadf.alloc(n * 8));
  HH(hipMemset(dbad.p, 0, n * 4));
  HH(hipMemset(dbadf.p, 0, n * 8));
  HH(darr.alloc(n * 8));
  HH(dlen.alloc(n * 4));
  HH(dtag.alloc(n * 4));
  HH(dcs.alloc(n * 4));
  HH(dprod.alloc(n * 4));
  HH(dfirst.alloc(n * 8));
  HH(ddone.alloc(n * 8));
  HH(dlat.alloc((size_t)(io->lat_cap ? io->lat_cap : 1) * 4));
  HH(dtab.alloc((size_t)io->slots * max_pages * 8));
  HH(dst.alloc(sizeof(SrvState)));
  HH(hipMemcpy(darr.p, arr_tk.data(), n * 8, hipMemcpyHostToDevice));
  HH(hipMemcpy(dlen.p, io->length, n * 4, hipMemcpyHostToDevice));
  HH(hipMemcpy(dtag.p, io->tag, n * 4, hipMemcpyHostToDevice));
  HH(hipMemset(dcs.p, 0, n * 4));
  HH(hipMemset(dprod.p, 0, n * 4));
  HH(hipMemset(dfirst.p, 0, n * 8));
  HH(hipMemset(ddone.p, 0, n * 8));
  SrvState init;
  std::memset(&init, 0, sizeof init);
  init.free_est = P.arena_bytes;
  HH(hipMemcpy(dst.p, &init, sizeof init, hipMemcpyHostToDevice));
  SrvArgs a;
  a.arrival = (const u64*)darr.p;
  a.length = (const u32*)dlen.p;
  a.tag = (const u32*)dtag.p;
  a.checksum = (u32*)dcs.p;
  a.produced = (u32*)dprod.p;
  a.first_tk = (u64*)dfirst.p;
  a.done_tk = (u64*)ddone.p;
  a.lat_tk = (u32*)dlat.p;
  a.bad = (u32*)dbad.p;
  a.bad_first = (u64*)dbadf.p;
  a.page_tab = (void**)dtab.p;
  a.st = (SrvState*)dst.p;
  a.watermark = io->watermark;
  a.deadline_tk = arr_tk[nreq - 1] + (u64)(30e9 * tk_per_ns);
  a.nreq = (u32)nreq;
  a.max_pages = max_pages;
  a.sample_every = (u32)io->sample_every;
  a.lat_cap = (u32)io->lat_cap;
  Stream s;
  HH(hipStreamCreate(&s.s));
  hipEvent_t e0, e1;
  HH(hipEventCreate(&e0));
  HH(hipEventCreate(&e1));
  HH(hipEventRecord(e0, s.s));
  srv_start_k<<<1, 1, 0, s.s>>>(a.st);
  srv_loop_k<<<io->slots / 256, 256, 0, s.s>>>(a);
  HH(hipEventRecord(e1, s.s));
  HH(hipStreamSynchronize(s.s));
  HH(hipGetLastError());
  float t = 0;
  HH(hipEventElapsedTime(&t, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  SrvState fin;
  HH(hipMemcpy(&fin, dst.p, sizeof fin, hipMemcpyDeviceToHost));
  HH(hipMemcpy(io->checksum, dcs.p, n * 4, hipMemcpyDeviceToHost));
  HH(hipMemcpy(io->produced, dprod.p, n * 4, hipMemcpyDeviceToHost));
  const double ns_per_tk = 1.0 / tk_per_ns;
  std::vector<u64> tk(n);
  if (io->first_ns) {
    HH(hipMemcpy(tk.data(), dfirst.p, n * 8, hipMemcpyDeviceToHost));
    for (size_t i = 0; i < n; ++i) io->first_ns[i] = (uint64_t)((double)tk[i] * ns_per_tk);
  }
  if (io->done_ns) {
    HH(hipMemcpy(tk.data(), ddone.p, n * 8, hipMemcpyDeviceToHost));
    for (size_t i = 0; i < n; ++i) io->done_ns[i] = (uint64_t)((double)tk[i] * ns_per_tk);
  }
  if (io->bad) HH(hipMemcpy(io->bad, dbad.p, n * 4, hipMemcpyDeviceToHost));
  if (io->bad_first) HH(hipMemcpy(io->bad_first, dbadf.p, n * 8, hipMemcpyDeviceToHost));
  const uint64_t ns = fin.samples < (unsigned long long)io->lat_cap ? fin.samples : (uint64_t)io->lat_cap;
  if (ns) {
    std::vector<u32> l(ns);
    HH(hipMemcpy(l.data(), dlat.p, ns * 4, hipMemcpyDeviceToHost));
    for (uint64_t i = 0; i < ns; ++i) io->lat_ns[i] = (uint32_t)((double)l[i] * ns_per_tk);
  }
  res->kernel_ms = t;
  res->pages = fin.pages;
  res->nulls = fin.nulls;
  res->stalls = fin.stalls;
  res->samples = ns;
  res->completed = fin.done;
  res->admitted = fin.admitted;
  res->clock_khz = (uint32_t)khz;
  res->pad = 0;
  return HA_OK;
}

}  // extern "C"
