// ha_internal.h -- device-side state and helpers of the MI355X halloc allocator.
//
// Layout in HBM (all allocated by ha_init):
//   arena       nsbs * sb_sz bytes, one allocation; slab b = [b*sb_sz, (b+1)*sb_sz)
//   bits        nsbs * (sb_sz/1024) u64: one bit per 16 B chunk      (src/halloc.cu:345)
//   nibs        nsbs * (sb_sz/256)  u64: 4-bit block length per chunk (src/slab.cuh:60-83)
//   ctr         nsbs * CTR_STRIDE   u64: slab counter, one per 128 B line (src/slab.h:52-132):
//               count bits 32..63, head references 16..23, chunk id 8..11
//   desc        nsbs Desc: size id, chunk id, chunk size, head references (src/slab.h:16-29)
//   sb_lock     nsbs u32                                                (src/slab.cuh:18)
//   sets        NSETS * nsw u64 slab bitmaps: free, roomy[16], busy[16], sparse[4]
//   head/cached/head_lock/ticket: nheads * 16 u32 each, one per 128 B line
//   lbits       ceil(lunits/64) u64: one bit per 4 KiB unit of the large region
//   lstart      lunits u32: class + 1 of the live large block starting at the unit
//   lticket     LARGE_NCLASSES u32, one per 128 B line
// Read-only parameters (pointers, sizes, the size table) sit in __constant__ memory.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

#include "ha_bits.h"

namespace ha {

using u32 = uint32_t;
using u64 = uint64_t;

constexpr u32 NSIZES = 16;
constexpr u32 NCHUNK_IDS = 4;
constexpr u32 SZ_NONE = 0xffffffffu;
constexpr u32 SB_NONE = 0xffffffffu;
constexpr u32 MIN_BLOCK_SZ = 16;
constexpr u32 MAX_BLOCK_SZ = 3072;
constexpr u32 BLOCK_STEP = 16;
constexpr u32 MAX_NTRIES = 32;       // src/slab.h:36
constexpr u32 CHECK_NTRIES = 2;      // src/slab.h:39
constexpr u32 CROWDED_NTRIES = 8;   // failed probe rounds of one sequence that mark a head crowded
constexpr u32 STEP_FREQ = 64;        // src/slab.h:48
constexpr u32 THREAD_FREQ = 11;      // src/halloc.cu:40
constexpr u32 RUN = 64;              // blocks per class-group run (one per wave lane)
constexpr u32 MAXB = 16;             // most blocks one lane takes in one batched call
constexpr u32 TICKET_INIT = 0x0B0B0B0Bu;  // src/halloc.cu:416
constexpr u32 NSETS = 1 + 2 * NSIZES + NCHUNK_IDS;
constexpr u32 HOT_STRIDE = 32;       // u32 words: one 128 B line per hot word
constexpr u32 CTR_STRIDE = 16;       // u64 words: one 128 B line per slab counter
constexpr u32 MAX_NHEADS = 64;
constexpr u32 MAX_NSBS = 1u << 20;
constexpr u32 MIN_SB_SZ_SH = 21;     // 2 MiB: the smallest slab whose 64-block-run size table exists
constexpr u32 MAX_SB_SZ_SH = 26;     // tst/common/common.cu:169 allows -b 20..26

// a head word is a slab id, plus HEAD_EXT when the slab came from the busy set:
// such a head is filled up to (busy + nchunks) / 2 instead of busy
constexpr u32 HEAD_EXT = 0x80000000u;
// large blocks (above MAX_BLOCK_SZ): class c is 2^c units of LARGE_UNIT bytes,
// aligned to its size inside the large region; one bit per unit marks it used
constexpr u32 LARGE_UNIT_SH = 12;     // 4 KiB
constexpr u32 LARGE_NCLASSES = 40;    // classes 4 KiB << c
constexpr u32 LARGE_SB = 0xfffffffeu; // the slab field of a replay record of a large block

constexpr u32 CTR_CHUNK_SH = 8;
constexpr u64 CTR_CHUNK_NONE = 0xF;
constexpr u64 CTR_HEAD = 1ull << 16;  // one head reference (bits 16..23 count them)
constexpr u32 CTR_COUNT_SH = 32;

struct SizeInfo {
  u32 nchunks, chunk_sz, chunk_id, ncib, sparse_thr, hash_step, roomy_thr, busy_thr;
  u32 block_sz, nblocks, tmod, chunk_bytes;
  u32 nchunks_m, hash_step_m;  // mod_magic of nchunks and hash_step
  u32 pad[2];
};

struct Desc {
  u32 size_id, chunk_id, chunk_sz, is_head;
};

struct Params {
  char* arena;
  u64 arena_bytes;
  u64* bits;
  u64* nibs;
  u64* ctr;
  Desc* desc;
  u32* sb_lock;
  u64* sets;
  u32* head;
  u32* cached;
  u32* head_lock;
  u32* ticket;
  u32* class_lock;  // one per size class, serializes head searches of its pairs
  u32* search_ep;   // per pair: searches done (bits 31..1) and whether the last failed (bit 0)
  u32* forced;      // per pair: the head whose forced replacement found no slab, and the free epoch then
  u32* active;      // pairs that hold a head
  u32* free_ep;     // slabs that became free so far
  u32* spread;      // head sets in use now (adaptive sets): waves fold their set below it
  u32* class_seen;  // size classes that have held a head, one bit each
  // large-block region: one bit per 4 KiB unit (lbits, the bits past lunits
  // set), per unit the class + 1 of the live block that starts there (lstart,
  // 0 elsewhere), per class a ticket that spreads the first word probed
  char* large;
  u64 large_bytes;
  u64* lbits;
  u32* lstart;
  u32* lticket;
  u32 lunits, lwords, lnclasses, lpad;
  u32 nsbs, sb_sz, sb_sz_sh, nheads;
  u32 nsw, bit_words, nib_words, initialized;
  u32 cached_arena;       // hafree releases before publishing
  u32 nh_m, nh8_m;        // mod_magic of nheads and of nheads / 8
  u32 fixed_sets;         // 1: every wave keeps its own head set (no spread)
  u64 cuda_memory;
  SizeInfo si[NSIZES];
};

extern __constant__ Params ha_p;

// --------------------------------------------------------------- pure helpers
__host__ __device__ inline u32 size_id_from_nbytes(u32 nbytes) {
  // blocks are 2*unit or 3*unit, unit = 8..1024 (src/halloc.cu:53-67)
  nbytes = nbytes < MIN_BLOCK_SZ ? MIN_BLOCK_SZ : nbytes;
  u32 iunit = (32 - __builtin_clz(nbytes)) - 5, unit = 1u << (iunit + 3);
  if (nbytes > 3 * unit) { iunit += 1; unit *= 2; }
  return nbytes <= 2 * unit ? 2 * iunit : 2 * iunit + 1;
}
// chunk size is stored as (div3 << 16 | shift) (src/size-info.h:50-70)
__host__ __device__ inline u32 chunk_mul(u32 v, u32 cs) { return (v << (cs & 0xffffu)) * (cs >> 16); }
__host__ __device__ inline u32 chunk_div(u32 v, u32 cs) {
  if ((cs >> 16) == 3u) v /= 3u;
  return v >> (cs & 0xffffu);
}
__host__ __device__ inline u32 ctr_count(u64 c) { return (u32)(c >> CTR_COUNT_SH); }
__host__ __device__ inline u32 ctr_chunk(u64 c) { return (u32)((c >> CTR_CHUNK_SH) & 0xF); }
__host__ __device__ inline bool ctr_is_head(u64 c) { return ((c >> 16) & 0xFF) != 0; }
__host__ __device__ inline u64 ctr_chunk_delta(u32 chunk_id) {
  return ((CTR_CHUNK_NONE - chunk_id) & 0xF) << CTR_CHUNK_SH;
}

__host__ __device__ inline u32 head_limit(u32 head_word, const SizeInfo& s) {
  return (head_word & HEAD_EXT) ? (s.busy_thr + s.nchunks) / 2 : s.busy_thr;
}

// slab class by count: free, sparse, roomy, busy
enum : u32 { CLS_FREE = 0, CLS_SPARSE = 1, CLS_ROOMY = 2, CLS_BUSY = 3 };
__host__ __device__ inline u32 sb_class(u32 count, const SizeInfo& s) {
  if (count == 0) return CLS_FREE;
  if (count <= s.sparse_thr) return CLS_SPARSE;
  if (count <= s.roomy_thr) return CLS_ROOMY;
  return CLS_BUSY;
}
__host__ __device__ inline u32 set_roomy(u32 s) { return 1 + s; }
__host__ __device__ inline u32 set_busy(u32 s) { return 1 + NSIZES + s; }
__host__ __device__ inline u32 set_sparse(u32 c) { return 1 + 2 * NSIZES + c; }
__host__ __device__ inline u32 class_set(u32 cls, u32 size_id, u32 chunk_id) {
  return cls == CLS_SPARSE ? set_sparse(chunk_id) : cls == CLS_ROOMY ? set_roomy(size_id) : set_busy(size_id);
}

}  // namespace ha
