// halloc_model.hpp -- host C++ model of the halloc slab / bit-array / hash scheme.
//
// TEST INFRASTRUCTURE ONLY.  This file is the correctness oracle for the
// MI355X device allocator in halloc_amd/csrc.  Only tests/, the smoke check and
// the CPU-baseline leg of bench.py may load it; the product library never links it.
//
// It restates the reference allocator's algorithm (canonizer/halloc, src/halloc.cu,
// src/slab.cuh, src/slab.h, src/sbset.*, src/size-info.h, src/utils.cu) on the host:
//   * size classes, chunk arithmetic and the size table   (src/halloc.cu:53-67, 363-394;
//     src/size-info.h:50-70; src/utils.cu:8-24)
//   * per-class tickets and the initial hash              (src/halloc.cu:71-122)
//   * the hashed probe over a slab's bit array           (src/slab.cuh:453-535)
//   * slab counters with chunk id and head flag           (src/slab.h:52-132)
//   * head replacement, cached head, slab search order    (src/slab.cuh:227-445)
//   * busy / roomy / sparse / free slab sets              (src/sbset.h, src/sbset.cuh)
//   * free path and count transitions                     (src/halloc.cu:211-256, src/slab.cuh:161-224)
//
// The model emulates one wavefront in lock step with explicit lane loops, so
//   - one active lane reproduces the reference's per-lane semantics, and
//   - one 64-lane wave pins the wave-aggregation rules the GPU follows:
//     lanes are processed in size-class groups ordered by their lowest lane,
//     and within a group every step (ticket, probe, counter update, head
//     replacement) is group-uniform.
// Every shared word is a std::atomic, so several host threads may drive waves
// concurrently: that is the multithreaded CPU baseline.
//
// Deliberate differences from the reference, shared with the GPU:
//   * 64-bit bit, nibble and set words (the reference uses 32-bit words).
//   * 64-bit slab counter: count in bits 32..63, head references bits 16..23,
//     chunk id bits 8..11 (0xF = none).  A slab may be the head of several
//     (head, class) pairs of ONE class at once: when no set holds a slab for a
//     new head, the head of another XCD's pair of the same class is shared
//     (its reference count goes up); a slab is detached when its last head
//     reference goes.
//   * one contiguous arena: slab = (offset >> sb_sz_sh); no address grid.
//   * hash_mode 1 ("wave run"): the reference's initial hash with tmod = 64
//     (nblocks rounded down to a multiple of 64), and lane l of a class group
//     takes ticket cv = old + l where the group adds (highest lane + 1) to the
//     ticket counter.  A wave's blocks then fill at most two aligned 64-block
//     runs, contiguously by lane.  hash_mode 0 keeps the reference's tmod
//     (4/2/1) and tickets cv = old + rank; it exists in the model only.
//   * the probe counter itry is group-uniform, and the probe offset
//     (itry+1)*step is added in 64-bit arithmetic before the modulo.
//   * slab set membership is complete: every non-head, non-free slab is in a
//     set.  Count transitions and head detaches file a slab by its count
//     (sparse <= sparse thr < roomy <= roomy thr < busy); the reference files
//     detached heads above the roomy threshold nowhere (src/slab.cuh:236).
//   * the busy-set search keeps slabs it rejects (the reference drops them),
//     and a head it returns is filled up to the limit it was accepted at,
//     (busy + nchunks) / 2, instead of being swapped out at busy again.
//   * a free of a pointer in a free slab, or whose nibble is 0, is a no-op.
//   * when no set holds a slab for a new head, an EMPTY head or cached head of
//     another (head, class) pair is taken over (under a try-lock of its head
//     lock), so slabs parked as idle heads do not make hamalloc return NULL.
//   * with several head sets, free slabs are "scarce" below the largest of
//     nheads, half the (head set, class) pairs holding a head, and a quarter of
//     the heap: then a cached head is not refilled from the free set, and a new
//     head first frees one slab another class holds idle (its cached head, or
//     its empty head on a pair outside the spread) and, when there is none,
//     shares its class's head before taking a free slab.
//   * one head search per class at a time (a class lock); a search frees at
//     most 4 idle heads of other pairs; waiters on a head lock whose holder's
//     search failed return NULL instead of searching again.
//   * hash_step_frac: the probe stride's prime may be bounded by a fraction of
//     nblocks instead of the reference's nblocks/256 + nblocks/64.
//   * batched calls (hamalloc_batch): one ticket for all rounds, the first
//     probes of 4 rounds made before any is checked, one counter update.
//
//   * nblocks is rounded down to a multiple of 64 and then by one more run of
//     64 when THREAD_FREQ (11) would divide the number of runs (first probes
//     step 11 runs at a time and would reach only every 11th run).
//   * a head in which a group's probes ran out of tries (31) is replaced even
//     below its busy limit ("forced"); when no other slab exists it is kept,
//     and the pair records the head and the count of slabs freed so far: a
//     later forced replacement of that head keeps it without a search until
//     a slab becomes free.
//     The reference replaces a head only when it is missing or at its busy
//     threshold (src/slab.cuh:385-387).
//   * a head in which one probe sequence of a group failed CROWDED_NTRIES (8)
//     rounds is "crowded" and replaced the same way, even when the group was
//     served: a group waits for its last lane, about log(lanes) / log(1/fill)
//     rounds, so a wave pays far more than one thread for a head near its
//     limit or full of scattered holes.
//   * requests above 3072 B: the reference hands them to the CUDA device heap
//     (src/halloc.cu:185-187, 233-235), which is not in its tree.  Here they
//     come from the large region, the share of the pool outside the slabs
//     (src/halloc.cu:285-289): class c = ceil(log2(units of 4 KiB)), a block of
//     2^c units aligned to its size, one bit per unit, per unit the class + 1
//     of the live block starting there.  A class group of a wave takes one
//     ticket; blocks of up to 64 units are taken from the free fields of the
//     words in order from large_start(ticket), the lane of rank r taking the
//     r-th free field, staying on a word while it shows a free field; larger blocks
//     take aligned runs of zero words lane by lane.  A free exchanges the start
//     word with 0 and clears the block's bits only when it read a class, so a
//     double free, an interior pointer or a unit never handed out is a no-op.
//     The model states the field search with loops over aligned positions, the
//     device with word-parallel bit arithmetic (ha_bits.h).
//   * adaptive head sets (fixed_sets 0): waves fold their set onto the first
//     `spread` sets; nheads is cut to at most the largest power of two s
//     with 3 * s <= slabs (fixed sets too), the spread starts at nheads and shrinks when a class
//     takes its first head, to the largest power of two s with
//     3 * s * classes seen <= slabs.  A pair that needs a head adopts a head of its class from a
//     pair outside the spread (after the roomy and sparse sets, before a free
//     slab); a pair outside the spread is served from the pair it folds onto.
#pragma once

#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <thread>
#include <vector>

namespace hamodel {

using u32 = uint32_t;
using u64 = uint64_t;

constexpr u32 NSIZES = 16;          // 8 units x {2,3}            (src/size-info.h:58)
constexpr u32 NCHUNK_IDS = 4;       // chunk sizes 16, 24, 256, 384
constexpr u32 SZ_NONE = 0xffffffffu;
constexpr u32 SB_NONE = 0xffffffffu;
constexpr u32 MIN_BLOCK_SZ = 16;    // src/size-info.h:60
constexpr u32 MAX_BLOCK_SZ = 3072;  // src/size-info.h:62
constexpr u32 BLOCK_STEP = 16;      // bit-array granularity, src/size-info.h:50
constexpr u32 MAX_NTRIES = 32;      // src/slab.h:36
constexpr u32 CHECK_NTRIES = 2;     // src/slab.h:39
constexpr u32 CROWDED_NTRIES = 8;  // failed probe rounds of one sequence that mark a head crowded
constexpr u32 STEP_FREQ = 64;       // src/slab.h:48
constexpr u32 THREAD_FREQ = 11;     // src/halloc.cu:40
constexpr u32 THREAD_MOD_SH = 2;    // src/halloc.cu:44
constexpr u32 TICKET_INIT = 0x0B0B0B0Bu;  // cudaMemset(counters_g, 11), src/halloc.cu:416
constexpr u32 RUN = 64;             // wave-run length (hash_mode 1)
constexpr u32 WAVE = 64;
constexpr u32 MAXB = 16;            // most blocks one lane takes in one batched call
// a head word is a slab id, plus HEAD_EXT when the slab came from the busy set:
// such a head is filled up to (busy + nchunks) / 2 instead of busy
constexpr u32 HEAD_EXT = 0x80000000u;
constexpr u32 MAX_STEALS = 4;       // idle heads one head search may free before it gives up
constexpr u32 BATCH_PIPE = MAXB;    // rounds of a batched call whose first probes go out together (all)
// large region
constexpr u32 LARGE_UNIT_SH = 12;   // 4 KiB units
constexpr u32 LARGE_NCLASSES = 40;
constexpr u32 LARGE_SB = 0xfffffffeu;  // slab field of a large block's record
constexpr u64 UNIT_NONE = ~0ull;

// slab counter layout (64-bit)
constexpr u32 CTR_CHUNK_SH = 8;
constexpr u64 CTR_CHUNK_NONE = 0xF;
constexpr u64 CTR_HEAD = 1ull << 16;  // one head reference
constexpr u32 CTR_COUNT_SH = 32;
inline u32 ctr_count(u64 c) { return (u32)(c >> CTR_COUNT_SH); }
inline u32 ctr_chunk(u64 c) { return (u32)((c >> CTR_CHUNK_SH) & 0xF); }
inline bool ctr_is_head(u64 c) { return ((c >> 16) & 0xFF) != 0; }
inline u32 ctr_head_refs(u64 c) { return (u32)((c >> 16) & 0xFF); }
inline u64 ctr_chunk_delta(u32 chunk_id) { return ((CTR_CHUNK_NONE - chunk_id) & 0xF) << CTR_CHUNK_SH; }

// ---------------------------------------------------------------- closed forms
// size class from request size: blocks are 2*unit or 3*unit, unit = 8..1024
// (src/halloc.cu:53-67)
inline u32 size_id_from_nbytes(u32 nbytes) {
  if (nbytes < MIN_BLOCK_SZ) nbytes = MIN_BLOCK_SZ;
  u32 iunit = (32 - __builtin_clz(nbytes)) - 5, unit = 1u << (iunit + 3);
  if (nbytes > 3 * unit) { iunit += 1; unit *= 2; }
  return nbytes <= 2 * unit ? 2 * iunit : 2 * iunit + 1;
}
inline u32 block_sz_of(u32 sid) { u32 unit = 1u << (sid / 2 + 3); return (sid % 2 ? 3 : 2) * unit; }
// reference thread modulus by size (src/halloc.cu:101-104)
inline u32 tmod_ref(u32 sid) { int sh = (int)THREAD_MOD_SH - (int)sid / 2; return 1u << (sh > 0 ? sh : 0); }
// chunk size encoding (div3 << 16 | shift) (src/size-info.h:50-70)
inline u32 chunk_val(u32 chunk_sz) {
  u32 div3 = chunk_sz % 3 ? 1 : 3;
  if (chunk_sz % 3 == 0) chunk_sz /= 3;
  u32 sh = 0;
  for (; (1u << sh) < chunk_sz; sh++) {}
  return div3 << 16 | sh;
}
inline u32 chunk_mul(u32 v, u32 cs) { return (v << (cs & 0xffffu)) * (cs >> 16); }
inline u32 chunk_div(u32 v, u32 cs) { if ((cs >> 16) == 3u) v /= 3u; return v >> (cs & 0xffffu); }
// largest prime p in [3, n) with n % p != 0 and nb % p != 0; 0 if none
// (src/utils.cu:8-24 exits the process instead of returning 0)
inline u32 max_prime_below(u32 n, u32 nb) {
  for (u32 p = n - 1; p >= 3 && p < n; p--) {
    u32 max_d = (u32)std::floor(std::sqrt((double)p));
    bool is_prime = true;
    for (u32 d = 2; d <= max_d; d++) if (p % d == 0) { is_prime = false; break; }
    if (is_prime && n % p && nb % p) return p;
  }
  return 0;
}

struct Opts {
  u64 memory = 512ull << 20;     // src/halloc.h:30-32 defaults
  double halloc_fraction = 0.75;
  double busy_fraction = 0.835;
  double roomy_fraction = 0.6;
  double sparse_fraction = 0.012;
  int sb_sz_sh = 22;
  int nheads = 1;
  int hash_mode = 1;
  double hash_step_frac = 0.0;   // 0: the reference's prime bound
  int fixed_sets = 0;            // 0: adaptive head sets (the spread shrinks as classes appear)
  u64 device_memory = 0;         // 0: no 3/4-of-device cap
};

struct SizeInfo {
  u32 nchunks, chunk_sz, chunk_id, ncib, sparse_thr, hash_step, roomy_thr, busy_thr;
  u32 block_sz, nblocks, tmod, chunk_bytes;
};

// the size table (src/halloc.cu:363-394); hash_mode 1 rounds nblocks to a multiple of RUN
// whose run count THREAD_FREQ does not divide
inline bool build_size_infos(u32 sb_sz, const Opts& o, SizeInfo out[NSIZES]) {
  for (u32 s = 0; s < NSIZES; ++s) {
    SizeInfo& si = out[s];
    std::memset(&si, 0, sizeof si);
    si.block_sz = block_sz_of(s);
    u32 nblocks = sb_sz / si.block_sz;
    si.tmod = o.hash_mode == 0 ? tmod_ref(s) : RUN;
    nblocks = nblocks / si.tmod * si.tmod;
    // runs advance THREAD_FREQ at a time: one run fewer when that would
    // divide their number, so that first probes reach every run
    if (o.hash_mode != 0 && (nblocks / RUN) % THREAD_FREQ == 0) nblocks -= RUN;
    si.nblocks = nblocks;
    si.chunk_id = s % 2 + (s < NSIZES / 2 ? 0 : 2);
    si.chunk_bytes = (si.chunk_id % 2 ? 3 : 2) * (si.chunk_id / 2 ? 128 : 8);
    si.chunk_sz = chunk_val(si.chunk_bytes);
    si.ncib = si.block_sz / si.chunk_bytes;
    si.nchunks = nblocks * si.ncib;
    const u32 bound = o.hash_step_frac > 0 ? (u32)(o.hash_step_frac * nblocks) : nblocks / 256 + nblocks / 64;
    u32 p = max_prime_below(bound, nblocks);
    if (p == 0 || nblocks == 0) return false;
    si.hash_step = si.ncib * p;
    si.roomy_thr = (u32)(o.roomy_fraction * si.nchunks);
    si.busy_thr = (u32)(o.busy_fraction * si.nchunks);
    si.sparse_thr = (u32)(o.sparse_fraction * si.nchunks);
  }
  return true;
}

// initial chunk of a lane (src/halloc.cu:107-122).  hash_mode 0: cv is the
// lane's own ticket.  hash_mode 1: cv = G*RUN + rank (the group's run ticket G).
inline u32 ichunk_init(u32 cv, const SizeInfo& si) {
  u32 t = si.tmod, tm = t - 1;
  u32 ich = (cv & ~tm) * THREAD_FREQ + (cv & tm);
  return ich * si.ncib % si.nchunks;
}
// next probe position (src/slab.cuh:492-495)
inline u32 probe_next(u32 ich0, u32 itry, const SizeInfo& si) {
  u32 step = ((itry + 1) * STEP_FREQ * si.ncib) % si.hash_step;
  return (u32)(((u64)ich0 + (u64)(itry + 1) * step) % si.nchunks);
}

// slab classes by count (complete membership)
enum SbClass : u32 { CLS_FREE = 0, CLS_SPARSE = 1, CLS_ROOMY = 2, CLS_BUSY = 3 };
inline u32 head_limit(u32 head_word, const SizeInfo& si) {
  return (head_word & HEAD_EXT) ? (si.busy_thr + si.nchunks) / 2 : si.busy_thr;
}
inline u32 sb_class(u32 count, const SizeInfo& si) {
  if (count == 0) return CLS_FREE;
  if (count <= si.sparse_thr) return CLS_SPARSE;
  if (count <= si.roomy_thr) return CLS_ROOMY;
  return CLS_BUSY;
}

// one record per replayed operation
// malloc: (size class, slab, chunk), a large block (16 + class, LARGE_SB, unit),
// NULL (REC_NULL x 3); free: (chunks or units freed, slab or LARGE_SB, chunk or unit)
struct Rec { u32 a, sb, chunk, extra; };
constexpr u32 REC_NULL = 0xffffffffu;

// class of a large request: blocks of 2^c units of 4 KiB
inline u32 large_class(u64 nbytes) {
  u64 units = (nbytes >> LARGE_UNIT_SH) + ((nbytes & ((1ull << LARGE_UNIT_SH) - 1)) != 0);
  u32 c = 0;
  while ((1ull << c) < units) ++c;
  return c;
}
// first word (or run of words) a group of class c with this ticket looks at
inline u32 large_start(u32 ticket, u32 c, u32 n) {
  u32 h = (ticket * 0x9E3779B1u) ^ (c * 0x85EBCA77u);
  return (u32)(((u64)h * n) >> 32);
}

// -------------------------------------------------------------------- the heap
class Heap {
 public:
  Opts opts;
  u32 sb_sz = 0, sb_sz_sh = 0, nsbs = 0, nheads = 1, nsw = 0;  // nsw: words per slab set
  u64 arena_bytes = 0;
  u32 bit_words = 0, nib_words = 0;  // per slab
  SizeInfo si[NSIZES];
  bool ok = false;

  std::vector<std::atomic<u64>> bits, nibs, ctr, sets;
  std::vector<std::atomic<u32>> d_size, d_chunk, d_csz, d_head, sb_lock;
  std::vector<std::atomic<u32>> head, cached, head_lock, ticket, search_ep;
  // per pair: the head whose forced replacement found no other slab, and the
  // free-slab epoch at that time; free_ep counts slabs that became free
  std::vector<std::atomic<u32>> forced_head, forced_ep;
  std::atomic<u32> free_ep{0};
  std::atomic<u64> n_crowded{0};  // probe sequences that marked their head crowded
  // adaptive head sets: waves fold their set onto the first `spread` sets;
  // class_seen has a bit per size class that has held a head
  std::atomic<u32> spread{0}, class_seen{0};
  std::atomic<u32> class_lock[NSIZES] = {};
  std::atomic<u32> active{0};  // pairs that hold a head
  std::vector<uint8_t> arena;  // payload memory (allocated lazily by users that touch data)
  // large region: lbits one bit per unit (the bits past lunits set), lstart
  // per unit the class + 1 of the live block starting there, a ticket per class
  u32 lunits = 0, lwords = 0, lnclasses = 0;
  u64 large_bytes = 0;
  std::vector<std::atomic<u64>> lbits;
  std::vector<std::atomic<u32>> lstart;
  std::atomic<u32> lticket[LARGE_NCLASSES] = {};

  // set index helpers: 0 = free, 1+s = roomy[s], 17+s = busy[s], 33+c = sparse[c]
  static constexpr u32 NSETS = 1 + 2 * NSIZES + NCHUNK_IDS;
  u32 set_roomy(u32 s) const { return 1 + s; }
  u32 set_busy(u32 s) const { return 1 + NSIZES + s; }
  u32 set_sparse(u32 c) const { return 1 + 2 * NSIZES + c; }

  explicit Heap(const Opts& o, bool with_arena = false) : opts(o) {
    u64 memory = o.memory;
    if (o.device_memory) memory = std::min<u64>(memory, 3ull * o.device_memory / 4ull);
    sb_sz_sh = (u32)o.sb_sz_sh;
    sb_sz = 1u << sb_sz_sh;
    u64 halloc_memory = (u64)(o.halloc_fraction * (double)memory);
    nsbs = (u32)(halloc_memory / sb_sz);
    nheads = (u32)(o.nheads < 1 ? 1 : o.nheads);
    // at most the sets one size class can hold heads in (halloc.hip
    // ha_init_c, the same cut)
    if (nsbs / 3 < nheads) {
      u32 s = 1;
      while (2 * s <= nsbs / 3) s *= 2;
      nheads = s;
    }
    if (!build_size_infos(sb_sz, o, si) || nsbs == 0) return;
    nsw = (nsbs + 63) / 64;
    bit_words = sb_sz / BLOCK_STEP / 64;
    nib_words = sb_sz / BLOCK_STEP / 16;
    arena_bytes = (u64)nsbs * sb_sz;
    bits = std::vector<std::atomic<u64>>((size_t)nsbs * bit_words);
    nibs = std::vector<std::atomic<u64>>((size_t)nsbs * nib_words);
    ctr = std::vector<std::atomic<u64>>(nsbs);
    sets = std::vector<std::atomic<u64>>((size_t)NSETS * nsw);
    d_size = std::vector<std::atomic<u32>>(nsbs);
    d_chunk = std::vector<std::atomic<u32>>(nsbs);
    d_csz = std::vector<std::atomic<u32>>(nsbs);
    d_head = std::vector<std::atomic<u32>>(nsbs);
    sb_lock = std::vector<std::atomic<u32>>(nsbs);
    head = std::vector<std::atomic<u32>>((size_t)nheads * NSIZES);
    cached = std::vector<std::atomic<u32>>((size_t)nheads * NSIZES);
    head_lock = std::vector<std::atomic<u32>>((size_t)nheads * NSIZES);
    search_ep = std::vector<std::atomic<u32>>((size_t)nheads * NSIZES);
    ticket = std::vector<std::atomic<u32>>((size_t)nheads * NSIZES);
    forced_head = std::vector<std::atomic<u32>>((size_t)nheads * NSIZES);
    forced_ep = std::vector<std::atomic<u32>>((size_t)nheads * NSIZES);
    for (auto& x : bits) x.store(0, std::memory_order_relaxed);
    for (auto& x : nibs) x.store(0, std::memory_order_relaxed);
    for (auto& x : sets) x.store(0, std::memory_order_relaxed);
    for (u32 b = 0; b < nsbs; ++b) {
      ctr[b].store(CTR_CHUNK_NONE << CTR_CHUNK_SH, std::memory_order_relaxed);
      d_size[b].store(SZ_NONE); d_chunk[b].store(SZ_NONE); d_csz[b].store(0); d_head[b].store(0);
      sb_lock[b].store(0);
      sets[(size_t)0 * nsw + b / 64].fetch_or(1ull << (b % 64));
    }
    for (size_t i = 0; i < head.size(); ++i) {
      head[i].store(SB_NONE); cached[i].store(SB_NONE); head_lock[i].store(0); ticket[i].store(TICKET_INIT);
      search_ep[i].store(0);
      forced_head[i].store(SB_NONE); forced_ep[i].store(0);
    }
    // the rest of the pool is the large region (src/halloc.cu:285-289)
    const u64 large_memory = memory - halloc_memory;
    lunits = (u32)std::min<u64>(large_memory >> LARGE_UNIT_SH, 0xfffffff0ull);
    large_bytes = (u64)lunits << LARGE_UNIT_SH;
    lwords = (lunits + 63) / 64;
    while (lnclasses < LARGE_NCLASSES && (1ull << lnclasses) <= lunits) ++lnclasses;
    lbits = std::vector<std::atomic<u64>>(lwords);
    lstart = std::vector<std::atomic<u32>>(lunits);
    for (u32 w = 0; w < lwords; ++w) {
      u64 v = 0;
      for (u32 b = 0; b < 64; ++b) if ((u64)w * 64 + b >= lunits) v |= 1ull << b;
      lbits[w].store(v);
    }
    for (auto& x : lstart) x.store(0);
    spread.store(nheads);
    if (with_arena) arena.assign((size_t)arena_bytes, 0);
    ok = true;
  }

  // ------------------------------------------------------------- sets (sbset)
  std::atomic<u64>* set_words(u32 set) { return &sets[(size_t)set * nsw]; }
  void set_add(u32 set, u32 sb) { set_words(set)[sb / 64].fetch_or(1ull << (sb % 64)); }
  void set_remove(u32 set, u32 sb) { set_words(set)[sb / 64].fetch_and(~(1ull << (sb % 64))); }
  // first-fit claim (src/sbset.cuh:5-35), scanning every word that holds a slab
  u32 set_pop(u32 set) {
    std::atomic<u64>* w = set_words(set);
    for (u32 i = 0; i < nsw; ++i) {
      u64 word = w[i].load();
      while (word) {
        u32 b = (u32)__builtin_ctzll(word);
        u64 m = 1ull << b;
        if (w[i].fetch_and(~m) & m) return i * 64 + b;
        word &= ~m;
      }
    }
    return SB_NONE;
  }
  u32 class_set(u32 cls, u32 size_id, u32 chunk_id) const {
    return cls == CLS_SPARSE ? set_sparse(chunk_id) : cls == CLS_ROOMY ? set_roomy(size_id) : set_busy(size_id);
  }

  // --------------------------------------------------------------- locks
  static void lock(std::atomic<u32>& m) {
    while (m.exchange(1, std::memory_order_acquire) != 0) std::this_thread::yield();
  }
  static bool try_lock(std::atomic<u32>& m) { return m.exchange(1, std::memory_order_acquire) == 0; }
  static void unlock(std::atomic<u32>& m) { m.store(0, std::memory_order_release); }

  // ----------------------------------------------- slab state transitions
  // sb_try_mark_free (src/slab.cuh:89-113); caller holds sb_lock[sb]
  void try_mark_free(u32 sb, u32 size_id, u32 chunk_id, bool from_head) {
    if (d_head[sb].load() || chunk_id == SZ_NONE) return;
    u64 delta = ctr_chunk_delta(chunk_id), old2;
    do {
      u64 old = ctr[sb].fetch_add(delta);
      if (ctr_count(old) == 0) {
        d_size[sb].store(SZ_NONE); d_chunk[sb].store(SZ_NONE); d_csz[sb].store(0);
        set_add(0, sb);
        free_ep.fetch_add(1);
        set_remove(set_sparse(chunk_id), sb);
        if (size_id != SZ_NONE) { set_remove(set_roomy(size_id), sb); set_remove(set_busy(size_id), sb); }
        return;
      }
      old2 = ctr[sb].fetch_sub(delta);
    } while (ctr_count(old2) == 0);
    if (from_head && size_id != SZ_NONE) {
      u32 c = sb_class(ctr_count(old2), si[size_id]);
      if (c != CLS_FREE) set_add(class_set(c, size_id, chunk_id), sb);
    }
  }

  // sb_ctr_dec leader part (src/slab.cuh:161-224) for one slab group
  void ctr_dec(u32 sb, u32 change) {
    u64 old = ctr[sb].fetch_sub((u64)change << CTR_COUNT_SH);
    if (ctr_is_head(old)) return;
    u32 size_id = d_size[sb].load(), chunk_id = d_chunk[sb].load();
    if (size_id == SZ_NONE || chunk_id == SZ_NONE) return;
    u32 oc = ctr_count(old), nc = oc - change;
    u32 co = sb_class(oc, si[size_id]), cn = sb_class(nc, si[size_id]);
    if (co == cn) return;
    if (cn == CLS_FREE) {
      if (ctr_chunk(old) != CTR_CHUNK_NONE) {
        lock(sb_lock[sb]);
        size_id = d_size[sb].load(); chunk_id = d_chunk[sb].load();
        try_mark_free(sb, size_id, chunk_id, false);
        unlock(sb_lock[sb]);
      }
    } else {
      set_add(class_set(cn, size_id, chunk_id), sb);
      if (co != CLS_FREE) set_remove(class_set(co, size_id, chunk_id), sb);
    }
  }

  // drop one head reference (detach_head, src/slab.cuh:227-255); the last one
  // files the slab by its count.  Caller holds sb_lock[sb].  True if the slab
  // became free.
  bool detach_head(u32 sb) {
    u32 refs = d_head[sb].load() - 1;
    d_head[sb].store(refs);
    u64 old = ctr[sb].fetch_sub(CTR_HEAD);
    if (refs) return false;
    u32 count = ctr_count(old);
    u32 size_id = d_size[sb].load(), chunk_id = d_chunk[sb].load();
    if (size_id == SZ_NONE) return false;
    u32 c = sb_class(count, si[size_id]);
    if (c == CLS_FREE) {
      try_mark_free(sb, size_id, chunk_id, true);
      return d_chunk[sb].load() == SZ_NONE;
    }
    set_add(class_set(c, size_id, chunk_id), sb);
    return false;
  }

  // share the head of another XCD's pair of the same class if it has room
  // (caller holds the head lock of (ihead, size_id))
  u32 share_head(u32 ihead, u32 size_id, u32 chunk_id) {
    const SizeInfo& s = si[size_id];
    for (u32 ih = 0; ih < nheads; ++ih) {
      if (ih == ihead) continue;
      u32 w = head[(size_t)ih * NSIZES + size_id].load();
      if (w == SB_NONE) continue;
      u32 sb = w & ~HEAD_EXT;
      if (ctr_count(ctr[sb].load()) >= head_limit(w, s)) continue;
      lock(sb_lock[sb]);
      u32 refs = d_head[sb].load();
      bool ok = refs > 0 && refs < 255 && d_size[sb].load() == size_id && ctr_chunk(ctr[sb].load()) == chunk_id;
      if (ok) { d_head[sb].store(refs + 1); ctr[sb].fetch_add(CTR_HEAD); }
      unlock(sb_lock[sb]);
      if (ok) return w;
    }
    return SB_NONE;
  }

  // take over empty heads and cached heads of other (head, class) pairs until
  // one slab becomes free; false if none does.  other_idle: only slabs
  // another class holds without using them (its cached heads, and its empty
  // heads on pairs outside the spread).  Caller holds the head lock of
  // (ihead, size_id).
  bool steal_idle_head(u32 ihead, u32 size_id, bool other_idle = false) {
    const u32 sp = opts.fixed_sets ? nheads : spread.load();
    for (u32 ih = 0; ih < nheads; ++ih) {
      for (u32 c = 0; c < NSIZES; ++c) {
        if (ih == ihead && c == size_id) continue;
        if (other_idle && c == size_id) continue;
        size_t hi = (size_t)ih * NSIZES + c;
        for (int which = 0; which < 2; ++which) {
          if (other_idle && which == 1 && ih < sp) continue;
          std::atomic<u32>& word = which == 0 ? cached[hi] : head[hi];
          u32 w = word.load();
          if (w == SB_NONE) continue;
          u32 sb = w & ~HEAD_EXT;
          if (ctr_count(ctr[sb].load()) != 0) continue;
          if (!try_lock(head_lock[hi])) continue;
          bool ok = word.load() == w && ctr_count(ctr[sb].load()) == 0;
          bool freed = false;
          if (ok) {
            word.store(SB_NONE);
            if (which == 1) active.fetch_sub(1);  // the pair no longer holds a head
            lock(sb_lock[sb]);
            freed = detach_head(sb);
            unlock(sb_lock[sb]);
          }
          unlock(head_lock[hi]);
          if (freed) return true;
        }
      }
    }
    return false;
  }

  u32 free_slab_count() {
    u32 n = 0;
    std::atomic<u64>* w = set_words(0);
    for (u32 i = 0; i < nsw; ++i) n += (u32)__builtin_popcountll(w[i].load());
    return n;
  }
  // with several head sets, free slabs are scarce below the largest of
  // nheads, half the (head set, class) pairs that hold a head, and a quarter
  // of the heap
  bool free_scarce() {
    if (nheads <= 1) return false;
    u32 lim = active.load() / 2;
    lim = lim > nheads ? lim : nheads;
    lim = lim > nsbs / 4 ? lim : nsbs / 4;
    return free_slab_count() < lim;
  }

  // find_sb_for_size (src/slab.cuh:259-372): roomy -> sparse -> free -> busy;
  // then, for a head needed now, another XCD's head of the class, then an
  // idle head of another class.  A cache refill (now == false) skips the free
  // set when free slabs are scarce; a head needed now then first frees one
  // slab another class holds idle and, if none, tries sharing
  u32 find_sb_for_size(u32 size_id, u32 chunk_id, u32 ihead = 0, bool now = false) {
    const SizeInfo& s = si[size_id];
    u32 sb;
    bool tried_share = false, tried_idle = false;
    u32 nsteal = 0;
  again:
    while ((sb = set_pop(set_roomy(size_id))) != SB_NONE) {
      lock(sb_lock[sb]);
      bool found = false;
      if (!d_head[sb].load() && d_size[sb].load() == size_id) {
        if (ctr_count(ctr[sb].load()) <= s.roomy_thr) {
          found = true; d_head[sb].store(1); ctr[sb].fetch_add(CTR_HEAD);
        } else {
          set_add(set_busy(size_id), sb);
        }
      }
      unlock(sb_lock[sb]);
      if (found) return sb;
    }
    while ((sb = set_pop(set_sparse(chunk_id))) != SB_NONE) {
      lock(sb_lock[sb]);
      bool found = false;
      if (!d_head[sb].load() && d_chunk[sb].load() == chunk_id) {
        u32 count = ctr_count(ctr[sb].load());
        if (count <= s.sparse_thr) {
          found = true; d_head[sb].store(1); d_size[sb].store(size_id); ctr[sb].fetch_add(CTR_HEAD);
        } else {
          // not sparse for this class: hand it to its own class's roomy set, as the
          // reference does (src/slab.cuh:307-308); re-adding it here would loop
          set_add(set_roomy(d_size[sb].load()), sb);
        }
      }
      unlock(sb_lock[sb]);
      if (found) return sb;
    }
    if (now) {
      const u32 a = adopt_head(ihead, size_id);
      if (a != SB_NONE) return a;
    }
    const bool scarce = free_scarce();
    if (scarce && !now) return SB_NONE;
    bool freed_idle = false;
    if (scarce && now && !tried_idle) {
      tried_idle = true;
      freed_idle = steal_idle_head(ihead, size_id, true);
    }
    if (scarce && now && !freed_idle && !tried_share) {
      tried_share = true;
      u32 shared = share_head(ihead, size_id, chunk_id);
      if (shared != SB_NONE) return shared;
    }
    while ((sb = set_pop(0)) != SB_NONE) {
      lock(sb_lock[sb]);
      bool found = false;
      if (!d_head[sb].load()) {
        found = true;
        d_head[sb].store(1); d_size[sb].store(size_id); d_chunk[sb].store(chunk_id); d_csz[sb].store(s.chunk_sz);
        ctr[sb].fetch_add(CTR_HEAD);
        ctr[sb].fetch_sub(ctr_chunk_delta(chunk_id));
      }
      unlock(sb_lock[sb]);
      if (found) return sb;
    }
    // busy: peek, claim only a slab with room, keep the others in the set
    std::atomic<u64>* w = set_words(set_busy(size_id));
    u32 limit = (s.busy_thr + s.nchunks) / 2;
    for (u32 i = 0; i < nsw; ++i) {
      u64 word = w[i].load();
      while (word) {
        u32 b = (u32)__builtin_ctzll(word);
        u64 m = 1ull << b;
        word &= ~m;
        sb = i * 64 + b;
        if (ctr_count(ctr[sb].load()) >= limit) continue;  // a head is swapped out at limit
        if (!(w[i].fetch_and(~m) & m)) continue;
        lock(sb_lock[sb]);
        bool found = false;
        if (!d_head[sb].load() && d_size[sb].load() == size_id) {
          found = true; d_head[sb].store(1); ctr[sb].fetch_add(CTR_HEAD);
        }
        unlock(sb_lock[sb]);
        if (found) return sb | HEAD_EXT;
      }
    }
    if (!now) return SB_NONE;
    if (!tried_share) {
      tried_share = true;
      u32 shared = share_head(ihead, size_id, chunk_id);
      if (shared != SB_NONE) return shared;
    }
    if (nsteal < MAX_STEALS && steal_idle_head(ihead, size_id)) { ++nsteal; goto again; }
    return SB_NONE;
  }

  // ------------------------------------------------------------ set spread
  // Adaptive sets (halloc.hip note_class / adopt_head): the spread starts at
  // nheads (at most what one class allows); when a size class takes its
  // first head, it becomes the largest power of two s (at most nheads) with
  // 3 * s * (classes seen) <= nsbs, if that is smaller.  Pairs at and above it keep their heads until a pair
  // below adopts them (or steals them when empty).
  u32 fold_set(u32 full, u32 s) const {
    if (s >= nheads) return full;
    if (s <= 8) return full % s;
    return full % 8 + 8 * ((full / 8) % (s / 8));
  }
  u32 fold_now(u32 full) const {
    const u32 f = full % nheads;
    return opts.fixed_sets ? f : fold_set(f, spread.load());
  }
  u32 spread_for(u32 ncl) const {
    u32 per = nsbs / (3 * (ncl ? ncl : 1)), best = 1;
    if (per >= nheads) return nheads;
    for (u32 s = 1; s <= nheads && s <= per; s *= 2) best = s;
    return best >= nheads ? nheads : best;
  }
  void note_class(u32 c) {
    if (opts.fixed_sets) return;
    const u32 bit = 1u << c;
    if (class_seen.load() & bit) return;
    const u32 old = class_seen.fetch_or(bit);
    if (old & bit) return;
    const u32 want = spread_for((u32)__builtin_popcount(old | bit));
    u32 cur = spread.load();
    while (want < cur && !spread.compare_exchange_strong(cur, want)) {}
  }
  // take over the head of class size_id held by a pair outside the spread
  // (the word moves, the slab keeps its reference); caller holds the head
  // lock of (ihead, size_id), the other pair's is only tried
  u32 adopt_head(u32 ihead, u32 size_id) {
    if (opts.fixed_sets) return SB_NONE;
    const SizeInfo& s = si[size_id];
    for (u32 ih = spread.load(); ih < nheads; ++ih) {
      if (ih == ihead) continue;
      const size_t hi = (size_t)ih * NSIZES + size_id;
      u32 w = head[hi].load();
      if (w == SB_NONE || ctr_count(ctr[w & ~HEAD_EXT].load()) >= head_limit(w, s)) continue;
      if (!try_lock(head_lock[hi])) continue;
      w = head[hi].load();
      const bool ok = w != SB_NONE;
      if (ok) { head[hi].store(SB_NONE); active.fetch_sub(1); }
      unlock(head_lock[hi]);
      if (ok) return w;
    }
    return SB_NONE;
  }

  // new_sb_for_size (src/slab.cuh:378-445).  `seen` is the head word the
  // caller last used: a head replaced since then is returned as it is; a
  // thread that finds the head lock taken waits for a change, and gets
  // SB_NONE if a slab search failed meanwhile, else looks again.
  // force: the caller's probes ran out of tries in `seen`, which is then
  // replaced even below its limit (halloc.hip new_sb_for_size)
  u32 new_sb_for_size(u32 size_id, u32 ihead, u32 seen, bool force = false) {
    const SizeInfo& s = si[size_id];
    if (!opts.fixed_sets) {
      // a pair outside the spread (a stale wave's) is served from the pair
      // its set folds onto, `seen` kept: when it is the folded pair's head it
      // is replaced as for any caller of that pair
      const u32 sp = spread.load();
      if (ihead >= sp) ihead = fold_set(ihead, sp);
    }
    size_t hi = (size_t)ihead * NSIZES + size_id;
    for (;;) {
      u32 cur = head[hi].load();
      if (cur != seen && cur != SB_NONE) return cur;
      const u32 e0 = search_ep[hi].load();
      if (try_lock(head_lock[hi])) {
        cur = head[hi].load();
        u32 nh;
        // below: a forced replacement of a head that still has room
        const bool below = cur != SB_NONE && ctr_count(ctr[cur & ~HEAD_EXT].load()) < head_limit(cur, s);
        const u32 fe = free_ep.load();
        if (below && forced_head[hi].load() == cur && forced_ep[hi].load() == fe) {
          nh = cur;  // this head's forced search failed and no slab became free since
        } else if (cur == SB_NONE || !below || (force && cur == seen)) {
          nh = cached[hi].load();
          cached[hi].store(SB_NONE);
          if (nh == SB_NONE) {
            // one head search per class at a time, so a pair that follows can
            // share the head the one before it took
            if (nheads > 1) lock(class_lock[size_id]);
            nh = find_sb_for_size(size_id, s.chunk_id, ihead, true);
            if (nheads > 1) unlock(class_lock[size_id]);
            search_ep[hi].store((((search_ep[hi].load() >> 1) + 1) << 1) | (nh == SB_NONE && !below ? 1u : 0u));
          }
          if (nh == SB_NONE && below) {
            nh = cur;  // no other slab: keep the head, which still has room
            forced_ep[hi].store(fe);
            forced_head[hi].store(cur);
          } else if (nh != SB_NONE) {
            head[hi].store(nh);
            if (cur == SB_NONE) { active.fetch_add(1); note_class(size_id); }
            if (cur != SB_NONE) {
              u32 csb = cur & ~HEAD_EXT;
              lock(sb_lock[csb]); detach_head(csb); unlock(sb_lock[csb]);
            }
            cached[hi].store(find_sb_for_size(size_id, s.chunk_id));
          }
        } else {
          nh = cur;
        }
        unlock(head_lock[hi]);
        return nh;
      }
      while (head[hi].load() == cur && head_lock[hi].load() != 0) std::this_thread::yield();
      const u32 e1 = search_ep[hi].load();
      if (e1 != e0 && (e1 & 1u) && head[hi].load() == cur) return SB_NONE;
    }
  }

  // ---------------------------------------------------------- malloc (wave)
  static u32 rank_of(u64 g, u32 l) { return (u32)__builtin_popcountll(g & ((1ull << l) - 1)); }

  // One probe sequence (sb_alloc_in, src/slab.cuh:453-535) in slab hsb for the
  // lanes P with initial chunks ich0; ich receives each lane's last position.
  // Returns the lanes that reserved; itry advances for the whole group.
  // skip_first: the lanes P already failed a first attempt at ich0 (made and
  // rolled back by the caller); go on from the second probe.
  u64 probe_in(const SizeInfo& s, u32 hsb, u32 limit, u64 P, const u32* ich0, u32* ich, u32& itry,
               bool skip_first = false, bool* spent = nullptr) {
    const u64 mask1 = (1ull << s.ncib) - 1;
    std::atomic<u64>* bw = &bits[(size_t)hsb * bit_words];
    u64 probing = P, reserved = 0;
    for (u32 l = 0; l < WAVE; ++l) if (probing >> l & 1) ich[l] = ich0[l];
    for (bool first = true;; first = false) {
      if (skip_first && first) goto next_probe;
      {
      u64 oldw[WAVE];
      for (u32 l = 0; l < WAVE; ++l) if (probing >> l & 1) {
        u64 m = mask1 << (ich[l] % 64);
        oldw[l] = bw[ich[l] / 64].fetch_or(m);
      }
      u64 now = 0;
      for (u32 l = 0; l < WAVE; ++l) if (probing >> l & 1) {
        u64 m = mask1 << (ich[l] % 64);
        if (!(oldw[l] & m)) now |= 1ull << l;
        else if (~oldw[l] & m) bw[ich[l] / 64].fetch_and(~(m & ~oldw[l]));
      }
      reserved |= now;
      probing &= ~now;
      if (!probing) break;
      }
    next_probe:
      if (itry % CHECK_NTRIES == CHECK_NTRIES - 1) {
        if (ctr_count(ctr[hsb].load()) >= limit) break;
      }
      for (u32 l = 0; l < WAVE; ++l) if (probing >> l & 1) ich[l] = probe_next(ich0[l], itry, s);
      if (++itry % MAX_NTRIES >= MAX_NTRIES - 1) {
        if (spent) *spent = true;
        break;
      }
    }
    return reserved;
  }

  // record blocks that passed the counter check
  void commit_blocks(u32 size_id, u32 hsb, u64 reserved, const u32* ich, u64* out, Rec* rec) {
    const SizeInfo& s = si[size_id];
    std::atomic<u64>* nw = &nibs[(size_t)hsb * nib_words];
    for (u32 l = 0; l < WAVE; ++l) if (reserved >> l & 1) {
      out[l] = (u64)hsb * sb_sz + chunk_mul(ich[l], s.chunk_sz) + 1;
      nw[ich[l] / 16].fetch_or((u64)s.ncib << (4 * (ich[l] % 16)));
      if (rec) rec[l] = Rec{size_id, hsb, ich[l], 0};
    }
  }
  void release_bits(const SizeInfo& s, u32 hsb, u64 reserved, const u32* ich) {
    const u64 mask1 = (1ull << s.ncib) - 1;
    std::atomic<u64>* bw = &bits[(size_t)hsb * bit_words];
    for (u32 l = 0; l < WAVE; ++l) if (reserved >> l & 1) bw[ich[l] / 64].fetch_and(~(mask1 << (ich[l] % 64)));
  }

  // one size-class group of one wave (hamalloc_small, src/halloc.cu:125-182)
  void malloc_group(u32 ihead, u32 size_id, u64 g, u64* out, Rec* rec) {
    const SizeInfo& s = si[size_id];
    size_t hi = (size_t)ihead * NSIZES + size_id;
    u32 hword = head[hi].load();
    u32 ich0[WAVE] = {0}, ich[WAVE] = {0};
    // tickets (size_ctr_inc, src/halloc.cu:71-98)
    if (opts.hash_mode == 0) {
      u32 base = ticket[hi].fetch_add((u32)__builtin_popcountll(g));
      for (u32 l = 0; l < WAVE; ++l) if (g >> l & 1) ich0[l] = ichunk_init(base + rank_of(g, l), s);
    } else {
      u32 last = 63 - (u32)__builtin_clzll(g);
      u32 base = ticket[hi].fetch_add(last + 1);
      for (u32 l = 0; l < WAVE; ++l) if (g >> l & 1) ich0[l] = ichunk_init(base + l, s);
    }
    u32 itry = 0;
    u64 want = g;
    while (want) {
      bool need = false, spent = false, crowded = false;
      const u32 hsb = hword == SB_NONE ? SB_NONE : (hword & ~HEAD_EXT);
      const u32 limit = head_limit(hword, s);
      if (hsb == SB_NONE) {
        need = true;
      } else {
        const u32 itry0 = itry;
        u64 reserved = probe_in(s, hsb, limit, want, ich0, ich, itry, false, &spent);
        if (reserved != want) need = true;
        if (itry - itry0 >= CROWDED_NTRIES) {
          need = crowded = true;
          n_crowded.fetch_add(1);
        }
        if (reserved) {
          // sb_ctr_inc (src/slab.cuh:122-155), one group per (slab, ncib)
          u32 change = s.ncib * (u32)__builtin_popcountll(reserved);
          u64 old = ctr[hsb].fetch_add((u64)change << CTR_COUNT_SH);
          if (ctr_chunk(old) != s.chunk_id) {
            release_bits(s, hsb, reserved, ich);
            ctr_dec(hsb, change);
            need = true;
          } else {
            u32 oc = ctr_count(old);
            if (oc < limit && oc + change >= limit) need = true;
            commit_blocks(size_id, hsb, reserved, ich, out, rec);
            want &= ~reserved;
          }
        }
      }
      if (need) {
        u32 nh = new_sb_for_size(size_id, ihead, hword, (spent && want != 0) || crowded);
        hword = nh;
        if (nh == SB_NONE) {
          for (u32 l = 0; l < WAVE; ++l) if (want >> l & 1) {
            out[l] = 0;
            if (rec) rec[l] = Rec{REC_NULL, REC_NULL, REC_NULL, 0};
          }
          want = 0;
        }
      }
    }
  }

  // hamalloc_batch for one size-class group: lane l wants cnt[l] (<= MAXB)
  // blocks.  The group takes ONE ticket for all R rounds (round j is based at
  // old + j*(highest lane + 1), where R back-to-back calls would be), probes every
  // round in the current head, and adds all its chunks to the slab counter in
  // ONE atomic.  Lanes a round leaves unserved, and everything after a rollback,
  // go through plain per-call rounds.  out/rec: [round][lane].
  void malloc_batch_group(u32 ihead, u32 size_id, u64 g, const u32* cnt, u64* out, Rec* rec) {
    const SizeInfo& s = si[size_id];
    size_t hi = (size_t)ihead * NSIZES + size_id;
    u32 hword = head[hi].load();
    if (hword == SB_NONE) hword = new_sb_for_size(size_id, ihead, SB_NONE);  // no head yet: get one first
    u32 R = 0;
    for (u32 l = 0; l < WAVE; ++l) if ((g >> l & 1) && cnt[l] > R) R = cnt[l];
    const u32 L = 64 - (u32)__builtin_clzll(g);
    const u32 base = ticket[hi].fetch_add(R * L);
    u64 pend[MAXB] = {0}, resv[MAXB] = {0};
    static thread_local u32 ich_r[MAXB][WAVE];
    for (u32 j = 0; j < R; ++j)
      for (u32 l = 0; l < WAVE; ++l) if ((g >> l & 1) && cnt[l] > j) pend[j] |= 1ull << l;
    bool need = false;
    const u32 hsb = hword == SB_NONE ? SB_NONE : (hword & ~HEAD_EXT);
    const u32 limit = head_limit(hword, s);
    if (hsb == SB_NONE) {
      need = true;
    } else {
      u32 total = 0;
      const u64 mask1 = (1ull << s.ncib) - 1;
      std::atomic<u64>* bw = &bits[(size_t)hsb * bit_words];
      // rounds in groups of BATCH_PIPE: the first probe of every round in the
      // group is made before any is checked; then, round by round, the lanes
      // that failed give their new bits back and go on from the second probe
      for (u32 j0 = 0; j0 < R; j0 += BATCH_PIPE) {
        const u32 j1 = j0 + BATCH_PIPE < R ? j0 + BATCH_PIPE : R;
        static thread_local u32 ich0_r[BATCH_PIPE][WAVE];
        static thread_local u64 oldw_r[BATCH_PIPE][WAVE];
        for (u32 j = j0; j < j1; ++j)
          for (u32 l = 0; l < WAVE; ++l) if (pend[j] >> l & 1) {
            const u32 c = ichunk_init(base + j * L + l, s);
            ich0_r[j - j0][l] = c;
            ich_r[j][l] = c;
            oldw_r[j - j0][l] = bw[c / 64].fetch_or(mask1 << (c % 64));
          }
        for (u32 j = j0; j < j1; ++j) {
          u64 ok = 0;
          for (u32 l = 0; l < WAVE; ++l) if (pend[j] >> l & 1) {
            const u32 c = ich0_r[j - j0][l];
            const u64 m = mask1 << (c % 64), ow = oldw_r[j - j0][l];
            if (!(ow & m)) ok |= 1ull << l;
            else if (~ow & m) bw[c / 64].fetch_and(~(m & ~ow));
          }
          resv[j] = ok;
          const u64 fail = pend[j] & ~ok;
          if (fail) {
            u32 itry = 0;
            u32 ich_f[WAVE];
            resv[j] |= probe_in(s, hsb, limit, fail, ich0_r[j - j0], ich_f, itry, true);
            for (u32 l = 0; l < WAVE; ++l) if (fail >> l & 1) ich_r[j][l] = ich_f[l];
          }
          if (resv[j] != pend[j]) need = true;
          total += (u32)__builtin_popcountll(resv[j]);
        }
      }
      if (total) {
        u32 change = s.ncib * total;
        u64 old = ctr[hsb].fetch_add((u64)change << CTR_COUNT_SH);
        if (ctr_chunk(old) != s.chunk_id) {
          for (u32 j = 0; j < R; ++j) release_bits(s, hsb, resv[j], ich_r[j]);
          ctr_dec(hsb, change);
          need = true;
        } else {
          u32 oc = ctr_count(old);
          if (oc < limit && oc + change >= limit) need = true;
          for (u32 j = 0; j < R; ++j) {
            commit_blocks(size_id, hsb, resv[j], ich_r[j], out + j * WAVE, rec ? rec + j * WAVE : nullptr);
            pend[j] &= ~resv[j];
          }
        }
      }
    }
    if (need) new_sb_for_size(size_id, ihead, hword);
    for (u32 j = 0; j < R; ++j)
      if (pend[j]) malloc_group(ihead, size_id, pend[j], out + j * WAVE, rec ? rec + j * WAVE : nullptr);
  }

  // ----------------------------------------------------------- large region
  // One class group of large requests (halloc.hip large_group): lanes g want a
  // block of class c.  out[l] = arena offset + 1 (0 = NULL).
  void large_group(u32 c, u64 g, u64* out, Rec* rec) {
    const u32 nw = lwords;
    u32 old = lticket[c].fetch_add((u32)__builtin_popcountll(g));
    u64 unit[WAVE];
    for (u32 l = 0; l < WAVE; ++l) unit[l] = UNIT_NONE;
    if (c <= 6) {
      const u32 k = 1u << c;
      const u64 fm = k == 64 ? ~0ull : ((1ull << k) - 1);  // one field's bits at position 0
      u32 w = large_start(old, c, nw), moved = 0;
      u64 v = lbits[w].load();
      u64 pend = g;
      while (pend) {
        // free aligned fields of the known word value, lowest first
        u32 freep[64], n = 0;
        for (u32 p = 0; p < 64; p += k) if (((v >> p) & fm) == 0) freep[n++] = p;
        if (n == 0) {
          if (++moved >= nw) break;
          w = w + 1 == nw ? 0 : w + 1;
          v = lbits[w].load();
          continue;
        }
        // the lane of rank r among the pending lanes takes field r
        u32 lanes[64], np = 0;
        for (u32 l = 0; l < WAVE; ++l) if (pend >> l & 1) lanes[np++] = l;
        const u32 npick = np < n ? np : n;
        u64 U = 0;
        for (u32 r = 0; r < npick; ++r) U |= fm << freep[r];
        const u64 o = lbits[w].fetch_or(U);
        u64 back = 0, got = 0;
        v = o;  // the word as it is now known: what the OR found, plus the fields won
        for (u32 r = 0; r < npick; ++r) {
          const u64 f = fm << freep[r];
          if (o & f) {
            back |= f & ~o;  // give back the bits set in a field found taken
          } else {
            unit[lanes[r]] = (u64)w * 64 + freep[r];
            got |= 1ull << lanes[r];
            v |= f;
          }
        }
        if (back) lbits[w].fetch_and(~back);
        pend &= ~got;
      }
    } else {
      const u32 m = 1u << (c - 6);  // words per block
      const u32 ngr = nw / m;
      u32 r = 0;
      for (u32 l = 0; l < WAVE; ++l) if (g >> l & 1) {
        const u32 rank = r++;
        if (ngr == 0) continue;
        const u32 g0 = large_start(old + rank, c, ngr);
        for (u32 i = 0; i < ngr && unit[l] == UNIT_NONE; ++i) {
          const u32 grp = (g0 + i) % ngr;
          const size_t base = (size_t)grp * m;
          bool zero = true;
          for (u32 j = 0; j < m && zero; ++j) zero = lbits[base + j].load() == 0;
          if (!zero) continue;
          u32 j = 0;
          for (; j < m; ++j) {
            u64 e = 0;
            if (!lbits[base + j].compare_exchange_strong(e, ~0ull)) break;
          }
          if (j == m) unit[l] = (u64)base * 64;
          else while (j--) lbits[base + j].store(0);
        }
      }
    }
    for (u32 l = 0; l < WAVE; ++l) if (g >> l & 1) {
      if (unit[l] != UNIT_NONE) {
        lstart[unit[l]].store(c + 1);
        out[l] = arena_bytes + (unit[l] << LARGE_UNIT_SH) + 1;
        if (rec) rec[l] = Rec{NSIZES + c, LARGE_SB, (u32)unit[l], 0};
      } else {
        out[l] = 0;
        if (rec) rec[l] = Rec{REC_NULL, REC_NULL, REC_NULL, 0};
      }
    }
  }

  // hamalloc above 3072 B for the lanes `active` of one wave: one class group
  // at a time, groups ordered by their lowest lane; a class past the region's
  // largest gets NULL
  void wave_large(u64 active, const u64* nbytes, u64* out, Rec* rec) {
    u32 cl[WAVE] = {0};
    u64 rem = 0;
    for (u32 l = 0; l < WAVE; ++l) if (active >> l & 1) {
      cl[l] = large_class(nbytes[l]);
      if (cl[l] < lnclasses) {
        rem |= 1ull << l;
      } else {
        out[l] = 0;
        if (rec) rec[l] = Rec{REC_NULL, REC_NULL, REC_NULL, 0};
      }
    }
    while (rem) {
      const u32 l0 = (u32)__builtin_ctzll(rem);
      u64 g = 0;
      for (u32 l = 0; l < WAVE; ++l) if ((rem >> l & 1) && cl[l] == cl[l0]) g |= 1ull << l;
      rem &= ~g;
      large_group(cl[l0], g, out, rec);
    }
  }

  // hafree of a large block at region offset loff (halloc.hip large_free)
  void large_free(u64 loff) {
    if (loff & ((1ull << LARGE_UNIT_SH) - 1)) return;
    const u64 unit = loff >> LARGE_UNIT_SH;
    const u32 st = lstart[unit].exchange(0);
    if (st == 0) return;
    const u32 c = st - 1;
    for (u64 u = unit; u < unit + (1ull << c); ++u) lbits[u / 64].fetch_and(~(1ull << (u % 64)));
  }

  // hamalloc_batch for the active lanes of one wave: lane l asks for cnt[l]
  // blocks of nbytes[l] (classes above 3072 B go to the device heap one by one)
  void wave_malloc_batch(u32 ihead, u64 active, const u32* nbytes, const u32* cnt, u64* out, Rec* rec) {
    const u32 ih = fold_now(ihead);  // once per call, as the device folds it at entry
    u32 sid[WAVE] = {0};
    u64 small = 0, large = 0;
    for (u32 l = 0; l < WAVE; ++l) if (active >> l & 1) {
      u32 c = cnt[l] > MAXB ? MAXB : cnt[l];
      if (!c) continue;
      if (nbytes[l] > MAX_BLOCK_SZ) {
        large |= 1ull << l;
      } else {
        sid[l] = size_id_from_nbytes(nbytes[l]);
        small |= 1ull << l;
      }
    }
    u32 c2[WAVE];
    for (u32 l = 0; l < WAVE; ++l) c2[l] = cnt[l] > MAXB ? MAXB : cnt[l];
    // large requests: round j serves the lanes that want more than j blocks
    u64 nb64[WAVE];
    for (u32 l = 0; l < WAVE; ++l) nb64[l] = nbytes[l];
    for (u32 j = 0; j < MAXB; ++j) {
      u64 a = 0;
      for (u32 l = 0; l < WAVE; ++l) if ((large >> l & 1) && c2[l] > j) a |= 1ull << l;
      if (a) wave_large(a, nb64, out + j * WAVE, rec ? rec + j * WAVE : nullptr);
    }
    while (small) {
      u32 l0 = (u32)__builtin_ctzll(small);
      u64 g = 0;
      for (u32 l = 0; l < WAVE; ++l) if ((small >> l & 1) && sid[l] == sid[l0]) g |= 1ull << l;
      small &= ~g;
      malloc_batch_group(ih, sid[l0], g, c2, out, rec);
    }
  }

  // hamalloc for the active lanes of one wave; out[] = arena offset + 1 (0 = NULL).
  // raw: use pair ihead as it is, without folding it onto the spread (a wave
  // that read the spread before it shrank)
  void wave_malloc(u32 ihead, u64 active, const u32* nbytes, u64* out, Rec* rec, bool raw = false) {
    const u32 ih = raw ? ihead % nheads : fold_now(ihead);  // once per call, as the device folds it at entry
    u32 sid[WAVE] = {0};
    u64 small = 0, large = 0;
    u64 nb64[WAVE] = {0};
    for (u32 l = 0; l < WAVE; ++l) if (active >> l & 1) {
      if (nbytes[l] > MAX_BLOCK_SZ) {
        large |= 1ull << l;
        nb64[l] = nbytes[l];
      } else {
        sid[l] = size_id_from_nbytes(nbytes[l]);
        small |= 1ull << l;
      }
    }
    if (large) wave_large(large, nb64, out, rec);
    while (small) {
      u32 l0 = (u32)__builtin_ctzll(small);
      u64 g = 0;
      for (u32 l = 0; l < WAVE; ++l) if ((small >> l & 1) && sid[l] == sid[l0]) g |= 1ull << l;
      small &= ~g;
      malloc_group(ih, sid[l0], g, out, rec);
    }
  }

  // hafree for the active lanes of one wave (src/halloc.cu:211-256)
  void wave_free(u64 active, const u64* ptr, Rec* rec) {
    u32 sbl[WAVE] = {0};
    u64 small = 0, large = 0;
    for (u32 l = 0; l < WAVE; ++l) if (active >> l & 1) {
      if (rec) rec[l] = Rec{0, REC_NULL, REC_NULL, 0};
      u64 p = ptr[l];
      if (p == 0) continue;
      u64 off = p - 1;
      if (off >= arena_bytes) {
        if (off - arena_bytes < large_bytes) large |= 1ull << l;
        continue;
      }
      sbl[l] = (u32)(off >> sb_sz_sh);
      small |= 1ull << l;
    }
    // large blocks: every record is read before any of the wave's frees
    for (u32 l = 0; l < WAVE; ++l) if ((large >> l & 1) && rec) {
      const u64 loff = ptr[l] - 1 - arena_bytes;
      if (loff & ((1ull << LARGE_UNIT_SH) - 1)) continue;
      const u32 st = lstart[loff >> LARGE_UNIT_SH].load();
      if (st) rec[l] = Rec{1u << (st - 1), LARGE_SB, (u32)(loff >> LARGE_UNIT_SH), 0};
    }
    for (u32 l = 0; l < WAVE; ++l) if (large >> l & 1) large_free(ptr[l] - 1 - arena_bytes);
    while (small) {
      u32 l0 = (u32)__builtin_ctzll(small);
      u32 sb = sbl[l0];
      u64 g = 0;
      for (u32 l = 0; l < WAVE; ++l) if ((small >> l & 1) && sbl[l] == sb) g |= 1ull << l;
      small &= ~g;
      u32 cs = d_csz[sb].load();
      if (cs == 0) continue;  // slab is free: nothing is allocated in it
      std::atomic<u64>* nw = &nibs[(size_t)sb * nib_words];
      std::atomic<u64>* bw = &bits[(size_t)sb * bit_words];
      u32 change = 0;
      u32 ich[WAVE] = {0}, nch[WAVE] = {0};
      for (u32 l = 0; l < WAVE; ++l) if (g >> l & 1) {
        u32 in_sb = (u32)((ptr[l] - 1) & (sb_sz - 1));
        ich[l] = chunk_div(in_sb, cs);
        u32 sh = 4 * (ich[l] % 16);
        nch[l] = (u32)((nw[ich[l] / 16].fetch_and(~(0xFull << sh)) >> sh) & 0xF);
      }
      for (u32 l = 0; l < WAVE; ++l) if ((g >> l & 1) && nch[l]) {
        bw[ich[l] / 64].fetch_and(~(((1ull << nch[l]) - 1) << (ich[l] % 64)));
        change += nch[l];
        if (rec) rec[l] = Rec{nch[l], sb, ich[l], 0};
      }
      if (change) ctr_dec(sb, change);
    }
  }

  // -------------------------------------------------------------- statistics
  // free bytes over slabs (src/statistics.cuh:34-48)
  u64 free_bytes() const {
    u64 total = 0;
    for (u32 b = 0; b < nsbs; ++b) {
      u32 cs = d_csz[b].load();
      u32 used = cs ? chunk_mul(ctr_count(ctr[b].load()), cs) : 0;
      total += sb_sz - used;
    }
    return total;
  }
};

}  // namespace hamodel
