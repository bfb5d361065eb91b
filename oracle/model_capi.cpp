// model_capi.cpp -- C entry points of the host model (TEST INFRASTRUCTURE ONLY).
//
// Loaded through ctypes by tests/ and by bench.py's CPU-baseline leg.  Nothing
// in the product library links or calls this file.
#include <algorithm>
#include <chrono>
#include <cstdlib>
#include <thread>
#include <vector>

#include "halloc_model.hpp"

using namespace hamodel;
using u8 = uint8_t;

extern "C" {

// mirrors ha_model_opts in tests/oracle.py
struct hm_opts_t {
  uint64_t memory;
  double halloc_fraction, busy_fraction, roomy_fraction, sparse_fraction;
  int32_t sb_sz_sh, nheads, hash_mode, fixed_sets;
  double hash_step_frac;
};
// kind 0 none, 1 malloc(arg = bytes), 2 free(arg = op index), 3 free of the
// pointer 4096 bytes into op arg's block (an interior pointer of a large block)
struct hm_op_t { uint32_t kind, arg; };

static Opts to_opts(const hm_opts_t* o) {
  Opts r;
  r.memory = o->memory; r.halloc_fraction = o->halloc_fraction; r.busy_fraction = o->busy_fraction;
  r.roomy_fraction = o->roomy_fraction; r.sparse_fraction = o->sparse_fraction;
  r.sb_sz_sh = o->sb_sz_sh; r.nheads = o->nheads; r.hash_mode = o->hash_mode;
  r.hash_step_frac = o->hash_step_frac;
  r.fixed_sets = o->fixed_sets;
  return r;
}

// ---------------------------------------------------------------- closed forms
uint32_t hm_size_id(uint32_t nbytes) { return size_id_from_nbytes(nbytes); }
uint32_t hm_chunk_val(uint32_t sz) { return chunk_val(sz); }
uint32_t hm_chunk_mul(uint32_t v, uint32_t cs) { return chunk_mul(v, cs); }
uint32_t hm_chunk_div(uint32_t v, uint32_t cs) { return chunk_div(v, cs); }
uint32_t hm_max_prime_below(uint32_t n, uint32_t nb) { return max_prime_below(n, nb); }
uint32_t hm_ichunk_init(uint32_t cv, uint32_t sb_sz_sh, uint32_t size_id, int32_t hash_mode) {
  Opts o; o.sb_sz_sh = (int)sb_sz_sh; o.hash_mode = hash_mode;
  SizeInfo si[NSIZES];
  if (!build_size_infos(1u << sb_sz_sh, o, si)) return 0xffffffffu;
  return ichunk_init(cv, si[size_id]);
}
uint32_t hm_probe_next(uint32_t ich0, uint32_t itry, uint32_t sb_sz_sh, uint32_t size_id, int32_t hash_mode) {
  Opts o; o.sb_sz_sh = (int)sb_sz_sh; o.hash_mode = hash_mode;
  SizeInfo si[NSIZES];
  if (!build_size_infos(1u << sb_sz_sh, o, si)) return 0xffffffffu;
  return probe_next(ich0, itry, si[size_id]);
}
// 16 rows x 12 u32: nchunks chunk_sz chunk_id ncib sparse hash_step roomy busy block_sz nblocks tmod chunk_bytes
int hm_size_table(const hm_opts_t* o, uint32_t* out) {
  Opts op = to_opts(o);
  SizeInfo si[NSIZES];
  if (!build_size_infos(1u << op.sb_sz_sh, op, si)) return -1;
  for (u32 s = 0; s < NSIZES; ++s) std::memcpy(out + 12 * s, &si[s], 12 * sizeof(u32));
  return 0;
}

// ---------------------------------------------------------------- heap objects
void* hm_create(const hm_opts_t* o) {
  Heap* h = new Heap(to_opts(o));
  if (!h->ok) { delete h; return nullptr; }
  return h;
}
void hm_destroy(void* h) { delete (Heap*)h; }
uint32_t hm_nsbs(void* h) { return ((Heap*)h)->nsbs; }
// probe sequences so far that marked their head crowded (the crowded-head rule)
uint64_t hm_crowded(void* h) { return ((Heap*)h)->n_crowded.load(); }

// Replay nsteps x lanes operations: per step, the malloc lanes run as one wave
// call, then the free lanes run as one wave call.  recs: 4 u32 per op.
// ptrs: per-op arena offset + 1 (0 = NULL), kept so later frees can refer to it.
int hm_replay(void* hv, const hm_op_t* ops, const uint32_t* iheads, int nsteps, int lanes,
              uint32_t* recs, uint64_t* ptrs) {
  Heap* h = (Heap*)hv;
  if (lanes < 1 || lanes > 64) return -1;
  for (int st = 0; st < nsteps; ++st) {
    const hm_op_t* row = ops + (size_t)st * lanes;
    u64 am = 0, fm = 0;
    u32 nbytes[WAVE] = {0};
    u64 outp[WAVE] = {0}, fptr[WAVE] = {0};
    Rec mrec[WAVE], frec[WAVE];
    std::memset(mrec, 0, sizeof mrec); std::memset(frec, 0, sizeof frec);
    for (int l = 0; l < lanes; ++l) {
      if (row[l].kind == 1) { am |= 1ull << l; nbytes[l] = row[l].arg; }
      else if (row[l].kind == 2 || row[l].kind == 3) {
        if ((size_t)row[l].arg >= (size_t)st * lanes) return -2;  // must name an earlier op
        fm |= 1ull << l; fptr[l] = ptrs[row[l].arg];
        // kind 3: a pointer one large unit (4 KiB) into the block
        if (row[l].kind == 3 && fptr[l]) fptr[l] += 1ull << LARGE_UNIT_SH;
      }
    }
    if (am) h->wave_malloc(iheads ? iheads[st] : 0, am, nbytes, outp, mrec);
    for (int l = 0; l < lanes; ++l) if (am >> l & 1) {
      size_t i = (size_t)st * lanes + l;
      ptrs[i] = outp[l];
      std::memcpy(recs + 4 * i, &mrec[l], 16);
    }
    if (fm) h->wave_free(fm, fptr, frec);
    for (int l = 0; l < lanes; ++l) if (fm >> l & 1) {
      size_t i = (size_t)st * lanes + l;
      ptrs[i] = 0;
      std::memcpy(recs + 4 * i, &frec[l], 16);
    }
  }
  return 0;
}

// One wave step with explicit pointers: the lanes with kind 1 call malloc of
// arg bytes together, then the lanes with kind 2 free the pointer arg (arena
// offset + 1, as returned in ptrs; 0 is NULL).  recs: 4 u32 per lane;
// ptrs[l] receives lane l's malloc result.
int hm_wave(void* hv, const uint32_t* kinds, const uint64_t* args, int lanes, uint32_t* recs, uint64_t* ptrs) {
  Heap* h = (Heap*)hv;
  if (lanes < 1 || lanes > 64) return -1;
  u64 am = 0, fm = 0;
  u32 nbytes[WAVE] = {0};
  u64 outp[WAVE] = {0}, fptr[WAVE] = {0};
  Rec mrec[WAVE], frec[WAVE];
  std::memset(mrec, 0, sizeof mrec); std::memset(frec, 0, sizeof frec);
  for (int l = 0; l < lanes; ++l) {
    if (kinds[l] == 1) { am |= 1ull << l; nbytes[l] = (u32)args[l]; }
    else if (kinds[l] == 2) { fm |= 1ull << l; fptr[l] = args[l]; }
  }
  if (am) h->wave_malloc(0, am, nbytes, outp, mrec);
  if (fm) h->wave_free(fm, fptr, frec);
  for (int l = 0; l < lanes; ++l) {
    ptrs[l] = (am >> l & 1) ? outp[l] : 0;
    std::memcpy(recs + 4 * l, (am >> l & 1) ? &mrec[l] : &frec[l], 16);
  }
  return 0;
}

// One wave of mallocs on head-set pair `pair` as given (not folded onto the
// spread): what a wave whose spread value is stale does.  ptrs: per lane.
int hm_wave_raw(void* hv, uint32_t pair, const uint32_t* nbytes, int lanes, uint64_t* ptrs) {
  Heap* h = (Heap*)hv;
  if (lanes < 1 || lanes > 64) return -1;
  u32 nb[WAVE] = {0};
  u64 out[WAVE] = {0};
  for (int l = 0; l < lanes; ++l) nb[l] = nbytes[l];
  h->wave_malloc(pair, lanes == 64 ? ~0ull : ((1ull << lanes) - 1), nb, out, nullptr, true);
  for (int l = 0; l < lanes; ++l) ptrs[l] = out[l];
  return 0;
}

// Concurrent mallocs: nthreads threads take the waves of `sizes` (64 lanes
// each) in turn, wave w on thread w % nthreads and head set
// (w / 4) % nheads, as blocks of 256 lanes map to the XCDs.  ptrs receives
// arena offset + 1 (0 = NULL).  The threads race on the heap as waves do on
// the GPU.
int hm_malloc_threads(void* hv, int nthreads, const uint32_t* sizes, int n, uint64_t* ptrs) {
  Heap* h = (Heap*)hv;
  if (nthreads < 1 || n < 0) return -1;
  const int nwaves = (n + (int)WAVE - 1) / (int)WAVE;
  auto body = [&](int t) {
    u32 nb[WAVE];
    u64 out[WAVE];
    for (int w = t; w < nwaves; w += nthreads) {
      const int base = w * (int)WAVE, m = std::min<int>((int)WAVE, n - base);
      const u64 act = m == 64 ? ~0ull : ((1ull << m) - 1);
      for (int l = 0; l < m; ++l) nb[l] = sizes[base + l];
      h->wave_malloc((u32)(w / 4) % h->nheads, act, nb, out, nullptr);
      for (int l = 0; l < m; ++l) ptrs[base + l] = out[l];
    }
  };
  std::vector<std::thread> th;
  for (int t = 0; t < nthreads; ++t) th.emplace_back(body, t);
  for (auto& x : th) x.join();
  return 0;
}

// Concurrent frees of ptrs (arena offset + 1, 0 skipped), split as above.
int hm_free_threads(void* hv, int nthreads, const uint64_t* ptrs, int n) {
  Heap* h = (Heap*)hv;
  if (nthreads < 1 || n < 0) return -1;
  const int nwaves = (n + (int)WAVE - 1) / (int)WAVE;
  auto body = [&](int t) {
    u64 p[WAVE];
    for (int w = t; w < nwaves; w += nthreads) {
      const int base = w * (int)WAVE, m = std::min<int>((int)WAVE, n - base);
      u64 act = 0;
      for (int l = 0; l < m; ++l) {
        p[l] = ptrs[base + l];
        if (p[l]) act |= 1ull << l;
      }
      if (act) h->wave_free(act, p, nullptr);
    }
  };
  std::vector<std::thread> th;
  for (int t = 0; t < nthreads; ++t) th.emplace_back(body, t);
  for (auto& x : th) x.join();
  return 0;
}

// Batched replay: ops are 4 u32 per (step, lane): kind (0 none, 1 hamalloc_batch,
// 2 hafree_batch, 3 hafree_batch that keeps the op's pointers so a later op
// frees the same blocks again), arg (bytes, or the index of an earlier kind-1
// op whose blocks are freed), count (blocks, kind 1), pad.  recs: 4 u32 per (op, j)
// for j < MAXB; ptrs: arena offset + 1 per (op, j).
int hm_replay_batch(void* hv, const uint32_t* ops, const uint32_t* iheads, int nsteps, int lanes,
                    uint32_t* recs, uint64_t* ptrs) {
  Heap* h = (Heap*)hv;
  if (lanes < 1 || lanes > 64) return -1;
  for (int st = 0; st < nsteps; ++st) {
    const uint32_t* row = ops + (size_t)st * lanes * 4;
    u64 am = 0;
    u32 nbytes[WAVE] = {0}, cnt[WAVE] = {0}, fref[WAVE] = {0};
    u64 fm = 0, fz = 0;  // lanes that free; of them, those that forget the pointers
    for (int l = 0; l < lanes; ++l) {
      if (row[4 * l] == 1) { am |= 1ull << l; nbytes[l] = row[4 * l + 1]; cnt[l] = row[4 * l + 2]; }
      else if (row[4 * l] == 2 || row[4 * l] == 3) {
        if ((size_t)row[4 * l + 1] >= (size_t)st * lanes) return -2;
        fm |= 1ull << l; fref[l] = row[4 * l + 1];
        if (row[4 * l] == 2) fz |= 1ull << l;
      }
    }
    if (am) {
      static thread_local u64 outp[MAXB * WAVE];
      static thread_local Rec mrec[MAXB * WAVE];
      std::memset(outp, 0, sizeof outp);
      std::memset(mrec, 0, sizeof mrec);
      h->wave_malloc_batch(iheads ? iheads[st] : 0, am, nbytes, cnt, outp, mrec);
      for (int l = 0; l < lanes; ++l) if (am >> l & 1) {
        size_t i = (size_t)st * lanes + l;
        u32 c = cnt[l] > MAXB ? MAXB : cnt[l];
        for (u32 j = 0; j < c; ++j) {
          ptrs[i * MAXB + j] = outp[j * WAVE + l];
          std::memcpy(recs + 4 * (i * MAXB + j), &mrec[j * WAVE + l], 16);
        }
      }
    }
    if (fm) {
      for (u32 j = 0; j < MAXB; ++j) {
        u64 m = 0;
        u64 fp[WAVE] = {0};
        for (int l = 0; l < lanes; ++l) if (fm >> l & 1) {
          u64 p = ptrs[(size_t)fref[l] * MAXB + j];
          if (p) { m |= 1ull << l; fp[l] = p; }
        }
        if (!m) continue;
        Rec frec[WAVE];
        std::memset(frec, 0, sizeof frec);
        h->wave_free(m, fp, frec);
        for (int l = 0; l < lanes; ++l) if (m >> l & 1) {
          size_t i = (size_t)st * lanes + l;
          if (fz >> l & 1) ptrs[(size_t)fref[l] * MAXB + j] = 0;
          std::memcpy(recs + 4 * (i * MAXB + j), &frec[l], 16);
        }
      }
    }
  }
  return 0;
}

// Canonical state dump (u64 words), shared with the GPU's ha_debug_dump:
//   [nsbs]            slab counters
//   [2*nsbs]          descriptors: size | chunk<<32, chunk_sz | is_head<<32
//   [NSETS*nsw]       slab sets: free, roomy[16], busy[16], sparse[4]
//   [3*nheads*16]     heads, cached heads, tickets
//   [nsbs*bit_words]  bit arrays
//   [nsbs*nib_words]  allocation-size nibbles
//   [2]               head sets in use (the spread), classes that held a head (bits)
//   [LARGE_NCLASSES]  large-region tickets
//   [lwords]          large-region unit bits
//   [lunits]          large-region start words (class + 1 of the block starting there)
static u64 dump_words(const Heap& h) {
  return (u64)h.nsbs * 3 + (u64)Heap::NSETS * h.nsw + 3ull * h.nheads * NSIZES + (u64)h.nsbs * h.bit_words +
         (u64)h.nsbs * h.nib_words + 2 + LARGE_NCLASSES + h.lwords + h.lunits;
}
uint64_t hm_dump_words(void* hv) { return dump_words(*(Heap*)hv); }
int hm_dump(void* hv, uint64_t* out) {
  Heap* h = (Heap*)hv;
  u64* p = out;
  for (u32 b = 0; b < h->nsbs; ++b) *p++ = h->ctr[b].load();
  for (u32 b = 0; b < h->nsbs; ++b) {
    *p++ = (u64)h->d_size[b].load() | (u64)h->d_chunk[b].load() << 32;
    *p++ = (u64)h->d_csz[b].load() | (u64)h->d_head[b].load() << 32;
  }
  for (auto& x : h->sets) *p++ = x.load();
  for (auto& x : h->head) *p++ = x.load();
  for (auto& x : h->cached) *p++ = x.load();
  for (auto& x : h->ticket) *p++ = x.load();
  for (auto& x : h->bits) *p++ = x.load();
  for (auto& x : h->nibs) *p++ = x.load();
  *p++ = h->spread.load();
  *p++ = h->class_seen.load();
  for (auto& x : h->lticket) *p++ = x.load();
  for (auto& x : h->lbits) *p++ = x.load();
  for (auto& x : h->lstart) *p++ = x.load();
  return 0;
}

// Invariants over a canonical dump (model's or GPU's), at quiescence:
//   1 per-slab popcount(bits) == count, and sum of nibbles == count
//   2 every set nibble marks a block whose chunks are all set in the bit array
//   3 free slabs (chunk NONE): count 0, no bits, no nibbles, in the free set
//   4 every non-head slab with count > 0 is in a set (busy/roomy/sparse)
//   5 a slab's head references (counter and descriptor) equal the number of
//     head and cached-head words naming it
//   6 a free non-head slab is in the free set; any other non-head slab has a
//     nonzero count, is not in the free set, and is in a sparse / roomy /
//     busy set of its class whose level is at most the one its count names
//     (a wave that loaded a head word before the head was detached can still
//     add blocks to it; only frees move a slab between sets, so its count may
//     have risen past its set's range, never fallen).  Further entries in
//     class sets are stale hints, which pops check and drop
//   7 large region: every live block (a nonzero start word) is aligned to
//     its size, inside the region, its units' bits set and not overlapping
//     another live block; every set bit inside the region belongs to a live
//     block; the bits past the region are set
// returns the number of violations; detail[0..6] counts per invariant
int64_t hm_check_dump(const hm_opts_t* o, const uint64_t* d, int64_t* detail) {
  Heap h(to_opts(o));
  if (!h.ok) return -1;
  for (int i = 0; i < 7; ++i) detail[i] = 0;
  const u64* ctr = d;
  const u64* desc = ctr + h.nsbs;
  const u64* sets = desc + 2ull * h.nsbs;
  const u64* heads = sets + (u64)Heap::NSETS * h.nsw;
  const u64* cached = heads + (u64)h.nheads * NSIZES;
  const u64* bits = cached + 2ull * h.nheads * NSIZES;
  const u64* nibs = bits + (u64)h.nsbs * h.bit_words;
  auto in_set = [&](u32 set, u32 sb) { return (sets[(u64)set * h.nsw + sb / 64] >> (sb % 64)) & 1; };
  for (u32 b = 0; b < h.nsbs; ++b) {
    u32 count = ctr_count(ctr[b]);
    u32 size_id = (u32)desc[2 * b], chunk_id = (u32)(desc[2 * b] >> 32);
    bool is_head = (desc[2 * b + 1] >> 32) != 0;
    u64 pop = 0, nsum = 0;
    const u64* bw = bits + (u64)b * h.bit_words;
    const u64* nw = nibs + (u64)b * h.nib_words;
    for (u32 w = 0; w < h.bit_words; ++w) pop += (u64)__builtin_popcountll(bw[w]);
    for (u32 w = 0; w < h.nib_words; ++w) {
      u64 x = nw[w];
      while (x) {
        u32 q = (u32)__builtin_ctzll(x) / 4;
        u32 v = (u32)(x >> (4 * q)) & 0xF;
        nsum += v;
        u32 ich = w * 16 + q;
        u64 m = ((1ull << v) - 1) << (ich % 64);
        if (v > 8 || ich / 64 >= h.bit_words || (bw[ich / 64] & m) != m) detail[1]++;
        x &= ~(0xFull << (4 * q));
      }
    }
    if (pop != count || nsum != count) detail[0]++;
    if (ctr_chunk(ctr[b]) == CTR_CHUNK_NONE && !is_head) {
      if (count || pop || nsum || !in_set(0, b)) detail[2]++;
    } else if (!is_head && count > 0 && size_id < NSIZES && chunk_id < NCHUNK_IDS) {
      if (!in_set(h.set_busy(size_id), b) && !in_set(h.set_roomy(size_id), b) &&
          !in_set(h.set_sparse(chunk_id), b))
        detail[3]++;
    }
  }
  std::vector<u32> refs(h.nsbs, 0);
  for (u32 i = 0; i < h.nheads * NSIZES; ++i) {
    for (const u64* arr : {heads, cached}) {
      u32 sb = (u32)arr[i];
      if (sb == SB_NONE) continue;
      sb &= ~HEAD_EXT;
      if (sb >= h.nsbs) { detail[4]++; continue; }
      refs[sb]++;
    }
  }
  for (u32 b = 0; b < h.nsbs; ++b) {
    u32 dref = (u32)(desc[2 * b + 1] >> 32);
    if (ctr_head_refs(ctr[b]) != refs[b] || dref != refs[b]) detail[4]++;
  }
  // 6: set membership of non-head slabs by their count.  A free slab is in
  // the free set.  Any other slab holds chunks, is not in the free set (whose
  // pop checks only is_head), and is in a set of its own class at a level no
  // higher than its count's, where a search of that level or a lower one
  // finds it.  Extra entries in class sets are stale hints that every pop
  // drops after checking the slab (a transition that files a slab by a count
  // read before a concurrent last free marked it free leaves one:
  // src/slab.cuh:176-221 files without the slab lock), so they are allowed.
  for (u32 b = 0; b < h.nsbs; ++b) {
    if (refs[b] || (desc[2 * b + 1] >> 32)) continue;
    const u32 size_id = (u32)desc[2 * b], chunk_id = (u32)(desc[2 * b] >> 32);
    const u32 count = ctr_count(ctr[b]);
    bool ok;
    if (ctr_chunk(ctr[b]) == CTR_CHUNK_NONE) {
      ok = in_set(0, b);
    } else {
      ok = false;
      if (!in_set(0, b) && size_id < NSIZES && chunk_id < NCHUNK_IDS && count > 0) {
        const u32 cls = sb_class(count, h.si[size_id]);
        for (u32 lvl = CLS_SPARSE; lvl <= cls && !ok; ++lvl) ok = in_set(h.class_set(lvl, size_id, chunk_id), b);
      }
    }
    if (!ok) detail[5]++;
  }
  // 7: the large region
  const u64* lbits = nibs + (u64)h.nsbs * h.nib_words + 2 + LARGE_NCLASSES;
  const u64* lstart = lbits + h.lwords;
  std::vector<u8> owned((size_t)h.lwords * 64, 0);
  for (u32 u = 0; u < h.lunits; ++u) {
    if (!lstart[u]) continue;
    const u64 c = lstart[u] - 1, k = 1ull << c;
    if (c >= h.lnclasses || u % k || u + k > h.lunits) { detail[6]++; continue; }
    for (u64 x = u; x < u + k; ++x) {
      if (owned[x] || !((lbits[x / 64] >> (x % 64)) & 1)) detail[6]++;
      owned[x] = 1;
    }
  }
  for (u64 x = 0; x < (u64)h.lwords * 64; ++x) {
    const bool bit = (lbits[x / 64] >> (x % 64)) & 1;
    if (x >= h.lunits ? !bit : (bit && !owned[x])) detail[6]++;
  }
  int64_t tot = 0;
  for (int i = 0; i < 7; ++i) tot += detail[i];
  return tot;
}

// ------------------------------------------------------------- CPU baselines
// Plain 16 B workload on the host model: nlanes logical lanes split over
// nthreads host threads; each thread drives 64-lane waves: malloc(16), write a
// 16 B pattern, read it back, free; iters passes.  Returns pairs/s.
double hm_bench_plain(const hm_opts_t* o, int nthreads, int nlanes, int iters, uint64_t* violations) {
  Opts op = to_opts(o);
  Heap h(op, true);
  if (!h.ok) return -1.0;
  std::atomic<u64> bad{0};
  auto body = [&](int t) {
    int lo = (int)((int64_t)nlanes * t / nthreads), hi = (int)((int64_t)nlanes * (t + 1) / nthreads);
    u32 nb[WAVE];
    for (u32 l = 0; l < WAVE; ++l) nb[l] = 16;
    u64 out[WAVE];
    u64 local_bad = 0;
    for (int it = 0; it < iters; ++it) {
      for (int base = lo; base < hi; base += WAVE) {
        int n = std::min<int>(WAVE, hi - base);
        u64 act = n == 64 ? ~0ull : ((1ull << n) - 1);
        h.wave_malloc((u32)t % h.nheads, act, nb, out, nullptr);
        for (int l = 0; l < n; ++l) {
          if (!out[l]) { local_bad++; continue; }
          u32* p = (u32*)(h.arena.data() + (out[l] - 1));
          u32 tid = (u32)(base + l);
          p[0] = tid; p[1] = ~tid; p[2] = tid ^ 0x9E3779B9u; p[3] = (u32)it;
        }
        for (int l = 0; l < n; ++l) {
          if (!out[l]) continue;
          u32* p = (u32*)(h.arena.data() + (out[l] - 1));
          u32 tid = (u32)(base + l);
          if (p[0] != tid || p[1] != ~tid || p[2] != (tid ^ 0x9E3779B9u) || p[3] != (u32)it) local_bad++;
        }
        h.wave_free(act, out, nullptr);
      }
    }
    bad += local_bad;
  };
  auto t0 = std::chrono::steady_clock::now();
  std::vector<std::thread> th;
  for (int t = 0; t < nthreads; ++t) th.emplace_back(body, t);
  for (auto& x : th) x.join();
  auto t1 = std::chrono::steady_clock::now();
  if (violations) *violations = bad.load();
  double s = std::chrono::duration<double>(t1 - t0).count();
  return (double)nlanes * iters / s;
}

// glibc malloc(16)/free with the same lane split and pattern check.
double hm_bench_glibc(int nthreads, int nlanes, int iters, uint64_t* violations) {
  std::atomic<u64> bad{0};
  auto body = [&](int t) {
    int lo = (int)((int64_t)nlanes * t / nthreads), hi = (int)((int64_t)nlanes * (t + 1) / nthreads);
    u64 local_bad = 0;
    u32* ptrs[WAVE];
    for (int it = 0; it < iters; ++it) {
      for (int base = lo; base < hi; base += WAVE) {
        int n = std::min<int>(WAVE, hi - base);
        for (int l = 0; l < n; ++l) {
          u32* p = (u32*)std::malloc(16);
          ptrs[l] = p;
          if (!p) { local_bad++; continue; }
          u32 tid = (u32)(base + l);
          p[0] = tid; p[1] = ~tid; p[2] = tid ^ 0x9E3779B9u; p[3] = (u32)it;
        }
        for (int l = 0; l < n; ++l) {
          u32* p = ptrs[l];
          if (!p) continue;
          u32 tid = (u32)(base + l);
          if (p[0] != tid || p[1] != ~tid || p[2] != (tid ^ 0x9E3779B9u) || p[3] != (u32)it) local_bad++;
          std::free(p);
        }
      }
    }
    bad += local_bad;
  };
  auto t0 = std::chrono::steady_clock::now();
  std::vector<std::thread> th;
  for (int t = 0; t < nthreads; ++t) th.emplace_back(body, t);
  for (auto& x : th) x.join();
  auto t1 = std::chrono::steady_clock::now();
  if (violations) *violations = bad.load();
  return (double)nlanes * iters / std::chrono::duration<double>(t1 - t0).count();
}

}  // extern "C"
