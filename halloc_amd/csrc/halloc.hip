// halloc.hip -- MI355X (gfx950, wave64) device allocator: hamalloc / hafree and
// the host API ha_init / ha_shutdown / ha_extfrag.
//
// Built with -fgpu-rdc into libhalloc.a so that user device code links against
// hamalloc / hafree (the drop-in boundary of src/halloc.h:38-55).
//
// Scheme (canonizer/halloc): size classes of 2 or 3 units, slabs with one bit
// per 16 B chunk searched by a hash, per-class head slabs with a cached next
// head, and busy / roomy / sparse / free slab sets.  MI355X design:
//   * a wave processes its lanes one size class at a time; inside a class group
//     the head slab, the size info and the probe counter are wave-uniform;
//   * a class group takes ONE ticket and its lanes fill one aligned run of 64
//     blocks, so a probe is one atomic per 64-bit bit word touched, its union
//     computed from the wave-uniform lane bitmap with scalar bit arithmetic;
//   * head slabs and tickets exist per XCD (head = xcc_id % nheads), so the
//     hot counters split 8 ways (one address sustains ~88 M atomics/s);
//   * one contiguous arena: slab id = offset >> sb_sz_sh, no address grid;
//   * slab counters, hot words and tickets each own a 128 B line.
// All shared metadata is accessed with agent-scope atomics (coherent across
// XCDs).  Fences (which write back / invalidate the L2) appear only on the slow
// path, around the slab and head locks.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>

#include "../../include/halloc.h"
#include "../../include/halloc_amd.h"
#include "ha_internal.h"

namespace ha {

__constant__ Params ha_p;

#define HA_DEV __device__ __forceinline__

// ------------------------------------------------------------ wave helpers
HA_DEV u32 lane_id() { return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u)); }
HA_DEV u64 ballot(bool p) { return __ballot(p); }
HA_DEV u32 rdl(u32 v, u32 lane) { return (u32)__builtin_amdgcn_readlane((int)v, (int)lane); }
HA_DEV u64 rdl64(u64 v, u32 lane) {
  return ((u64)rdl((u32)(v >> 32), lane) << 32) | rdl((u32)v, lane);
}
HA_DEV u32 rfl(u32 v) { return (u32)__builtin_amdgcn_readfirstlane((int)v); }
// v with lane `lane` replaced by the wave-uniform x
HA_DEV u32 wrl(u32 x, u32 lane, u32 v) { return lane_id() == lane ? x : v; }
HA_DEV u64 shfl64(u64 v, u32 src) {
  int a = (int)(src << 2);
  u32 lo = (u32)__builtin_amdgcn_ds_bpermute(a, (int)(u32)v);
  u32 hi = (u32)__builtin_amdgcn_ds_bpermute(a, (int)(u32)(v >> 32));
  return ((u64)hi << 32) | lo;
}
HA_DEV u32 xcc_id() { return __builtin_amdgcn_s_getreg((20) | (0 << 6) | ((4 - 1) << 11)); }
// HW_ID bits 8..15: the wave's CU, shader array and shader engine
HA_DEV u32 cu_bits() { return __builtin_amdgcn_s_getreg((4) | (8 << 6) | ((8 - 1) << 11)); }
// The head set a wave would use with every set in use: its XCD's, and with
// more than 8 sets (a multiple of 8), one of the XCD's nheads / 8 sets picked
// by the wave's CU.
HA_DEV u32 head_set() {
  const u32 x = xcc_id(), n = ha_p.nheads;
  if (n <= 8) return mod_by(x, n, ha_p.nh_m);
  return x + 8u * mod_by(cu_bits(), n / 8u, ha_p.nh8_m);  // x < 8: already below n
}
// Set `full` (< nheads) folded onto the first s sets (s = nheads, or a power
// of two below it): the XCD part keeps its low bits, the CU part its low bits.
HA_DEV u32 fold_set(u32 full, u32 s) {
  if (s >= ha_p.nheads) return full;
  if (s <= 8) return full & (s - 1);
  return (full & 7u) + 8u * ((full >> 3) & (s / 8u - 1));
}
// The sets in use now, read through the scalar cache: a stale value (the
// spread only shrinks while a kernel runs) still names valid sets.
HA_DEV u32 spread_cached() {
  if (ha_p.fixed_sets) return ha_p.nheads;
  return *(const __attribute__((address_space(4))) u32*)(uintptr_t)ha_p.spread;
}

// ------------------------------------------------ agent-scope atomic helpers
#define HA_SCOPE __HIP_MEMORY_SCOPE_AGENT
HA_DEV u64 a_or(u64* p, u64 v) { return __hip_atomic_fetch_or(p, v, __ATOMIC_RELAXED, HA_SCOPE); }
HA_DEV u64 a_and(u64* p, u64 v) { return __hip_atomic_fetch_and(p, v, __ATOMIC_RELAXED, HA_SCOPE); }
HA_DEV u64 a_add(u64* p, u64 v) { return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, HA_SCOPE); }
HA_DEV u64 a_sub(u64* p, u64 v) { return __hip_atomic_fetch_sub(p, v, __ATOMIC_RELAXED, HA_SCOPE); }
HA_DEV u32 a_add32(u32* p, u32 v) { return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, HA_SCOPE); }
HA_DEV u32 a_xchg32(u32* p, u32 v) { return __hip_atomic_exchange(p, v, __ATOMIC_RELAXED, HA_SCOPE); }
HA_DEV u64 ld64(u64* p) { return __hip_atomic_load(p, __ATOMIC_RELAXED, HA_SCOPE); }
HA_DEV u32 ld32(u32* p) { return __hip_atomic_load(p, __ATOMIC_RELAXED, HA_SCOPE); }
HA_DEV void st32(u32* p, u32 v) { __hip_atomic_store(p, v, __ATOMIC_RELAXED, HA_SCOPE); }

HA_DEV u64* ctr_of(u32 sb) { return ha_p.ctr + (u64)sb * CTR_STRIDE; }
HA_DEV u32* hot(u32* base, u32 ihead, u32 sid) { return base + (ihead * NSIZES + sid) * HOT_STRIDE; }
HA_DEV u64* set_words(u32 set) { return ha_p.sets + (u64)set * ha_p.nsw; }

// ----------------------------------------------------- single-lane slow path
// Everything below until malloc_group runs on exactly one lane of a wave.

HA_DEV bool try_lock(u32* m) {
  bool ok = a_xchg32(m, 1u) == 0;
  if (ok) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  return ok;
}
HA_DEV void lock(u32* m) {
  while (!try_lock(m)) __builtin_amdgcn_s_sleep(2);
}
HA_DEV void unlock(u32* m) {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
  st32(m, 0u);
}

HA_DEV void set_add(u32 set, u32 sb) { a_or(set_words(set) + sb / 64, 1ull << (sb % 64)); }
HA_DEV void set_remove(u32 set, u32 sb) { a_and(set_words(set) + sb / 64, ~(1ull << (sb % 64))); }
// first-fit claim over every word holding a slab (src/sbset.cuh:5-35)
HA_DEV u32 set_pop(u32 set) {
  u64* w = set_words(set);
  for (u32 i = 0; i < ha_p.nsw; ++i) {
    u64 word = ld64(w + i);
    while (word) {
      u32 b = (u32)__builtin_ctzll(word);
      u64 m = 1ull << b;
      if (a_and(w + i, ~m) & m) return i * 64 + b;
      word &= ~m;
    }
  }
  return SB_NONE;
}

// sb_try_mark_free (src/slab.cuh:89-113); caller holds sb_lock[sb]
HA_DEV void try_mark_free(u32 sb, u32 size_id, u32 chunk_id, bool from_head) {
  Desc* d = ha_p.desc + sb;
  if (ld32(&d->is_head) || chunk_id == SZ_NONE) return;
  u64 delta = ctr_chunk_delta(chunk_id), old2;
  u64* c = ctr_of(sb);
  do {
    u64 old = a_add(c, delta);
    if (ctr_count(old) == 0) {
      st32(&d->size_id, SZ_NONE);
      st32(&d->chunk_id, SZ_NONE);
      st32(&d->chunk_sz, 0);
      set_add(0, sb);
      a_add32(ha_p.free_ep, 1u);
      set_remove(set_sparse(chunk_id), sb);
      if (size_id != SZ_NONE) {
        set_remove(set_roomy(size_id), sb);
        set_remove(set_busy(size_id), sb);
      }
      return;
    }
    old2 = a_sub(c, delta);
  } while (ctr_count(old2) == 0);
  if (from_head && size_id != SZ_NONE) {
    u32 cls = sb_class(ctr_count(old2), ha_p.si[size_id]);
    if (cls != CLS_FREE) set_add(class_set(cls, size_id, chunk_id), sb);
  }
}

// count transitions after a decrement (src/slab.cuh:176-221): `old` is the
// counter before the decrement of `change` chunks
HA_DEV void ctr_transitions(u32 sb, u64 old, u32 change) {
  if (ctr_is_head(old)) return;
  Desc* d = ha_p.desc + sb;
  u32 size_id = ld32(&d->size_id), chunk_id = ld32(&d->chunk_id);
  if (size_id == SZ_NONE || chunk_id == SZ_NONE) return;
  u32 oc = ctr_count(old), nc = oc - change;
  const SizeInfo& s = ha_p.si[size_id];
  u32 co = sb_class(oc, s), cn = sb_class(nc, s);
  if (co == cn) return;
  if (cn == CLS_FREE) {
    if (ctr_chunk(old) != CTR_CHUNK_NONE) {
      u32* l = ha_p.sb_lock + sb;
      lock(l);
      size_id = ld32(&d->size_id);
      chunk_id = ld32(&d->chunk_id);
      try_mark_free(sb, size_id, chunk_id, false);
      unlock(l);
    }
  } else {
    set_add(class_set(cn, size_id, chunk_id), sb);
    if (co != CLS_FREE) set_remove(class_set(co, size_id, chunk_id), sb);
  }
}


// drop one head reference (detach_head, src/slab.cuh:227-255); the last one
// files the slab by its count.  Caller holds sb_lock[sb].  True if the slab
// became free.
HA_DEV bool detach_head(u32 sb) {
  Desc* d = ha_p.desc + sb;
  const u32 refs = ld32(&d->is_head) - 1;
  st32(&d->is_head, refs);
  u64 old = a_sub(ctr_of(sb), CTR_HEAD);
  if (refs) return false;
  u32 count = ctr_count(old);
  u32 size_id = ld32(&d->size_id), chunk_id = ld32(&d->chunk_id);
  if (size_id == SZ_NONE) return false;
  u32 cls = sb_class(count, ha_p.si[size_id]);
  if (cls == CLS_FREE) {
    try_mark_free(sb, size_id, chunk_id, true);
    return ld32(&d->chunk_id) == SZ_NONE;
  }
  set_add(class_set(cls, size_id, chunk_id), sb);
  return false;
}

// Shares the head of another XCD's pair of the same class if it has room.
// Caller holds the head lock of (ihead, size_id).
HA_DEV u32 share_head(u32 ihead, u32 size_id, u32 chunk_id) {
  const SizeInfo& s = ha_p.si[size_id];
  for (u32 ih = 0; ih < ha_p.nheads; ++ih) {
    if (ih == ihead) continue;
    u32 w = ld32(hot(ha_p.head, ih, size_id));
    if (w == SB_NONE) continue;
    u32 sb = w & ~HEAD_EXT;
    if (ctr_count(ld64(ctr_of(sb))) >= head_limit(w, s)) continue;
    Desc* d = ha_p.desc + sb;
    u32* l = ha_p.sb_lock + sb;
    lock(l);
    u32 refs = ld32(&d->is_head);
    bool ok = refs > 0 && refs < 255 && ld32(&d->size_id) == size_id && ctr_chunk(ld64(ctr_of(sb))) == chunk_id;
    if (ok) {
      st32(&d->is_head, refs + 1);
      a_add(ctr_of(sb), CTR_HEAD);
    }
    unlock(l);
    if (ok) return w;
  }
  return SB_NONE;
}

// Takes over empty heads and cached heads of other (head, class) pairs until
// one slab becomes free; false if none does.  other_idle: only slabs another
// class holds without using them -- its cached heads, and its empty heads on
// pairs outside the spread, which no wave serves.  Caller holds the head lock
// of (ihead, size_id); the victims' head locks are only try-locked.  A class
// that never took a head holds none, so only the classes seen are visited
// (in the same set-major order): with one class the idle search costs
// nothing.
HA_DEV bool steal_idle_head(u32 ihead, u32 size_id, bool other_idle = false) {
  const u32 nh = ha_p.nheads;
  const u32 sp = ha_p.fixed_sets ? nh : ld32(ha_p.spread);
  u32 cls = ha_p.fixed_sets ? (1u << NSIZES) - 1 : ld32(ha_p.class_seen) & ((1u << NSIZES) - 1);
  if (other_idle) cls &= ~(1u << size_id);
  if (!cls) return false;
  for (u32 ih = 0; ih < nh; ++ih) {
    for (u32 rem = cls; rem; rem &= rem - 1) {
      const u32 c = (u32)__builtin_ctz(rem);
      if (ih == ihead && c == size_id) continue;
      u32* lp = hot(ha_p.head_lock, ih, c);
      for (int which = 0; which < 2; ++which) {
        if (other_idle && which == 1 && ih < sp) continue;
        u32* wp = hot(which == 0 ? ha_p.cached : ha_p.head, ih, c);
        u32 w = ld32(wp);
        if (w == SB_NONE) continue;
        u32 sb = w & ~HEAD_EXT;
        if (ctr_count(ld64(ctr_of(sb))) != 0) continue;
        if (!try_lock(lp)) continue;
        bool ok = ld32(wp) == w && ctr_count(ld64(ctr_of(sb))) == 0;
        bool freed = false;
        if (ok) {
          st32(wp, SB_NONE);
          if (which == 1) a_add32(ha_p.active, ~0u);  // the pair no longer holds a head
          u32* l = ha_p.sb_lock + sb;
          lock(l);
          freed = detach_head(sb);
          unlock(l);
        }
        unlock(lp);
        if (freed) return true;
      }
    }
  }
  return false;
}

// A head search frees at most this many idle heads before it gives up.
constexpr u32 MAX_STEALS = 4;

// With several head sets, free slabs are scarce below the largest of nheads,
// half the (head set, class) pairs that hold a head, and a quarter of the heap.
HA_DEV bool free_scarce() {
  if (ha_p.nheads <= 1) return false;
  u32 n = 0;
  u64* w = set_words(0);
  for (u32 i = 0; i < ha_p.nsw; ++i) n += (u32)__builtin_popcountll(ld64(w + i));
  u32 lim = ld32(ha_p.active) / 2;
  lim = lim > ha_p.nheads ? lim : ha_p.nheads;
  lim = lim > ha_p.nsbs / 4 ? lim : ha_p.nsbs / 4;
  return n < lim;
}

HA_DEV u32 adopt_head(u32 ihead, u32 size_id);

// find_sb_for_size (src/slab.cuh:259-372): roomy -> sparse -> free -> busy;
// for a head needed now, first (after sparse) a head of the class that a pair
// outside the spread holds, and at the end another XCD's head of the class,
// then an idle head of another class.  A cache refill (now == false) skips the free set
// when free slabs are scarce; a head needed now then first frees one slab
// that another class holds idle (steal_idle_head, other_idle) and, if none,
// tries sharing: pairs that share a head put two ticket sequences into one
// slab and crowd it.
HA_DEV u32 find_sb_for_size(u32 size_id, u32 chunk_id, u32 ihead, bool now) {
  const SizeInfo& s = ha_p.si[size_id];
  u32 sb;
  bool tried_share = false, tried_idle = false;
  u32 nsteal = 0;
again:
  while ((sb = set_pop(set_roomy(size_id))) != SB_NONE) {
    Desc* d = ha_p.desc + sb;
    u32* l = ha_p.sb_lock + sb;
    lock(l);
    bool found = false;
    if (!ld32(&d->is_head) && ld32(&d->size_id) == size_id) {
      if (ctr_count(ld64(ctr_of(sb))) <= s.roomy_thr) {
        found = true;
        st32(&d->is_head, 1);
        a_add(ctr_of(sb), CTR_HEAD);
      } else {
        set_add(set_busy(size_id), sb);
      }
    }
    unlock(l);
    if (found) return sb;
  }
  while ((sb = set_pop(set_sparse(chunk_id))) != SB_NONE) {
    Desc* d = ha_p.desc + sb;
    u32* l = ha_p.sb_lock + sb;
    lock(l);
    bool found = false;
    if (!ld32(&d->is_head) && ld32(&d->chunk_id) == chunk_id) {
      if (ctr_count(ld64(ctr_of(sb))) <= s.sparse_thr) {
        found = true;
        st32(&d->is_head, 1);
        st32(&d->size_id, size_id);
        a_add(ctr_of(sb), CTR_HEAD);
      } else {
        set_add(set_roomy(ld32(&d->size_id)), sb);
      }
    }
    unlock(l);
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
    Desc* d = ha_p.desc + sb;
    u32* l = ha_p.sb_lock + sb;
    lock(l);
    bool found = false;
    if (!ld32(&d->is_head)) {
      found = true;
      st32(&d->is_head, 1);
      st32(&d->size_id, size_id);
      st32(&d->chunk_id, chunk_id);
      st32(&d->chunk_sz, s.chunk_sz);
      a_add(ctr_of(sb), CTR_HEAD);
      a_sub(ctr_of(sb), ctr_chunk_delta(chunk_id));
    }
    unlock(l);
    if (found) return sb;
  }
  u64* w = set_words(set_busy(size_id));
  u32 limit = (s.busy_thr + s.nchunks) / 2;
  for (u32 i = 0; i < ha_p.nsw; ++i) {
    u64 word = ld64(w + i);
    while (word) {
      u32 b = (u32)__builtin_ctzll(word);
      u64 m = 1ull << b;
      word &= ~m;
      sb = i * 64 + b;
      if (ctr_count(ld64(ctr_of(sb))) >= limit) continue;  // a head is swapped out at limit
      if (!(a_and(w + i, ~m) & m)) continue;
      Desc* d = ha_p.desc + sb;
      u32* l = ha_p.sb_lock + sb;
      lock(l);
      bool found = false;
      if (!ld32(&d->is_head) && ld32(&d->size_id) == size_id) {
        found = true;
        st32(&d->is_head, 1);
        a_add(ctr_of(sb), CTR_HEAD);
      }
      unlock(l);
      if (found) return sb | HEAD_EXT;
    }
  }
  if (!now) return SB_NONE;
  if (!tried_share) {
    tried_share = true;
    u32 shared = share_head(ihead, size_id, chunk_id);
    if (shared != SB_NONE) return shared;
  }
  if (nsteal < MAX_STEALS && steal_idle_head(ihead, size_id)) {
    ++nsteal;
    goto again;
  }
  return SB_NONE;
}

// ------------------------------------------------------------- set spread
// With adaptive sets (fixed_sets 0), waves use the first `spread` head sets
// only.  The spread starts at nheads (which init keeps to what one size
// class allows) and shrinks as classes take their first head: to the largest
// power of two s (at most nheads) with 3 * s * classes <= slabs, so that the
// heads of every class seen stay within a third of the heap.

// spread for ncl classes (a power of two, at most nheads, at least 1)
HA_DEV u32 spread_for(u32 ncl) {
  const u32 per = ha_p.nsbs / (3u * (ncl ? ncl : 1u));
  if (per >= ha_p.nheads) return ha_p.nheads;
  u32 s = 1;
  while (2 * s <= per && 2 * s <= ha_p.nheads) s *= 2;
  return s >= ha_p.nheads ? ha_p.nheads : s;
}

// Size class c holds a head for the first time: count it and shrink the
// spread if the classes seen need it.  The pairs at and above the new spread
// keep their heads; their waves fold onto the pairs below from then on, and
// a pair below that needs a head of the class adopts one of them
// (adopt_head), while their empty heads and cached heads go to whichever
// search needs a slab (steal_idle_head).  No head is detached under waves
// that may still be reserving in it.
HA_DEV void note_class(u32 c) {
  if (ha_p.fixed_sets) return;
  const u32 bit = 1u << c;
  if (ld32(ha_p.class_seen) & bit) return;
  const u32 old = __hip_atomic_fetch_or(ha_p.class_seen, bit, __ATOMIC_RELAXED, HA_SCOPE);
  if (old & bit) return;
  const u32 want = spread_for((u32)__builtin_popcount(old | bit));
  u32 cur = ld32(ha_p.spread);
  while (want < cur) {
    u32 e = cur;
    if (__hip_atomic_compare_exchange_strong(ha_p.spread, &e, want, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                             HA_SCOPE))
      break;
    cur = e;
  }
}

// Takes over the head of class size_id held by a pair outside the spread:
// the word moves to the caller's pair, the slab stays a head with the same
// reference, so waves still reserving in it stay consistent.  Caller holds
// the head lock of (ihead, size_id); the other pair's lock is only tried.
HA_DEV u32 adopt_head(u32 ihead, u32 size_id) {
  if (ha_p.fixed_sets) return SB_NONE;
  const SizeInfo& s = ha_p.si[size_id];
  for (u32 ih = ld32(ha_p.spread); ih < ha_p.nheads; ++ih) {
    if (ih == ihead) continue;
    u32* hp = hot(ha_p.head, ih, size_id);
    u32 w = ld32(hp);
    if (w == SB_NONE || ctr_count(ld64(ctr_of(w & ~HEAD_EXT))) >= head_limit(w, s)) continue;
    u32* lp = hot(ha_p.head_lock, ih, size_id);
    if (!try_lock(lp)) continue;
    w = ld32(hp);
    const bool ok = w != SB_NONE;
    if (ok) {
      st32(hp, SB_NONE);
      a_add32(ha_p.active, ~0u);  // that pair no longer holds a head
    }
    unlock(lp);
    if (ok) return w;
  }
  return SB_NONE;
}

// new_sb_for_size (src/slab.cuh:378-445).  `seen` is the head word the caller
// last used.  A head replaced since then is returned as it is; otherwise the
// wave that takes the head lock replaces a missing or full head, and a wave
// that finds the lock taken waits for the head to change or the lock to be
// released.  A waiter whose wait saw a slab search fail gets SB_NONE too
// rather than repeating the search; otherwise it looks again.  force: the
// caller's probes ran out of tries in `seen`, or one probe sequence of its
// group failed CROWDED_NTRIES rounds there (a crowded head), and `seen` is
// then replaced even below its limit -- a deliberate difference: the reference replaces a head
// only when it is missing or at its busy threshold (src/slab.cuh:385-387).
// When such a forced search finds no other slab the head is kept, and the
// pair records the head and the free-slab epoch: a later forced replacement
// of the same head before any slab has become free keeps it without
// searching again.
HA_DEV u32 new_sb_for_size(u32 size_id, u32 ihead, u32 seen, bool force = false) {
  const SizeInfo& s = ha_p.si[size_id];
  if (!ha_p.fixed_sets) {
    // a wave with a stale spread names a pair outside the spread: serve it
    // from the pair its set folds onto.  `seen` stays the head the caller
    // last used: when that is the folded pair's head, it is replaced as for
    // any caller of that pair
    const u32 sp = ld32(ha_p.spread);
    if (ihead >= sp) ihead = fold_set(ihead, sp);
  }
  u32* hp = hot(ha_p.head, ihead, size_id);
  u32* lp = hot(ha_p.head_lock, ihead, size_id);
  u32* cp = hot(ha_p.cached, ihead, size_id);
  u32* ep = hot(ha_p.search_ep, ihead, size_id);
  u32* fp = hot(ha_p.forced, ihead, size_id);  // [0] head word, [1] free epoch
  for (;;) {
    u32 cur = ld32(hp);
    if (cur != seen && cur != SB_NONE) return cur;
    const u32 e0 = ld32(ep);
    if (try_lock(lp)) {
      cur = ld32(hp);
      u32 nh;
      // below: a forced replacement of a head that still has room
      const bool below = cur != SB_NONE && ctr_count(ld64(ctr_of(cur & ~HEAD_EXT))) < head_limit(cur, s);
      const u32 fe = ld32(ha_p.free_ep);
      if (below && ld32(fp) == cur && ld32(fp + 1) == fe) {
        nh = cur;  // this head's forced search failed, and no slab became free since
      } else if (cur == SB_NONE || !below || (force && cur == seen)) {
        nh = ld32(cp);
        st32(cp, SB_NONE);
        if (nh == SB_NONE) {
          // one head search per class at a time: a pair that follows sees the
          // head the one before it took and can share it instead of taking
          // another slab
          u32* cl = ha_p.class_lock + size_id * HOT_STRIDE;
          if (ha_p.nheads > 1) lock(cl);
          nh = find_sb_for_size(size_id, s.chunk_id, ihead, true);
          if (ha_p.nheads > 1) unlock(cl);
          st32(ep, (((ld32(ep) >> 1) + 1) << 1) | (nh == SB_NONE && !below ? 1u : 0u));
        }
        if (nh == SB_NONE && below) {
          nh = cur;  // no other slab: keep the head, which still has room
          st32(fp + 1, fe);
          st32(fp, cur);
        } else if (nh != SB_NONE) {
          __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
          st32(hp, nh);
          if (cur == SB_NONE) {
            a_add32(ha_p.active, 1u);
            note_class(size_id);
          }
          if (cur != SB_NONE) {
            u32 csb = cur & ~HEAD_EXT;
            u32* l = ha_p.sb_lock + csb;
            lock(l);
            detach_head(csb);
            unlock(l);
          }
          st32(cp, find_sb_for_size(size_id, s.chunk_id, ihead, false));
        }
      } else {
        nh = cur;
      }
      unlock(lp);
      return nh;
    }
    while (ld32(hp) == cur && ld32(lp) != 0) __builtin_amdgcn_s_sleep(2);
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    const u32 e1 = ld32(ep);
    if (e1 != e0 && (e1 & 1u) && ld32(hp) == cur) return SB_NONE;
  }
}

// x mod nchunks for a probe position: the offset stays far below 2^32 in
// practice (the division-free form); the 64-bit remainder covers the rest
HA_DEV u32 mod_chunks(const SizeInfo& s, u64 x) {
  return (x >> 32) ? (u32)(x % s.nchunks) : mod_by((u32)x, s.nchunks, s.nchunks_m);
}

// ------------------------------------------------------------- malloc path
// Lane l of a class group takes ticket cv = old + l.  With runs of 64 blocks
// (ichunk_init, src/halloc.cu:107-122, tmod = 64) the lanes fill the tail of
// run A = (old >> 6) from slot (old & 63) and the head of run B = A + 1,
// contiguously: lane l sits at chunk st + l*k of its segment, where st is the
// segment's (virtual) start.  Each probe is then one atomic per 64-bit word
// touched, its mask the union of the segment's lanes in the word.
//
// probe_run is one probe sequence (sb_alloc_in, src/slab.cuh:453-535) in slab
// hsb for the lanes in P (wave-uniform), from the blocks of ticket base `old`.
// It returns the lanes that reserved, records their block lengths in the
// nibbles, and leaves each reserved lane's chunk in my_ich.  itry advances for
// the whole group; off0 is the probe offset of the first attempt (a run that
// resumes after a first attempt made elsewhere starts at itry 1).
HA_DEV u64 probe_run(const SizeInfo& s, u32 hsb, u32 limit, u32 old, u64 P, u32 me, u32& itry, u32& my_ich,
                     u64 off0 = 0, bool* spent = nullptr) {
  const u32 k = s.ncib, nchunks = s.nchunks, nwb = nchunks >> 6, nwn = nchunks >> 4;
  const u32 slot0 = old & (RUN - 1), split = RUN - slot0;
  const u32 runA = mod_by(((old & ~(RUN - 1)) * THREAD_FREQ) * k, nchunks, s.nchunks_m);
  const u32 runB = mod_by((((old & ~(RUN - 1)) + RUN) * THREAD_FREQ) * k, nchunks, s.nchunks_m);
  const bool inA = me < split;
  const u64 amask = split >= 64 ? ~0ull : ((1ull << split) - 1);
  const u64 myseg = inA ? amask : ~amask;
  u64* bw = ha_p.bits + (u64)hsb * ha_p.bit_words;
  u64* nw = ha_p.nibs + (u64)hsb * ha_p.nib_words;
  u64 reserved = 0;
  u64 off = off0;
  for (;;) {
    const u32 pA = mod_chunks(s, (u64)runA + off), pB = mod_chunks(s, (u64)runB + off);
    const u32 st = inA ? pA + slot0 * k : pB - split * k;
    u32 wb, ib;
    u64 ub;
    run_word_bits(P & myseg, me, st, k, wb, ub, ib);
    const u32 wreal = wb >= nwb ? wb - nwb : wb;
    const bool inP = (P >> me) & 1;
    u64 oldw = 0;
    if (inP && ib == me) oldw = a_or(bw + wreal, ub);
    oldw = shfl64(oldw, ib & 63);
    u32 c = st + me * k;
    c = c >= nchunks ? c - nchunks : c;
    const u64 mym = ((1ull << k) - 1) << (c & 63);
    const bool ok = inP && !(oldw & mym);
    const u64 S = ballot(ok) & P;  // P: this lane's group (other groups probe alongside)
    const u64 F = P & ~S;
    if (F) {
      // give back the bits the failing lanes set (src/slab.cuh:480-483)
      u32 wf, iff;
      u64 uf;
      run_word_bits(F & myseg, me, st, k, wf, uf, iff);
      const u64 clr = uf & ~oldw;
      if (((F >> me) & 1) && iff == me && clr) a_and(bw + wreal, ~clr);
    }
    if (S) {
      // record the block lengths (sb_set_alloc_size, src/slab.cuh:60-70)
      u32 wn, inn;
      u64 un;
      run_word_nibs(S & myseg, me, st, k, (u64)k, wn, un, inn);
      const u32 wnreal = wn >= nwn ? wn - nwn : wn;
      if (ok && inn == me) a_or(nw + wnreal, un);
      if (ok) my_ich = c;
    }
    reserved |= S;
    P = F;
    if (!P) break;
    if (itry % CHECK_NTRIES == CHECK_NTRIES - 1) {
      if (ctr_count(ld64(ctr_of(hsb))) >= limit) break;
    }
    const u32 step = mod_by((itry + 1) * STEP_FREQ * k, s.hash_step, s.hash_step_m);
    off = (u64)(itry + 1) * step;
    if (++itry % MAX_NTRIES >= MAX_NTRIES - 1) {
      if (spent) *spent = true;
      break;
    }
  }
  return reserved;
}

// gives back one reserved block of this lane: nibble first, then the bits
HA_DEV void release_block(u32 hsb, u32 ich, u32 k) {
  a_and(ha_p.nibs + (u64)hsb * ha_p.nib_words + ich / 16, ~(0xFull << (4 * (ich % 16))));
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
  a_and(ha_p.bits + (u64)hsb * ha_p.bit_words + ich / 64, ~(((1ull << k) - 1) << (ich % 64)));
}

// One size class of one wave (hamalloc_small, src/halloc.cu:125-182).  The
// active lanes are exactly the bitmap g; sid, g and ihead are wave-uniform.
HA_DEV void* malloc_group(u32 sid, u64 g, u32 ihead, u32 me) {
  const SizeInfo& s = ha_p.si[sid];
  const u32 k = s.ncib;
  const u32 lead = (u32)__builtin_ctzll(g);
  const u32 last = 63u - (u32)__builtin_clzll(g);
  u32 hword = rfl(ld32(hot(ha_p.head, ihead, sid)));
  // tickets (size_ctr_inc, src/halloc.cu:71-98): one atomic per group
  u32 old = 0;
  if (me == lead) old = a_add32(hot(ha_p.ticket, ihead, sid), last + 1);
  old = rdl(old, lead);
  u32 itry = 0;
  u64 want = g;
  void* p = nullptr;
  u32 my_ich = 0;
  while (want) {
    bool need = false, spent = false, crowded = false;
    const u32 hsb = hword == SB_NONE ? SB_NONE : (hword & ~HEAD_EXT);
    const u32 limit = head_limit(hword, s);
    if (hsb == SB_NONE) {
      need = true;
    } else {
      const u32 itry0 = itry;
      const u64 reserved = probe_run(s, hsb, limit, old, want, me, itry, my_ich, 0, &spent);
      if (reserved != want) need = true;
      // a probe sequence this long says the head is crowded: the group's last
      // lanes wait for the luckiest of log(64) / log(1 / fill) rounds, so the
      // head is replaced even below its limit (CROWDED_NTRIES)
      if (itry - itry0 >= CROWDED_NTRIES) need = crowded = true;
      if (reserved) {
        // sb_ctr_inc (src/slab.cuh:122-155): one atomic for the group
        const u32 change = k * (u32)__builtin_popcountll(reserved);
        u64 oldc = 0;
        if (me == lead) oldc = a_add(ctr_of(hsb), (u64)change << CTR_COUNT_SH);
        oldc = rdl64(oldc, lead);
        const bool mine = (reserved >> me) & 1;
        if (ctr_chunk(oldc) != s.chunk_id) {
          // the slab was freed under us: give the chunks back
          if (mine) release_block(hsb, my_ich, k);
          if (me == lead) {
            u64 o = a_sub(ctr_of(hsb), (u64)change << CTR_COUNT_SH);
            ctr_transitions(hsb, o, change);
          }
          need = true;
        } else {
          const u32 oc = ctr_count(oldc);
          if (oc < limit && oc + change >= limit) need = true;
          if (mine) p = ha_p.arena + (u64)hsb * ha_p.sb_sz + chunk_mul(my_ich, s.chunk_sz);
          want &= ~reserved;
        }
      }
    }
    if (need) {
      u32 nh = 0;
      if (me == lead) nh = new_sb_for_size(sid, ihead, hword, (spent && want != 0) || crowded);
      hword = rdl(nh, lead);
      if (hword == SB_NONE) want = 0;
    }
  }
  return p;
}

// The first probe of one round of a batched call: the round's ticket base
// `old` fixes where each lane's slot starts (st), the bit word it falls in
// (wreal, unwrapped modulo the slab), the union of the group's fields in that
// word (ub) and the lane that issues the word's atomic (ib); myseg = the lanes
// of this lane's run.
struct BatchProbe {
  u32 st, ib, wreal;
  u64 ub, myseg;
};
HA_DEV BatchProbe batch_probe(const SizeInfo& s, u32 old, u32 me, u64 P) {
  const u32 k = s.ncib, nchunks = s.nchunks, nwb = nchunks >> 6;
  const u32 slot0 = old & (RUN - 1), split = RUN - slot0;
  const u32 runA = mod_by(((old & ~(RUN - 1)) * THREAD_FREQ) * k, nchunks, s.nchunks_m);
  const u32 runB = mod_by((((old & ~(RUN - 1)) + RUN) * THREAD_FREQ) * k, nchunks, s.nchunks_m);
  const bool inA = me < split;
  const u64 amask = split >= 64 ? ~0ull : ((1ull << split) - 1);
  BatchProbe b;
  b.myseg = inA ? amask : ~amask;
  b.st = inA ? runA + slot0 * k : runB - split * k;
  u32 wb;
  run_word_bits(P & b.myseg, me, b.st, k, wb, b.ub, b.ib);
  b.wreal = wb >= nwb ? wb - nwb : wb;
  return b;
}

// hamalloc_batch for one class group: lane l wants cnt (<= MAXB) blocks into
// out[0..cnt).  One ticket for all R rounds (round j based where R back-to-back
// calls would be), every round probed in the current head, ONE counter atomic
// for all reserved chunks; unserved lanes and rollbacks go through plain rounds.
HA_DEV void malloc_batch_group(u32 sid, u64 g, u32 ihead, u32 me, u32 cnt, void** out) {
  const SizeInfo& s = ha_p.si[sid];
  const u32 k = s.ncib;
  const u32 lead = (u32)__builtin_ctzll(g);
  const u32 L = 64u - (u32)__builtin_clzll(g);
  u32 R = 0;  // the group's largest cnt
  for (u32 j = 0; j < MAXB; ++j)
    if (ballot(cnt > j)) R = j + 1;
  // the ticket and the head load go out together: no head search touches a
  // ticket, so taking it before a missing head is replaced changes nothing
  u32 base = 0;
  if (me == lead) base = a_add32(hot(ha_p.ticket, ihead, sid), R * L);
  u32 hword = rfl(ld32(hot(ha_p.head, ihead, sid)));
  base = rdl(base, lead);
  if (hword == SB_NONE) {  // no head yet: get one first
    u32 nh = 0;
    if (me == lead) nh = new_sb_for_size(sid, ihead, SB_NONE);
    hword = rdl(nh, lead);
  }
  for (u32 j = 0; j < cnt; ++j) out[j] = nullptr;
  bool need = false;
  const u32 hsb = hword == SB_NONE ? SB_NONE : (hword & ~HEAD_EXT);
  const u32 limit = head_limit(hword, s);
  if (hsb == SB_NONE) {
    need = true;
  } else {
    u32 total = 0;
    char* sbase = ha_p.arena + (u64)hsb * ha_p.sb_sz;
    const u32 nchunks = s.nchunks, nwn = nchunks >> 4;
    u64* bw = ha_p.bits + (u64)hsb * ha_p.bit_words;
    u64* nw = ha_p.nibs + (u64)hsb * ha_p.nib_words;
    // the first probe of every round is issued before any result is awaited;
    // then, round by round, the lanes that failed give their new bits back and
    // go on from the second probe.  ow is indexed by the wave-uniform round, so
    // it stays in registers.
    u64 ow[MAXB];
#pragma nounroll
    for (u32 j = 0; j < R; ++j) {
      const u64 P = ballot(cnt > j);
      const BatchProbe b = batch_probe(s, base + j * L, me, P);
      u64 v = 0;
      if (((P >> me) & 1) && b.ib == me) v = a_or(bw + b.wreal, b.ub);
      ow[j] = v;
    }
#pragma nounroll
    for (u32 j = 0; j < R; ++j) {
      const u64 P = ballot(cnt > j);
      const BatchProbe b = batch_probe(s, base + j * L, me, P);
      const u64 oww = shfl64(ow[j], b.ib & 63);
      u32 c = b.st + me * k;
      c = c >= nchunks ? c - nchunks : c;
      const u64 mym = ((1ull << k) - 1) << (c & 63);
      const bool ok = ((P >> me) & 1) && !(oww & mym);
      const u64 S = ballot(ok);
      const u64 F = P & ~S;
      if (F) {
        u32 wf, iff;
        u64 uf;
        run_word_bits(F & b.myseg, me, b.st, k, wf, uf, iff);
        const u64 clr = uf & ~oww;
        if (((F >> me) & 1) && iff == me && clr) a_and(bw + b.wreal, ~clr);
      }
      if (S) {
        u32 wn, inn;
        u64 un;
        run_word_nibs(S & b.myseg, me, b.st, k, (u64)k, wn, un, inn);
        const u32 wnreal = wn >= nwn ? wn - nwn : wn;
        if (ok && inn == me) a_or(nw + wnreal, un);
        if (ok) out[j] = sbase + chunk_mul(c, s.chunk_sz);
      }
      u64 Sall = S;
      if (F) {
        // resume at the second probe (itry 1) for the lanes left over
        u32 itry = 1, my_ich = 0;
        const u64 off1 = (u64)mod_by(STEP_FREQ * k, s.hash_step, s.hash_step_m);
        const u64 S2 = probe_run(s, hsb, limit, base + j * L, F, me, itry, my_ich, off1);
        if ((S2 >> me) & 1) out[j] = sbase + chunk_mul(my_ich, s.chunk_sz);
        Sall |= S2;
      }
      if (Sall != P) need = true;
      total += (u32)__builtin_popcountll(Sall);
    }
    if (total) {
      const u32 change = k * total;
      u64 oldc = 0;
      if (me == lead) oldc = a_add(ctr_of(hsb), (u64)change << CTR_COUNT_SH);
      oldc = rdl64(oldc, lead);
      if (ctr_chunk(oldc) != s.chunk_id) {
        for (u32 j = 0; j < cnt; ++j) {
          if (out[j]) release_block(hsb, chunk_div((u32)((char*)out[j] - sbase), s.chunk_sz), k);
          out[j] = nullptr;
        }
        if (me == lead) {
          u64 o = a_sub(ctr_of(hsb), (u64)change << CTR_COUNT_SH);
          ctr_transitions(hsb, o, change);
        }
        need = true;
      } else {
        const u32 oc = ctr_count(oldc);
        if (oc < limit && oc + change >= limit) need = true;
      }
    }
  }
  if (need && me == lead) new_sb_for_size(sid, ihead, hword);
  for (u32 j = 0; j < R; ++j) {
    const bool pend = cnt > j && out[j] == nullptr;
    const u64 Pj = ballot(pend);
    if (pend) out[j] = malloc_group(sid, Pj, ihead, me);
  }
}

HA_DEV void* malloc_small(u32 nbytes, u32 ihead) {
  const u32 me = lane_id();
  const u32 sid = size_id_from_nbytes(nbytes);
  void* res = nullptr;
  u64 rem = ballot(true);
  while (rem) {
    u32 l0 = (u32)__builtin_ctzll(rem);
    u32 gsid = rdl(sid, l0);
    u64 g = ballot(sid == gsid) & rem;
    if ((g >> me) & 1) res = malloc_group(gsid, g, ihead, me);
    rem &= ~g;
  }
  return res;
}

// --------------------------------------------------------------- free path
// hafree_small (src/halloc.cu:211-230) for NR rounds at once: in round j the
// lanes with act[j] free the block at arena offset off[j].  The slab
// descriptors of every round are loaded together, then every round's
// nibbles are cleared, then its bits; the counter decrements and the set
// transitions they cause follow round by round, in the order NR one-round
// calls would make them.
template <int NR>
HA_DEV void free_rounds(const u64* off, const bool* act) {
  // Each XCD has its own write-back L2.  With a cached arena the caller's
  // stores to the blocks must reach memory before another XCD can be handed
  // the blocks, or a later eviction of this XCD's dirty lines would overwrite
  // the new owner's data: one agent-scope release per wave call.  The default
  // uncached arena needs none.
  if (ha_p.cached_arena) __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
  const u32 me = lane_id();
  u32 sb[NR], cs[NR], sid[NR], cid[NR];
#pragma unroll
  for (int j = 0; j < NR; ++j) {
    sb[j] = act[j] ? (u32)(off[j] >> ha_p.sb_sz_sh) : 0u;
    cs[j] = sid[j] = cid[j] = 0;
    if (act[j]) {
      Desc* d = ha_p.desc + sb[j];
      cs[j] = ld32(&d->chunk_sz);
      sid[j] = ld32(&d->size_id);
      cid[j] = ld32(&d->chunk_id);
    }
  }
  // nibbles: read back the block lengths and clear them
  u32 ich[NR], nch[NR], k[NR], st[NR];
  u64 A[NR], myseg[NR], o[NR];
  bool fast[NR];
  u32 inn[NR];
#pragma unroll
  for (int j = 0; j < NR; ++j) {
    const bool valid = act[j] && cs[j] != 0;
    ich[j] = valid ? chunk_div((u32)(off[j] & (ha_p.sb_sz - 1)), cs[j]) : 0u;
    A[j] = ballot(valid);
    nch[j] = 0;
    k[j] = st[j] = 0;
    myseg[j] = o[j] = 0;
    inn[j] = 64;
    fast[j] = false;
    // fast path: the lanes free one slab's blocks laid out as at most two runs
    // (what one class group of a wave allocates): one atomic per word touched
    if (A[j]) {
      const u32 l0 = (u32)__builtin_ctzll(A[j]);
      const u32 sb0 = rdl(sb[j], l0);
      if (ballot(valid && sb[j] == sb0) == A[j]) {
        const u32 sid0 = rdl(sid[j], l0);
        if (sid0 < NSIZES) {
          k[j] = ha_p.si[sid0].ncib;
          const u32 sa = rdl(ich[j], l0) - l0 * k[j];
          const bool ina = valid && ich[j] == sa + me * k[j];
          const u64 Am = ballot(ina);
          const u64 Rm = A[j] & ~Am;
          u32 sbb = 0;
          u64 Bm = 0;
          if (Rm) {
            const u32 l1 = (u32)__builtin_ctzll(Rm);
            sbb = rdl(ich[j], l1) - l1 * k[j];
            Bm = ballot(valid && !ina && ich[j] == sbb + me * k[j]);
          }
          // the run helpers take slot starts on multiples of k, as every block
          // of this size has; a block of another size of the chunk (a slab
          // taken over from the sparse set) may start anywhere
          fast[j] = (Am | Bm) == A[j] && ((sa | sbb) & (k[j] - 1)) == 0;
          st[j] = ina ? sa : sbb;
          myseg[j] = ina ? Am : Bm;
        }
      }
    }
    u64* nw = ha_p.nibs + (u64)sb[j] * ha_p.nib_words;
    if (fast[j]) {
      u32 wn;
      u64 un;
      run_word_nibs(myseg[j], me, st[j], k[j], 0xFull, wn, un, inn[j]);
      if (valid && inn[j] == me) o[j] = a_and(nw + wn, ~un);
    } else if (valid) {
      o[j] = a_and(nw + ich[j] / 16, ~(0xFull << (4 * (ich[j] % 16))));
    }
  }
  // bits
#pragma unroll
  for (int j = 0; j < NR; ++j) {
    const bool valid = act[j] && cs[j] != 0;
    u64* bw = ha_p.bits + (u64)sb[j] * ha_p.bit_words;
    if (fast[j]) {
      const u64 oldn = shfl64(o[j], inn[j] & 63);
      nch[j] = valid ? (u32)((oldn >> (4 * (ich[j] & 15))) & 0xF) : 0u;
      if (ballot(valid && nch[j] == k[j]) == A[j]) {
        u32 wb, ib;
        u64 ub;
        run_word_bits(myseg[j], me, st[j], k[j], wb, ub, ib);
        if (valid && ib == me) a_and(bw + wb, ~ub);
      } else if (nch[j]) {
        a_and(bw + ich[j] / 64, ~(((1ull << nch[j]) - 1) << (ich[j] % 64)));
      }
    } else if (valid) {
      nch[j] = (u32)((o[j] >> (4 * (ich[j] % 16))) & 0xF);
      if (nch[j]) a_and(bw + ich[j] / 64, ~(((1ull << nch[j]) - 1) << (ich[j] % 64)));
    }
  }
  // sb_ctr_dec (src/slab.cuh:161-224): one atomic per slab touched, issued for
  // every round in round order; then the transitions, round by round, each
  // from its own round's old count (as NR one-round calls would see them)
  u64 oldc[NR];
  u32 change[NR];
  bool lead[NR];
#pragma unroll
  for (int j = 0; j < NR; ++j) {
    u64 rem = ballot(nch[j] != 0);
    u32 my_change = 0;
    bool leader = false;
    while (rem) {
      const u32 l = (u32)__builtin_ctzll(rem);
      const u32 b = rdl(sb[j], l);
      const u64 G = ballot(sb[j] == b) & rem;
      u32 ch = 0;
      u64 Gr = G;
      while (Gr) {
        const u32 l2 = (u32)__builtin_ctzll(Gr);
        const u32 v = rdl(nch[j], l2);
        const u64 V = ballot(nch[j] == v) & Gr;
        ch += v * (u32)__builtin_popcountll(V);
        Gr &= ~V;
      }
      if (me == l) {
        leader = true;
        my_change = ch;
      }
      rem &= ~G;
    }
    oldc[j] = 0;
    if (leader) oldc[j] = a_sub(ctr_of(sb[j]), (u64)my_change << CTR_COUNT_SH);
    change[j] = my_change;
    lead[j] = leader;
  }
#pragma unroll
  for (int j = 0; j < NR; ++j) {
    const bool tr = lead[j] && !ctr_is_head(oldc[j]) && sid[j] < NSIZES && cid[j] != SZ_NONE &&
                    sb_class(ctr_count(oldc[j]), ha_p.si[sid[j]]) !=
                        sb_class(ctr_count(oldc[j]) - change[j], ha_p.si[sid[j]]);
    u64 T = ballot(tr);
    while (T) {
      const u32 l = (u32)__builtin_ctzll(T);
      if (me == l) ctr_transitions(sb[j], oldc[j], change[j]);
      T &= T - 1;
    }
  }
}

// Consecutive run rounds of hafree_batch (the rounds in the wave-uniform mask
// F).  A run round frees blocks of one slab laid out as at most two runs of
// consecutive k-chunk slots, one block per lane -- what one class group of
// one wave allocates in one round.  Lane j holds round j's slab (sb), block
// length (k) and run starts (sa, sbb); bit j of act/ina says whether this
// lane frees a block in round j and whether it lies in the first run.
// Every round's nibbles are cleared before any result is awaited, then every
// round's bits; then ONE counter decrement per slab for the total of its
// rounds, each round's old count recovered in round order (what one-round
// calls in round order would see), and the transitions round by round.
struct RunRounds {
  u32 sb, k, sa, sbb, sid, cid, act, ina;
};

// round j's wave-uniform values and this lane's place in it
struct RunRound {
  u32 sb, k, st, ich;
  bool act, ina;
  u64 Am, Bm;
};
HA_DEV RunRound run_round(const RunRounds& r, u32 j) {
  RunRound x;
  x.sb = rdl(r.sb, j);
  x.k = rdl(r.k, j);
  x.act = (r.act >> j) & 1;
  x.ina = (r.ina >> j) & 1;
  x.Am = ballot(x.act && x.ina);
  x.Bm = ballot(x.act && !x.ina);
  x.st = x.ina ? rdl(r.sa, j) : rdl(r.sbb, j);
  x.ich = x.st + lane_id() * x.k;
  return x;
}

HA_DEV void free_runs(u32 F, RunRounds r) {
  const u32 me = lane_id();
  // every round's nibbles, no result awaited (o is indexed by the
  // wave-uniform round: it stays in registers)
  u64 o[MAXB];
#pragma nounroll
  for (u32 j = 0; j < MAXB; ++j) {
    u64 v = 0;
    if ((F >> j) & 1) {
      const RunRound x = run_round(r, j);
      u32 wn, inn;
      u64 un;
      run_word_nibs(x.ina ? x.Am : x.Bm, me, x.st, x.k, 0xFull, wn, un, inn);
      if (x.act && inn == me) v = a_and(ha_p.nibs + (u64)x.sb * ha_p.nib_words + wn, ~un);
    }
    o[j] = v;
  }
  // then every round's bits and change
  u32 rch = 0;  // lane j: chunks freed in round j
#pragma nounroll
  for (u32 j = 0; j < MAXB; ++j) {
    if (!((F >> j) & 1)) continue;
    const RunRound x = run_round(r, j);
    const u64 myseg = x.ina ? x.Am : x.Bm;
    u32 wn, inn;
    u64 un;
    run_word_nibs(myseg, me, x.st, x.k, 0xFull, wn, un, inn);
    const u64 oj = o[j];
    // each issuer checks its own word: every block it cleared was k chunks
    const bool bad = x.act && inn == me && (oj & un) != (un / 0xFull) * x.k;
    u64* bw = ha_p.bits + (u64)x.sb * ha_p.bit_words;
    u32 change;
    if (!ballot(bad)) {
      u32 wb, ib;
      u64 ub;
      run_word_bits(myseg, me, x.st, x.k, wb, ub, ib);
      if (x.act && ib == me) a_and(bw + wb, ~ub);
      change = x.k * (u32)__builtin_popcountll(x.Am | x.Bm);
    } else {
      // lengths other than the class's (a block freed twice, a pointer into
      // a block): each lane clears what its own nibble says
      const u64 oldn = shfl64(oj, inn & 63);
      const u32 nch = x.act ? (u32)((oldn >> (4 * (x.ich & 15))) & 0xF) : 0u;
      if (nch) a_and(bw + x.ich / 64, ~(((1ull << nch) - 1) << (x.ich % 64)));
      change = 0;
#pragma unroll
      for (u32 b = 0; b < 4; ++b) change += (u32)__builtin_popcountll(ballot((nch >> b) & 1)) << b;
    }
    rch = wrl(change, j, rch);
  }
  const u32 rsb = r.sb, rsid = r.sid, rcid = r.cid;
  // one decrement per slab, by the lane of its first round; pre = the
  // changes of the slab's earlier rounds
  const bool inF = me < MAXB && ((F >> me) & 1) && rch != 0;
  u32 pre = 0, tot = 0, ldr = 0;
  for (u64 rem = ballot(inF); rem;) {
    const u32 l = (u32)__builtin_ctzll(rem);
    const u32 b = rdl(rsb, l);
    const u64 G = ballot(inF && rsb == b) & rem;
    u32 acc = 0;
    for (u64 Gi = G; Gi; Gi &= Gi - 1) {
      const u32 i = (u32)__builtin_ctzll(Gi);
      if (me == i) {
        pre = acc;
        ldr = l;
      }
      acc += rdl(rch, i);
    }
    if (me == l) tot = acc;
    rem &= ~G;
  }
  u64 told = 0;
  if (inF && me == ldr) told = a_sub(ctr_of(rsb), (u64)tot << CTR_COUNT_SH);
  const u64 oldc = shfl64(told, ldr) - ((u64)pre << CTR_COUNT_SH);
  const bool tr = inF && !ctr_is_head(oldc) && rsid < NSIZES && rcid != SZ_NONE &&
                  sb_class(ctr_count(oldc), ha_p.si[rsid < NSIZES ? rsid : 0]) !=
                      sb_class(ctr_count(oldc) - rch, ha_p.si[rsid < NSIZES ? rsid : 0]);
  for (u64 T = ballot(tr); T; T &= T - 1) {
    const u32 l = (u32)__builtin_ctzll(T);
    if (me == l) ctr_transitions(rsb, oldc, rch);
  }
}

HA_DEV void free_small(u64 off) {
  const bool act = true;
  free_rounds<1>(&off, &act);
}

// ------------------------------------------------------------ large blocks
// Requests above MAX_BLOCK_SZ come from the large region, the share of the
// pool outside the slabs (src/halloc.cu:285-289, which the reference hands to
// the CUDA device heap).  Class c is a block of 2^c units of 4 KiB aligned to
// its size.  One bit per unit marks it used, so the units of a freed block
// serve any class again: blocks split and merge without lists.  lstart[u]
// holds c + 1 while a block of class c starts at unit u; a free claims it
// with one exchange, so a second free, an interior pointer or a unit never
// handed out reads 0 and changes nothing.
HA_DEV u32 large_class(u64 nbytes) {
  const u64 units = (nbytes >> LARGE_UNIT_SH) + ((nbytes & ((1ull << LARGE_UNIT_SH) - 1)) != 0);
  return units <= 1 ? 0u : 64u - (u32)__builtin_clzll(units - 1);
}

HA_DEV u64 a_cas(u64* p, u64 expected, u64 desired) {
  __hip_atomic_compare_exchange_strong(p, &expected, desired, __ATOMIC_RELAXED, __ATOMIC_RELAXED, HA_SCOPE);
  return expected;  // the value found: `expected` as passed when the exchange took place
}

constexpr u64 UNIT_NONE = ~0ull;

// One class group of large requests of one wave: the lanes g (wave-uniform)
// want a block of class c.  The group takes one ticket (one atomic).  Blocks
// of up to 64 units: the group starts at word large_start(ticket) and walks
// the words in order; in each word the pending lanes take its free fields
// lowest first (the lane of rank r the r-th field), reserved by ONE atomic OR
// for the group, and the fields another wave took meanwhile are given back by
// one AND.  The group stays on a word while the value its last atomic
// returned shows a free field (every field it loses is one another lane
// won), so in a kernel that only allocates, NULL means every word was full
// when the group looked at it.  Larger blocks take an
// aligned run of whole words, one lane at a time in lane order, from the run
// large_start(ticket + rank) on, each word claimed by a compare-and-swap from
// zero.  Returns this lane's unit (UNIT_NONE: nothing free was found).
HA_DEV u64 large_group(u32 c, u64 g, u32 me) {
  const u32 lead = (u32)__builtin_ctzll(g);
  const u32 nw = ha_p.lwords;
  u64* lb = ha_p.lbits;
  u32 old = 0;
  if (me == lead) old = a_add32(ha_p.lticket + c * HOT_STRIDE, (u32)__builtin_popcountll(g));
  old = rdl(old, lead);
  u64 unit = UNIT_NONE;
  if (c <= 6) {
    const u32 k = 1u << c;
    u32 w = large_start(old, c, nw), moved = 0;
    u64 v = rdl64(ld64(lb + w), lead);
    u64 pend = g;
    while (pend) {
      const u64 cand = free_fields(v, k);
      const u32 n = (u32)__builtin_popcountll(cand);
      if (n == 0) {
        if (++moved >= nw) break;  // every word looked at: nothing free
        w = w + 1 == nw ? 0u : w + 1;
        v = rdl64(ld64(lb + w), lead);
        continue;
      }
      const u64 pick = lowest_bits(cand, (u32)__builtin_popcountll(pend));
      const u64 U = widen_fields(pick, k);
      const u32 r = (u32)__builtin_popcountll(pend & ((1ull << me) - 1));
      const bool mine = ((pend >> me) & 1) && r < n;
      const u32 pos = mine ? select_bit(pick, r) : 0u;
      u64 o = 0;
      if (me == lead) o = a_or(lb + w, U);
      o = rdl64(o, lead);
      const u64 won = pick & ~used_fields(o, k);
      const u64 back = U & ~o & ~widen_fields(won, k);  // bits set in fields found taken
      if (me == lead && back) a_and(lb + w, ~back);
      const bool got = mine && ((won >> pos) & 1);
      if (got) unit = (u64)w * 64 + pos;
      pend &= ~ballot(got);
      v = o | widen_fields(won, k);
    }
  } else {
    const u32 m = 1u << (c - 6);  // words per block
    const u32 ngr = nw / m;        // aligned runs of m words
    for (u64 G = g; G; G &= G - 1) {
      const u32 l = (u32)__builtin_ctzll(G);
      if (me != l || ngr == 0) continue;
      const u32 r = (u32)__builtin_popcountll(g & ((1ull << l) - 1));
      const u32 g0 = large_start(old + r, c, ngr);
      for (u32 i = 0; i < ngr && unit == UNIT_NONE; ++i) {
        const u32 grp = g0 + i >= ngr ? g0 + i - ngr : g0 + i;
        u64* base = lb + (u64)grp * m;
        u32 j = 0;
        while (j < m && ld64(base + j) == 0) ++j;
        if (j < m) continue;
        for (j = 0; j < m; ++j)
          if (a_cas(base + j, 0ull, ~0ull) != 0ull) break;
        if (j == m) {
          unit = (u64)grp * m * 64;
        } else {
          while (j--) a_and(base + j, 0ull);  // give back the words taken
        }
      }
    }
  }
  if (unit != UNIT_NONE) st32(ha_p.lstart + unit, c + 1);
  return unit;
}

// hamalloc for the active lanes whose requests exceed MAX_BLOCK_SZ: one
// class group at a time, groups ordered by their lowest lane
HA_DEV void* large_alloc(u64 nbytes) {
  const u32 me = lane_id();
  const u32 c = large_class(nbytes);
  const bool fits = c < ha_p.lnclasses;
  void* res = nullptr;
  u64 rem = ballot(fits);
  while (rem) {
    const u32 l0 = (u32)__builtin_ctzll(rem);
    const u32 gc = rdl(c, l0);
    const u64 g = ballot(fits && c == gc) & rem;
    if ((g >> me) & 1) {
      const u64 u = large_group(gc, g, me);
      if (u != UNIT_NONE) res = ha_p.large + (u << LARGE_UNIT_SH);
    }
    rem &= ~g;
  }
  return res;
}

HA_DEV void large_free(u64 off) {
  if (off & ((1ull << LARGE_UNIT_SH) - 1)) return;  // not a unit start: not a block
  const u64 unit = off >> LARGE_UNIT_SH;
  const u32 s = a_xchg32(ha_p.lstart + unit, 0u);
  if (s == 0) return;  // freed already, inside a block, or never handed out
  // with a cached arena the caller's stores to the block reach memory before
  // its units can be handed out again (see free_rounds)
  if (ha_p.cached_arena) __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
  const u32 c = s - 1;
  if (c <= 6) {
    a_and(ha_p.lbits + unit / 64, ~(low_mask(1u << c) << (unit % 64)));
  } else {
    for (u64 j = 0; j < (1ull << (c - 6)); ++j) a_and(ha_p.lbits + unit / 64 + j, 0ull);
  }
}

}  // namespace ha

// ---------------------------------------------------------- public device API
// The entry points are inlined into the calling kernel (link-time inlining
// under -fgpu-rdc): a call would make every kernel keep its live values in
// callee-saved registers and pay the callee's register saves per call; inlined,
// the 16 B malloc/free kernel is about 5% faster.  Built with
// -DHALLOC_OUT_OF_LINE (libhalloc_ool.a) they are out-of-line calls, as the
// reference's __noinline__ hamalloc / hafree are (src/halloc.h:38,41): one
// copy of the allocator however many call sites a program has.
#ifdef HALLOC_OUT_OF_LINE
#define HA_API __device__ __attribute__((noinline))
#else
#define HA_API __device__ __attribute__((always_inline))
#endif

HA_API void* hamalloc(size_t nbytes) {
  // requests above 3072 B go to the large-block region (src/halloc.cu:185-195
  // sends them to the CUDA device heap)
  if (nbytes > ha::MAX_BLOCK_SZ) return ha::large_alloc(nbytes);
  return ha::malloc_small((ha::u32)nbytes, ha::rfl(ha::fold_set(ha::head_set(), ha::spread_cached())));
}

// Batched allocation: lane l gets cnt (<= HA_MAX_BATCH) blocks of nbytes into
// out[0..cnt) (NULL where memory ran out); returns the number obtained.
namespace ha {
// hamalloc_batch on head set ihead (already folded to the spread)
HA_DEV int malloc_batch_on(size_t nbytes, int cnt, void** out, u32 ihead) {
  if (cnt <= 0) return 0;
  ha::u32 c = cnt > (int)ha::MAXB ? ha::MAXB : (ha::u32)cnt;
  int got = 0;
  if (nbytes > ha::MAX_BLOCK_SZ) {
    for (ha::u32 j = 0; j < c; ++j) got += (out[j] = ha::large_alloc(nbytes)) != nullptr;
    return got;
  }
  const ha::u32 me = ha::lane_id();
  const ha::u32 sid = ha::size_id_from_nbytes((ha::u32)nbytes);
  ha::u64 rem = ha::ballot(true);
  while (rem) {
    ha::u32 l0 = (ha::u32)__builtin_ctzll(rem);
    ha::u32 gsid = ha::rdl(sid, l0);
    ha::u64 g = ha::ballot(sid == gsid) & rem;
    if ((g >> me) & 1) ha::malloc_batch_group(gsid, g, ihead, me, c, out);
    rem &= ~g;
  }
  for (ha::u32 j = 0; j < c; ++j) got += out[j] != nullptr;
  return got;
}
}  // namespace ha

HA_API int hamalloc_batch(size_t nbytes, int cnt, void** out) {
  return ha::malloc_batch_on(nbytes, cnt, out, ha::rfl(ha::fold_set(ha::head_set(), ha::spread_cached())));
}

// Batched free: lane l frees ptrs[0..cnt) (NULL entries are skipped); round
// j frees every lane's ptrs[j], in round order.  Rounds that free one slab's
// runs (what a batched allocation hands out) go together through free_runs;
// any other round goes through the one-round free path in its place in the
// order.  Large blocks and foreign pointers go first, one by one.
HA_API void hafree_batch(void* const* ptrs, int cnt) {
  using namespace ha;
  if (cnt <= 0) return;
  const u32 c = cnt > (int)MAXB ? MAXB : (u32)cnt;
  u32 R = 0;  // the wave's largest count
  for (u32 j = 0; j < MAXB; ++j)
    if (ballot(c > j)) R = j + 1;
  for (u32 j = 0; j < c; ++j) {
    void* p = ptrs[j];
    if (p && (u64)((char*)p - ha_p.arena) >= ha_p.arena_bytes) hafree(p);
  }
  // see free_rounds: one release per call for a cached arena
  if (ha_p.cached_arena) __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
  const u32 me = lane_id();
  // every round's slab, where all its blocks lie in one (lane j: round j's)
  u32 ino[MAXB];  // in-slab offsets
  u32 actm = 0, single = 0, rsb = 0;
#pragma nounroll
  for (u32 j = 0; j < MAXB; ++j) {
    ino[j] = 0;
    if (j >= R) continue;
    void* p = j < c ? ptrs[j] : nullptr;
    const u64 off = (u64)((char*)p - ha_p.arena);
    const bool act = p && off < ha_p.arena_bytes;
    const u32 sb = act ? (u32)(off >> ha_p.sb_sz_sh) : 0u;
    ino[j] = (u32)(off & (ha_p.sb_sz - 1));
    actm |= (u32)act << j;
    const u64 A = ballot(act);
    if (!A) continue;
    const u32 sb0 = rdl(sb, (u32)__builtin_ctzll(A));
    if (ballot(act && sb == sb0) == A) {
      single |= 1u << j;
      rsb = wrl(sb0, j, rsb);
    }
  }
  // their descriptors: one load per round
  u32 rcs = 0, rsid = SZ_NONE, rcid = SZ_NONE;
  if (me < MAXB && ((single >> me) & 1)) {
    Desc* d = ha_p.desc + rsb;
    rcs = ld32(&d->chunk_sz);
    rsid = ld32(&d->size_id);
    rcid = ld32(&d->chunk_id);
  }
  // which of them are run rounds
  u32 fast = 0, inam = 0, rk = 0, rsa = 0, rsbb = 0;
#pragma nounroll
  for (u32 j = 0; j < MAXB; ++j) {
    if (!((single >> j) & 1)) continue;
    const u32 cs = rdl(rcs, j), sid = rdl(rsid, j);
    if (cs == 0 || sid >= NSIZES) continue;
    const bool act = (actm >> j) & 1;
    const u32 ich = act ? chunk_div(ino[j], cs) : 0u;
    const u32 k = ha_p.si[sid].ncib;
    const u64 A = ballot(act);
    const u32 l0 = (u32)__builtin_ctzll(A);
    const u32 sa = rdl(ich, l0) - l0 * k;
    const bool ina = act && ich == sa + me * k;
    const u64 Am = ballot(ina), Rm = A & ~Am;
    u32 sbb = 0;
    u64 Bm = 0;
    if (Rm) {
      const u32 l1 = (u32)__builtin_ctzll(Rm);
      sbb = rdl(ich, l1) - l1 * k;
      Bm = ballot(act && !ina && ich == sbb + me * k);
    }
    // run starts on multiples of k only (see free_rounds)
    if ((Am | Bm) == A && ((sa | sbb) & (k - 1)) == 0) {
      fast |= 1u << j;
      inam |= (u32)ina << j;
      rk = wrl(k, j, rk);
      rsa = wrl(sa, j, rsa);
      rsbb = wrl(sbb, j, rsbb);
    }
  }
  // round order: runs of run rounds together, the others one by one
  u32 F = 0;
  for (u32 j = 0; j <= R; ++j) {
    if (j < R && ((fast >> j) & 1)) {
      F |= 1u << j;
      continue;
    }
    if (F) free_runs(F, RunRounds{rsb, rk, rsa, rsbb, rsid, rcid, actm, inam});
    F = 0;
    if (j < R && ballot((actm >> j) & 1)) {
      const bool act = (actm >> j) & 1;
      u64 off = act ? (u64)((char*)ptrs[j] - ha_p.arena) : 0ull;
      free_rounds<1>(&off, &act);
    }
  }
}

// hamalloc and hamalloc_batch with the full-spread head set chosen by the
// caller instead of by the XCD id, folded onto the spread read coherently;
// the replay harness uses them to pin behaviour across head sets
namespace ha {
HA_DEV u32 fold_coherent(unsigned full) {
  const u32 f = full % ha_p.nheads;
  return ha_p.fixed_sets ? f : fold_set(f, ld32(ha_p.spread));
}
}  // namespace ha
__device__ void* ha_malloc_on_head(size_t nbytes, unsigned full) {
  if (nbytes > ha::MAX_BLOCK_SZ) return ha::large_alloc(nbytes);
  return ha::malloc_small((ha::u32)nbytes, ha::rfl(ha::fold_coherent(full)));
}
__device__ int ha_malloc_batch_on_head(size_t nbytes, int cnt, void** out, unsigned full) {
  return ha::malloc_batch_on(nbytes, cnt, out, ha::rfl(ha::fold_coherent(full)));
}
// the full-spread head set of the calling wave (what hamalloc folds)
__device__ unsigned ha_head_set_full() { return ha::rfl(ha::head_set()); }

HA_API void hafree(void* p) {
  if (!p) return;  // src/halloc.cu:240
  const ha::u64 off = (ha::u64)((char*)p - ha::ha_p.arena);
  const ha::u64 loff = (ha::u64)((char*)p - ha::ha_p.large);
  if (off < ha::ha_p.arena_bytes) ha::free_small(off);
  else if (loff < ha::ha_p.large_bytes) ha::large_free(loff);
  else free(p);  // not from this heap: the device heap's (src/halloc.cu:244-246)
}

// ------------------------------------------------------------------ host side
namespace ha {

static std::mutex g_mu;
static Params g_host;          // host copy of ha_p
static void* g_meta = nullptr;  // one allocation for all metadata

static thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}
#define HA_HIP(call)                                                                              \
  do {                                                                                            \
    hipError_t e_ = (call);                                                                       \
    if (e_ != hipSuccess) return ::ha::fail(HA_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

// Uncached arenas are never freed once a kernel has touched them (freeing
// such memory has corrupted later allocations on this platform); shutdown
// hands the arena back to this pool and a later init on the same device
// reuses the smallest one that is large enough.
struct UcArena {
  int dev;
  void* p;
  u64 bytes;
  bool in_use;
};
static std::vector<UcArena> g_uc;

static int arena_get(int dev, u64 bytes, bool cached, void** out) {
  *out = nullptr;
  if (cached) {
    if (hipMalloc(out, bytes) != hipSuccess) return fail(HA_E_NOMEM, "hipMalloc of the arena failed");
    return HA_OK;
  }
  UcArena* best = nullptr;
  for (auto& a : g_uc)
    if (a.dev == dev && !a.in_use && a.bytes >= bytes && (!best || a.bytes < best->bytes)) best = &a;
  if (!best) {
    void* p = nullptr;
    if (hipExtMallocWithFlags(&p, bytes, hipDeviceMallocUncached) != hipSuccess)
      return fail(HA_E_NOMEM, "allocation of the uncached arena failed");
    g_uc.push_back(UcArena{dev, p, bytes, false});
    best = &g_uc.back();
  }
  best->in_use = true;
  *out = best->p;
  return HA_OK;
}

static int arena_put(void* p, bool cached) {
  if (!p) return HA_OK;
  if (cached) {
    HA_HIP(hipFree(p));
    return HA_OK;
  }
  for (auto& a : g_uc)
    if (a.p == p) a.in_use = false;
  return HA_OK;
}

// largest prime p in [3, n) with n % p != 0 and nb % p != 0 (src/utils.cu:8-24)
static u32 max_prime_below(u32 n, u32 nb) {
  for (u32 p = n - 1; p >= 3 && p < n; p--) {
    u32 max_d = (u32)std::floor(std::sqrt((double)p));
    bool is_prime = true;
    for (u32 dd = 2; dd <= max_d; dd++)
      if (p % dd == 0) { is_prime = false; break; }
    if (is_prime && n % p && nb % p) return p;
  }
  return 0;
}
static u32 chunk_val(u32 chunk_sz) {
  u32 div3 = chunk_sz % 3 ? 1 : 3;
  if (chunk_sz % 3 == 0) chunk_sz /= 3;
  u32 sh = 0;
  for (; (1u << sh) < chunk_sz; sh++) {}
  return div3 << 16 | sh;
}
// size table (src/halloc.cu:363-394) with blocks per slab rounded to RUN, and
// to a run count THREAD_FREQ does not divide
static bool build_size_infos(u32 sb_sz, const ha_opts_c& o, SizeInfo* out) {
  for (u32 sid = 0; sid < NSIZES; ++sid) {
    SizeInfo& s = out[sid];
    std::memset(&s, 0, sizeof s);
    u32 unit = 1u << (sid / 2 + 3);
    s.block_sz = (sid % 2 ? 3 : 2) * unit;
    s.tmod = RUN;
    s.nblocks = sb_sz / s.block_sz / RUN * RUN;
    // a lane group's first probe starts at run (ticket / RUN) * THREAD_FREQ:
    // with THREAD_FREQ dividing the number of runs, first probes would reach
    // only every THREAD_FREQ-th run (31 of 341 for 192 B at 4 MiB slabs)
    if ((s.nblocks / RUN) % THREAD_FREQ == 0) s.nblocks -= RUN;
    s.chunk_id = sid % 2 + (sid < NSIZES / 2 ? 0 : 2);
    s.chunk_bytes = (s.chunk_id % 2 ? 3 : 2) * (s.chunk_id / 2 ? 128 : 8);
    s.chunk_sz = chunk_val(s.chunk_bytes);
    s.ncib = s.block_sz / s.chunk_bytes;
    s.nchunks = s.nblocks * s.ncib;
    const u32 bound = o.hash_step_frac > 0 ? (u32)(o.hash_step_frac * s.nblocks) : s.nblocks / 256 + s.nblocks / 64;
    u32 p = max_prime_below(bound, s.nblocks);
    if (!p || !s.nblocks) return false;
    s.hash_step = s.ncib * p;
    s.nchunks_m = mod_magic(s.nchunks);
    s.hash_step_m = mod_magic(s.hash_step);
    s.roomy_thr = (u32)(o.roomy_fraction * s.nchunks);
    s.busy_thr = (u32)(o.busy_fraction * s.nchunks);
    s.sparse_thr = (u32)(o.sparse_fraction * s.nchunks);
  }
  return true;
}

__global__ void init_state_k(Params P) {
  u32 i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < P.nsbs) {
    P.ctr[(u64)i * CTR_STRIDE] = CTR_CHUNK_NONE << CTR_CHUNK_SH;
    P.desc[i] = Desc{SZ_NONE, SZ_NONE, 0u, 0u};
    P.sb_lock[i] = 0;
  }
  if (i < P.nsw) {
    u32 lo = i * 64;
    u64 w = 0;
    if (lo < P.nsbs) w = P.nsbs - lo >= 64 ? ~0ull : ((1ull << (P.nsbs - lo)) - 1);
    P.sets[i] = w;  // the free set
  }
  // the units past the end of the large region are never free
  if (i == 0 && P.lwords && P.lunits % 64) P.lbits[P.lwords - 1] = ~low_mask(P.lunits % 64);
  if (i == 0) *P.spread = P.nheads;  // every set in use until size classes appear
  if (i < P.nheads * NSIZES) {
    P.head[i * HOT_STRIDE] = SB_NONE;
    P.cached[i * HOT_STRIDE] = SB_NONE;
    P.head_lock[i * HOT_STRIDE] = 0;
    P.ticket[i * HOT_STRIDE] = TICKET_INIT;
    P.forced[i * HOT_STRIDE] = SB_NONE;
  }
}

static u64 align_up(u64 x, u64 a) { return (x + a - 1) / a * a; }
static int scratch_ptr(u64** out);
static int g_heap_dev = -1;  // the device ha_init initialized

static int shutdown_locked() {
  if (!g_host.initialized) return HA_OK;
  Params zero;
  std::memset(&zero, 0, sizeof zero);
  HA_HIP(hipDeviceSynchronize());
  HA_HIP(hipMemcpyToSymbol(HIP_SYMBOL(ha_p), &zero, sizeof zero));
  int rc = arena_put(g_host.arena, g_host.cached_arena != 0);
  if (rc) return rc;
  if (g_meta) HA_HIP(hipFree(g_meta));
  g_meta = nullptr;
  std::memset(&g_host, 0, sizeof g_host);
  return HA_OK;
}

static int init_locked(const ha_opts_c& o) {
  if (g_host.initialized) {
    int rc = shutdown_locked();
    if (rc) return rc;
  }
  if (o.sb_sz_sh < (int)MIN_SB_SZ_SH || o.sb_sz_sh > (int)MAX_SB_SZ_SH)
    return fail(HA_E_INVALID, "sb_sz_sh must be in 21..26");
  if (o.nheads < 0 || o.nheads > (int)MAX_NHEADS) return fail(HA_E_INVALID, "nheads must be in 0..64");
  if (!(o.halloc_fraction > 0 && o.halloc_fraction <= 1) || !(o.sparse_fraction >= 0) ||
      !(o.sparse_fraction <= o.roomy_fraction) || !(o.roomy_fraction <= o.busy_fraction) ||
      !(o.busy_fraction <= 1))
    return fail(HA_E_INVALID, "fractions must satisfy 0 <= sparse <= roomy <= busy <= 1, 0 < halloc <= 1");
  if (!(o.hash_step_frac >= 0 && o.hash_step_frac < 1)) return fail(HA_E_INVALID, "hash_step_frac must be in [0, 1)");
  if (o.cached_arena != 0 && o.cached_arena != 1) return fail(HA_E_INVALID, "cached_arena must be 0 or 1");
  if (o.fixed_sets != 0 && o.fixed_sets != 1) return fail(HA_E_INVALID, "fixed_sets must be 0 or 1");
  int dev = 0;
  HA_HIP(hipGetDevice(&dev));
  u64* sp = nullptr;
  if (int rc = scratch_ptr(&sp)) return rc;  // the scratch address of this device, once
  hipDeviceProp_t prop;
  HA_HIP(hipGetDeviceProperties(&prop, dev));
  u64 memory = std::min<u64>(o.memory, 3ull * (u64)prop.totalGlobalMem / 4ull);  // src/halloc.cu:282
  u64 halloc_memory = (u64)(o.halloc_fraction * (double)memory);
  u64 cuda_memory = memory - halloc_memory;                                        // src/halloc.cu:285-289
  Params P;
  std::memset(&P, 0, sizeof P);
  P.sb_sz_sh = (u32)o.sb_sz_sh;
  P.sb_sz = 1u << P.sb_sz_sh;
  P.nsbs = (u32)(halloc_memory / P.sb_sz);
  if (P.nsbs == 0 || P.nsbs > MAX_NSBS) return fail(HA_E_INVALID, "memory gives no slab or too many slabs");
  // nheads 0: as many sets (a power of two, 1..16) as keep one head of every
  // size class in every set within a third of the slabs
  P.nheads = (u32)o.nheads;
  if (P.nheads == 0) {
    P.nheads = 1;
    while (2 * P.nheads <= 16 && 2 * P.nheads * NSIZES * 3 <= P.nsbs) P.nheads *= 2;
  }
  P.fixed_sets = (u32)o.fixed_sets;
  // no more sets than one size class can hold heads in (an adaptive spread
  // never exceeds that), so that free_scarce and the head searches count
  // only sets that can be in use. With 64 adaptive sets on 96 slabs, the 32
  // that could never be used made 32 free slabs look scarce: pairs shared
  // heads instead of taking free slabs and the 16 B kernel ran 4-6x slower;
  // 64 fixed sets there, each pair wanting a head and a cached head, took
  // 86-362 ms per launch.
  if (P.nsbs / 3u < P.nheads) {
    u32 s = 1;
    while (2 * s <= P.nsbs / 3u) s *= 2;
    P.nheads = s;
  }
  P.nh_m = mod_magic(P.nheads);
  P.nh8_m = mod_magic(P.nheads / 8 ? P.nheads / 8 : 1);
  P.cached_arena = (u32)o.cached_arena;
  P.nsw = (P.nsbs + 63) / 64;
  P.bit_words = P.sb_sz / BLOCK_STEP / 64;
  P.nib_words = P.sb_sz / BLOCK_STEP / 16;
  P.cuda_memory = cuda_memory;
  if (!build_size_infos(P.sb_sz, o, P.si)) return fail(HA_E_INVALID, "no size table for this slab size");
  P.arena_bytes = (u64)P.nsbs * P.sb_sz;
  // the large-block region follows the slabs in the same allocation
  P.lunits = (u32)std::min<u64>(cuda_memory >> LARGE_UNIT_SH, 0xfffffff0ull);
  P.large_bytes = (u64)P.lunits << LARGE_UNIT_SH;
  P.lwords = (P.lunits + 63) / 64;
  P.lnclasses = 0;
  while (P.lnclasses < LARGE_NCLASSES && (1ull << P.lnclasses) <= P.lunits) ++P.lnclasses;
  // metadata offsets, each section 256 B aligned
  u64 off = 0;
  u64 o_bits = off; off = align_up(off + (u64)P.nsbs * P.bit_words * 8, 256);
  u64 o_nibs = off; off = align_up(off + (u64)P.nsbs * P.nib_words * 8, 256);
  u64 o_ctr = off; off = align_up(off + (u64)P.nsbs * CTR_STRIDE * 8, 256);
  u64 o_desc = off; off = align_up(off + (u64)P.nsbs * sizeof(Desc), 256);
  u64 o_lock = off; off = align_up(off + (u64)P.nsbs * 4, 256);
  u64 o_sets = off; off = align_up(off + (u64)NSETS * P.nsw * 8, 256);
  u64 hot_bytes = (u64)P.nheads * NSIZES * HOT_STRIDE * 4;
  u64 o_head = off; off = align_up(off + hot_bytes, 256);
  u64 o_cached = off; off = align_up(off + hot_bytes, 256);
  u64 o_hlock = off; off = align_up(off + hot_bytes, 256);
  u64 o_ticket = off; off = align_up(off + hot_bytes, 256);
  u64 o_clock = off; off = align_up(off + (u64)NSIZES * HOT_STRIDE * 4, 256);
  u64 o_sep = off; off = align_up(off + hot_bytes, 256);
  u64 o_forced = off; off = align_up(off + hot_bytes, 256);
  u64 o_active = off; off = align_up(off + 4, 256);
  u64 o_free_ep = off; off = align_up(off + 4, 256);
  u64 o_spread = off; off = align_up(off + 4, 256);
  u64 o_seen = off; off = align_up(off + 4, 256);
  u64 o_lbits = off; off = align_up(off + (u64)P.lwords * 8, 256);
  u64 o_lstart = off; off = align_up(off + (u64)P.lunits * 4, 256);
  u64 o_ltick = off; off = align_up(off + (u64)LARGE_NCLASSES * HOT_STRIDE * 4, 256);
  void* arena = nullptr;
  int rc = arena_get(dev, P.arena_bytes + P.large_bytes, o.cached_arena != 0, &arena);
  if (rc) return rc;
  void* meta = nullptr;
  if (hipMalloc(&meta, off) != hipSuccess) {
    (void)arena_put(arena, o.cached_arena != 0);
    return fail(HA_E_NOMEM, "hipMalloc of the metadata failed");
  }
  char* m = (char*)meta;
  P.arena = (char*)arena;
  P.bits = (u64*)(m + o_bits);
  P.nibs = (u64*)(m + o_nibs);
  P.ctr = (u64*)(m + o_ctr);
  P.desc = (Desc*)(m + o_desc);
  P.sb_lock = (u32*)(m + o_lock);
  P.sets = (u64*)(m + o_sets);
  P.head = (u32*)(m + o_head);
  P.cached = (u32*)(m + o_cached);
  P.head_lock = (u32*)(m + o_hlock);
  P.ticket = (u32*)(m + o_ticket);
  P.class_lock = (u32*)(m + o_clock);
  P.search_ep = (u32*)(m + o_sep);
  P.forced = (u32*)(m + o_forced);
  P.active = (u32*)(m + o_active);
  P.free_ep = (u32*)(m + o_free_ep);
  P.spread = (u32*)(m + o_spread);
  P.class_seen = (u32*)(m + o_seen);
  P.large = P.arena + P.arena_bytes;
  P.lbits = (u64*)(m + o_lbits);
  P.lstart = (u32*)(m + o_lstart);
  P.lticket = (u32*)(m + o_ltick);
  P.initialized = 1;
  HA_HIP(hipMemset(meta, 0, off));
  u32 n = std::max(std::max(P.nsbs, P.nsw), P.nheads * NSIZES);
  init_state_k<<<(n + 255) / 256, 256>>>(P);
  HA_HIP(hipGetLastError());
  HA_HIP(hipMemcpyToSymbol(HIP_SYMBOL(ha_p), &P, sizeof P));
  HA_HIP(hipDeviceSynchronize());
  g_host = P;
  g_meta = meta;
  g_heap_dev = dev;
  return HA_OK;
}

const Params& host_params() { return g_host; }

// Free bytes and free-slab bytes over the heap (find_free_mem_k,
// src/statistics.cuh:34-48) as a two-stage reduction over a buffer of
// STAT_WORDS words: STAT_PARTS workgroups stride over the slabs and each
// writes its two partial sums (words 2b, 2b + 1); then one workgroup adds the
// partials and writes the totals (words 2 * STAT_PARTS and 2 * STAT_PARTS + 1).
// Every word a stage reads was written by the stage before it in the same
// call, so the buffer needs no clearing, and it holds STAT_WORDS words for any
// number of slabs.
constexpr u32 STAT_PARTS = 32, STAT_BS = 256;
constexpr u32 STAT_WORDS = 2 * STAT_PARTS + 2;

__global__ void __launch_bounds__(STAT_BS) free_mem_part_k(u64* buf) {
  __shared__ u64 red[2][STAT_BS];
  u64 f = 0, fs = 0;
  for (u32 i = blockIdx.x * STAT_BS + threadIdx.x; i < ha_p.nsbs; i += STAT_PARTS * STAT_BS) {
    const u32 cs = ld32(&ha_p.desc[i].chunk_sz);
    const u32 used = cs ? chunk_mul(ctr_count(ld64(ctr_of(i))), cs) : 0u;
    f += ha_p.sb_sz - used;
    if (cs == 0) fs += ha_p.sb_sz;
  }
  red[0][threadIdx.x] = f;
  red[1][threadIdx.x] = fs;
  __syncthreads();
  for (u32 h = STAT_BS / 2; h; h >>= 1) {
    if (threadIdx.x < h) {
      red[0][threadIdx.x] += red[0][threadIdx.x + h];
      red[1][threadIdx.x] += red[1][threadIdx.x + h];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    buf[2 * blockIdx.x] = red[0][0];
    buf[2 * blockIdx.x + 1] = red[1][0];
  }
}

__global__ void __launch_bounds__(64) free_mem_sum_k(u64* buf) {
  const u32 t = threadIdx.x;
  u64 f = t < STAT_PARTS ? buf[2 * t] : 0ull, fs = t < STAT_PARTS ? buf[2 * t + 1] : 0ull;
  for (u32 off = 32; off; off >>= 1) {
    f += __shfl_xor(f, off);
    fs += __shfl_xor(fs, off);
  }
  if (t == 0) {
    buf[2 * STAT_PARTS] = f;
    buf[2 * STAT_PARTS + 1] = fs;
  }
}

// both stages on the null stream over buf (STAT_WORDS device words), then the
// totals copied to the host
static int stats_on(u64* buf, u64* free_mem, u64* free_slab_mem) {
  free_mem_part_k<<<STAT_PARTS, STAT_BS>>>(buf);
  HA_HIP(hipGetLastError());
  free_mem_sum_k<<<1, 64>>>(buf);
  HA_HIP(hipGetLastError());
  u64 h[2];
  HA_HIP(hipMemcpy(h, buf + 2 * STAT_PARTS, 16, hipMemcpyDeviceToHost));
  *free_mem = h[0];
  if (free_slab_mem) *free_slab_mem = h[1];
  return HA_OK;
}

static int stats_locked(u64* free_mem, u64* free_slab_mem) {
  if (!g_host.initialized) return fail(HA_E_STATE, "allocator not initialized");
  u64* d = nullptr;
  HA_HIP(hipMalloc(&d, STAT_WORDS * 8));
  const int rc = stats_on(d, free_mem, free_slab_mem);
  HA_HIP(hipFree(d));
  return rc;
}

// ------------------------------------------- statistics without allocation
// ha_free_bytes_noalloc_c runs the statistics kernels over `scratch`, a
// file-scope device array, instead of a buffer from hipMalloc.  Each device
// holds its own copy of the array for the life of the process.  Host code
// never passes the array's name to a kernel (on the host it is a stand-in,
// not a device address): the address comes from hipGetSymbolAddress, taken
// once per device with that device current, at ha_init or at the first call.
constexpr u32 SCRATCH_WORDS = 128;  // 1 KiB per device
static_assert(STAT_WORDS <= SCRATCH_WORDS, "the statistics buffer must fit the scratch array");
static __device__ __align__(256) u64 scratch[SCRATCH_WORDS];

constexpr int MAX_DEVICES = 64;
static u64* g_scratch_ptr[MAX_DEVICES] = {};

// the scratch address on the current device (set up on the first use there)
static int scratch_ptr(u64** out) {
  int dev = 0;
  HA_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= MAX_DEVICES) return fail(HA_E_INVALID, "device ordinal out of range");
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    HA_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return fail(HA_E_HIP, "hipGetSymbolAddress returned no address for the scratch array");
    g_scratch_ptr[dev] = (u64*)p;
  }
  *out = g_scratch_ptr[dev];
  return HA_OK;
}

static int stats_noalloc_locked(u64* free_mem) {
  if (!g_host.initialized) return fail(HA_E_STATE, "allocator not initialized");
  int dev = 0;
  HA_HIP(hipGetDevice(&dev));
  if (dev != g_heap_dev) return fail(HA_E_STATE, "the current device is not the one ha_init initialized");
  if (STAT_WORDS > SCRATCH_WORDS) return fail(HA_E_INVALID, "statistics buffer larger than the scratch array");
  u64* buf = nullptr;
  const int rc = scratch_ptr(&buf);
  if (rc) return rc;
  return stats_on(buf, free_mem, nullptr);
}

}  // namespace ha

// --------------------------------------------------------------------- C ABI
extern "C" {

void ha_default_opts(ha_opts_c* o) {
  halloc_opts_t d;
  o->memory = d.memory;
  o->halloc_fraction = d.halloc_fraction;
  o->busy_fraction = d.busy_fraction;
  o->roomy_fraction = d.roomy_fraction;
  o->sparse_fraction = d.sparse_fraction;
  o->sb_sz_sh = d.sb_sz_sh;
  o->nheads = 32;
  o->hash_step_frac = 0.0;
  o->cached_arena = 0;
  o->fixed_sets = 0;
}

int ha_init_c(const ha_opts_c* o) {
  if (!o) return ha::fail(HA_E_INVALID, "null options");
  std::lock_guard<std::mutex> g(ha::g_mu);
  return ha::init_locked(*o);
}

int ha_shutdown_c(void) {
  std::lock_guard<std::mutex> g(ha::g_mu);
  return ha::shutdown_locked();
}

int ha_free_bytes_c(uint64_t* out) {
  if (!out) return ha::fail(HA_E_INVALID, "null output");
  std::lock_guard<std::mutex> g(ha::g_mu);
  return ha::stats_locked((ha::u64*)out, nullptr);
}

int ha_free_bytes_noalloc_c(uint64_t* out) {
  if (!out) return ha::fail(HA_E_INVALID, "null output");
  std::lock_guard<std::mutex> g(ha::g_mu);
  return ha::stats_noalloc_locked((ha::u64*)out);
}

int ha_extfrag_c(int ideal, double* out) {
  if (!out) return ha::fail(HA_E_INVALID, "null output");
  std::lock_guard<std::mutex> g(ha::g_mu);
  ha::u64 free_mem = 0, free_slabs = 0;
  int rc = ha::stats_locked(&free_mem, &free_slabs);
  if (rc) return rc;
  // largest block the large region can hand out now (the counterpart of
  // find_max_alloc_k's bisection of the CUDA heap, src/statistics.cuh:14-30)
  const ha::Params& P = ha::g_host;
  ha::u64 max_alloc = 0;
  if (P.lunits) {
    std::vector<ha::u64> lb(P.lwords);
    HA_HIP(hipDeviceSynchronize());
    HA_HIP(hipMemcpy(lb.data(), P.lbits, lb.size() * 8, hipMemcpyDeviceToHost));
    max_alloc = (ha::u64)ha::large_max_units(lb.data(), P.lwords, P.lnclasses) << ha::LARGE_UNIT_SH;
  }
  if (ideal) max_alloc += free_slabs;
  free_mem += P.large_bytes;  // the whole non-slab share, as the reference adds cuda_mem_g
  *out = free_mem ? 1.0 - (double)max_alloc / (double)free_mem : 0.0;
  return HA_OK;
}

int ha_geometry_c(ha_geometry* g) {
  if (!g) return ha::fail(HA_E_INVALID, "null output");
  std::lock_guard<std::mutex> lk(ha::g_mu);
  const ha::Params& P = ha::g_host;
  if (!P.initialized) return ha::fail(HA_E_STATE, "allocator not initialized");
  std::memset(g, 0, sizeof *g);
  g->nsbs = P.nsbs; g->sb_sz = P.sb_sz; g->sb_sz_sh = P.sb_sz_sh; g->nheads = P.nheads;
  g->nsw = P.nsw; g->bit_words = P.bit_words; g->nib_words = P.nib_words;
  g->arena_bytes = P.arena_bytes; g->large_bytes = P.large_bytes;
  for (ha::u32 s = 0; s < ha::NSIZES; ++s) std::memcpy(g->size_table + 12 * s, &P.si[s], 12 * sizeof(uint32_t));
  g->arena_base = (uint64_t)(uintptr_t)P.arena;
  g->large_units = P.lunits;
  g->large_classes = P.lnclasses;
  return HA_OK;
}

const char* ha_last_error(void) { return ha::g_err.c_str(); }

}  // extern "C"

// ------------------------------------------------ C++ host API of halloc.h
void ha_init(halloc_opts_t opts) {
  ha_opts_c o;
  ha_default_opts(&o);
  o.memory = opts.memory;
  o.halloc_fraction = opts.halloc_fraction;
  o.busy_fraction = opts.busy_fraction;
  o.roomy_fraction = opts.roomy_fraction;
  o.sparse_fraction = opts.sparse_fraction;
  o.sb_sz_sh = opts.sb_sz_sh;
  if (ha_init_c(&o) != HA_OK) {
    std::fprintf(stderr, "ha_init: %s\n", ha_last_error());
    std::exit(-1);
  }
}

void ha_shutdown(void) {
  if (ha_shutdown_c() != HA_OK) std::fprintf(stderr, "ha_shutdown: %s\n", ha_last_error());
}

double ha_extfrag(bool ideal) {
  double v = 0;
  if (ha_extfrag_c(ideal ? 1 : 0, &v) != HA_OK) {
    std::fprintf(stderr, "ha_extfrag: %s\n", ha_last_error());
    std::exit(-1);
  }
  return v;
}
