// ha_bits.h -- pure bit arithmetic of the wave-run layout, shared by device
// and host code (the host side is unit-tested with g++ against brute force).
#pragma once

#include <cstdint>

#if defined(__HIPCC__) || defined(__HIP__)
#define HA_HD __host__ __device__
#else
#define HA_HD
#endif

namespace ha {

using u32 = uint32_t;
using u64 = uint64_t;

// bit i of x moves to bit f*i (x holds at most 64/f meaningful bits)
HA_HD inline u64 spread_bits(u64 x, u32 f) {
  switch (f) {
    case 1:
      return x;
    case 2:
      x &= 0xffffffffull;
      x = (x | x << 16) & 0x0000FFFF0000FFFFull;
      x = (x | x << 8) & 0x00FF00FF00FF00FFull;
      x = (x | x << 4) & 0x0F0F0F0F0F0F0F0Full;
      x = (x | x << 2) & 0x3333333333333333ull;
      return (x | x << 1) & 0x5555555555555555ull;
    case 4:
      x &= 0xffffull;
      x = (x | x << 24) & 0x000000FF000000FFull;
      x = (x | x << 12) & 0x000F000F000F000Full;
      x = (x | x << 6) & 0x0303030303030303ull;
      return (x | x << 3) & 0x1111111111111111ull;
    case 8:
      x &= 0xffull;
      x = (x | x << 28) & 0x0000000F0000000Full;
      x = (x | x << 14) & 0x0003000300030003ull;
      return (x | x << 7) & 0x0101010101010101ull;
    case 16:
      return (x & 1) | ((x & 2) << 15) | ((x & 4) << 30) | ((x & 8) << 45);
    case 32:
      return (x & 1) | ((x & 2) << 31);
    default:
      return 0;
  }
}

// A run of 64 k-chunk slots starts at chunk s (slot of lane l = s + l*k,
// unwrapped); k is a power of two (1, 2, 4 or 8), so the divisions by k are
// shifts.  s must be a multiple of k (mod 2^32): for other starts the unions
// below are wrong, so a caller that derives s from block positions checks it.  For lane `me` and the lanes in bitmap B, return the unwrapped
// index of the word holding me's slot, the union of B-lanes' fields inside that
// word, and the lowest B-lane in that word (64 if none).
//   bit array:    64 chunks per word, field = k one-bits
//   nibble array: 16 chunks per word, field = 4-bit value `val` at the slot start
HA_HD inline void run_word_bits(u64 B, u32 me, u32 s, u32 k, u32& word, u64& uni, u32& issuer) {
  u32 c = s + me * k;
  word = c >> 6;
  const u32 ks = (u32)__builtin_ctz(k);
  int lo_u = (int)(word << 6) - (int)s;
  u32 lo = lo_u <= 0 ? 0u : (u32)lo_u >> ks;
  u32 hi = (u32)(lo_u + 63) >> ks;
  hi = hi > 63 ? 63 : hi;
  u32 n = hi - lo + 1;
  u64 m = n >= 64 ? ~0ull : ((1ull << n) - 1);
  u64 sub = (B >> lo) & m;
  u32 sh = s + lo * k - (word << 6);
  uni = (spread_bits(sub, k) * ((1ull << k) - 1)) << sh;
  issuer = sub ? lo + (u32)__builtin_ctzll(sub) : 64u;
}
HA_HD inline void run_word_nibs(u64 B, u32 me, u32 s, u32 k, u64 val, u32& word, u64& uni,
                                              u32& issuer) {
  u32 c = s + me * k;
  word = c >> 4;
  const u32 ks = (u32)__builtin_ctz(k);
  int lo_u = (int)(word << 4) - (int)s;
  u32 lo = lo_u <= 0 ? 0u : (u32)lo_u >> ks;
  u32 hi = (u32)(lo_u + 15) >> ks;
  hi = hi > 63 ? 63 : hi;
  u32 n = hi - lo + 1;
  u64 m = n >= 64 ? ~0ull : ((1ull << n) - 1);
  u64 sub = (B >> lo) & m;
  u32 sh = s + lo * k - (word << 4);
  uni = (spread_bits(sub, 4 * k) * val) << (4 * sh);
  issuer = sub ? lo + (u32)__builtin_ctzll(sub) : 64u;
}

// ----------------------------------------------------------- unit fields
// The large-block region keeps one bit per 4 KiB unit in 64-bit words; a block
// of k = 2^c units (k <= 64) is a k-bit field aligned to k inside one word.

// k one-bits (k in 1..64)
HA_HD inline u64 low_mask(u32 k) { return k >= 64 ? ~0ull : ((1ull << k) - 1); }
// one bit at every multiple of k: the start bits of the aligned k-bit fields
HA_HD inline u64 field_starts(u32 k) { return k >= 64 ? 1ull : ~0ull / low_mask(k); }
// start bits of the aligned k-bit fields of v that are entirely zero
HA_HD inline u64 free_fields(u64 v, u32 k) {
  u64 f = ~v;
  for (u32 s = 1; s < k; s <<= 1) f &= f >> s;
  return f & field_starts(k);
}
// start bits of the aligned k-bit fields of v that hold at least one one-bit
HA_HD inline u64 used_fields(u64 v, u32 k) {
  u64 f = v;
  for (u32 s = 1; s < k; s <<= 1) f |= f >> s;
  return f & field_starts(k);
}
// every start bit of b (aligned to k) widened to its k-bit field
HA_HD inline u64 widen_fields(u64 b, u32 k) { return b * low_mask(k); }
// position of the r-th lowest one-bit of x (r < popcount(x))
HA_HD inline u32 select_bit(u64 x, u32 r) {
  u32 pos = 0;
  for (u32 w = 32; w; w >>= 1) {
    const u32 c = (u32)__builtin_popcountll(x & low_mask(w));
    if (r >= c) {
      r -= c;
      x >>= w;
      pos += w;
    }
  }
  return pos;
}
// the n lowest one-bits of x
HA_HD inline u64 lowest_bits(u64 x, u32 n) {
  if (n == 0) return 0;
  if (n >= (u32)__builtin_popcountll(x)) return x;
  return x & low_mask(select_bit(x, n - 1) + 1);
}
// The first word a group of large requests of class c looks at: its ticket
// spread over the nw words by a multiplicative (Fibonacci) hash.
HA_HD inline u32 large_start(u32 ticket, u32 c, u32 nw) {
  const u32 h = (ticket * 0x9E3779B1u) ^ (c * 0x85EBCA77u);
  return (u32)(((u64)h * nw) >> 32);
}

// The largest block (in units) the large region's bit words lb[0..nw) can
// hand out now, over classes 0..ncl-1: a free aligned field inside a word
// for up to 64 units, an aligned run of zero words above.
inline u64 large_max_units(const u64* lb, u32 nw, u32 ncl) {
  u64 best = 0;
  for (u32 c = 0; c < ncl; ++c) {
    bool found = false;
    if (c <= 6) {
      for (u32 w = 0; w < nw && !found; ++w) found = free_fields(lb[w], 1u << c) != 0;
    } else {
      const u32 m = 1u << (c - 6);
      for (u32 g = 0; g + 1 <= nw / m && !found; ++g) {
        bool z = true;
        for (u32 j = 0; j < m && z; ++j) z = lb[(u64)g * m + j] == 0;
        found = z;
      }
    }
    if (!found) break;  // no class above fits either
    best = 1ull << c;
  }
  return best;
}

// x mod n without a division, for x < 2^32 and 1 <= n, given m = mod_magic(n)
// = floor(2^32 / n) (0 for n = 1): q = floor(x * m / 2^32) is floor(x / n) or
// one less, since x * m / 2^32 lies in (x / n - 1, x / n]; one correction step
// finishes
HA_HD inline u32 mod_magic(u32 n) { return n < 2 ? 0u : (u32)(0x100000000ull / n); }
HA_HD inline u32 mod_by(u32 x, u32 n, u32 m) {
  const u32 q = (u32)(((u64)x * m) >> 32);
  u32 r = x - q * n;
  r = r >= n ? r - n : r;
  return n < 2 ? 0u : r;
}

}  // namespace ha
