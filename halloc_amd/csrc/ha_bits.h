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
// shifts.  For lane `me` and the lanes in bitmap B, return the unwrapped
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
