// Brute-force check of the wave-run union arithmetic (halloc_amd/csrc/ha_bits.h)
// used by the device malloc probe and the free fast path.  Built with g++ by
// tests/test_native.py; prints "ok" and exits 0 when every case agrees.
#include <cstdio>
#include <cstdlib>
#include <map>
#include <random>

#include "../../halloc_amd/csrc/ha_bits.h"

using ha::u32;
using ha::u64;

static int fails = 0;
#define CHECK(c)                                                                \
  do {                                                                          \
    if (!(c)) {                                                                 \
      if (fails++ < 20) std::printf("FAIL line %d: %s\n", __LINE__, #c);        \
    }                                                                           \
  } while (0)

int main() {
  std::mt19937_64 rng(12345);
  const u32 ks[4] = {1, 2, 4, 8};
  for (int iter = 0; iter < 200000; ++iter) {
    u32 k = ks[rng() % 4];
    u32 nruns = 2 + rng() % 40;
    u32 nchunks = nruns * 64 * k;
    u32 old = (u32)rng();
    u32 slot0 = old & 63, split = 64 - slot0;
    u32 runA = (u32)(rng() % nruns) * 64 * k, runB = (u32)(rng() % nruns) * 64 * k;
    u64 off = (rng() % 4 == 0) ? 0 : (rng() % (8ull * nchunks)) / k * k;
    u64 P = rng() & rng();
    if (rng() % 3 == 0) P = ~0ull;
    u64 amask = split >= 64 ? ~0ull : ((1ull << split) - 1);
    u32 pA = (u32)((runA + off) % nchunks), pB = (u32)((runB + off) % nchunks);
    // per-lane expected positions (the model's per-lane formula)
    u32 pos[64];
    for (u32 l = 0; l < 64; ++l) {
      u32 slot = l < split ? slot0 + l : l - split;
      u32 rb = l < split ? runA : runB;
      pos[l] = (u32)(((u64)rb + slot * k + off) % nchunks);
    }
    u64 ufield = (1ull << k) - 1;
    // ---- bit array: per word union restricted to the lane's segment
    for (u32 me = 0; me < 64; ++me) {
      if (!((P >> me) & 1)) continue;
      bool inA = me < split;
      u64 myseg = inA ? amask : ~amask;
      u32 st = inA ? pA + slot0 * k : pB - split * k;
      u32 word, issuer;
      u64 uni;
      ha::run_word_bits(P & myseg, me, st, k, word, uni, issuer);
      u32 c = st + me * k;
      CHECK((c >= nchunks ? c - nchunks : c) == pos[me]);
      u32 wreal = word >= nchunks / 64 ? word - nchunks / 64 : word;
      CHECK(wreal == pos[me] / 64);
      u64 want = 0;
      u32 wiss = 64;
      for (u32 l = 0; l < 64; ++l) {
        if (!((P & myseg) >> l & 1)) continue;
        if (pos[l] / 64 != pos[me] / 64) continue;
        want |= ufield << (pos[l] % 64);
        if (l < wiss) wiss = l;
      }
      CHECK(uni == want);
      CHECK(issuer == wiss);
      // ---- nibble array with value k
      u32 nword, nissuer;
      u64 nuni;
      ha::run_word_nibs(P & myseg, me, st, k, (u64)k, nword, nuni, nissuer);
      u32 nreal = nword >= nchunks / 16 ? nword - nchunks / 16 : nword;
      CHECK(nreal == pos[me] / 16);
      u64 nwant = 0;
      u32 niss = 64;
      for (u32 l = 0; l < 64; ++l) {
        if (!((P & myseg) >> l & 1)) continue;
        if (pos[l] / 16 != pos[me] / 16) continue;
        nwant |= (u64)k << (4 * (pos[l] % 16));
        if (l < niss) niss = l;
      }
      CHECK(nuni == nwant);
      CHECK(nissuer == niss);
    }
    // ---- free fast path: real positions, virtual start ich[l0] - l0*k
    u64 A = P & amask;
    if (A) {
      u32 l0 = (u32)__builtin_ctzll(A);
      u32 sa = pos[l0] - l0 * k;
      for (u32 me = 0; me < 64; ++me) {
        if (!((A >> me) & 1)) continue;
        if (pos[me] != sa + me * k) continue;  // the run wrapped around nchunks
        u32 word, issuer;
        u64 uni;
        ha::run_word_nibs(A, me, sa, k, 0xFull, word, uni, issuer);
        CHECK(word == pos[me] / 16);
        u64 nwant = 0;
        for (u32 l = 0; l < 64; ++l)
          if ((A >> l & 1) && pos[l] / 16 == pos[me] / 16) nwant |= 0xFull << (4 * (pos[l] % 16));
        bool contiguous = true;
        for (u32 l = 0; l < 64; ++l)
          if ((A >> l & 1) && pos[l] != sa + l * k) contiguous = false;
        if (contiguous) CHECK(uni == nwant);
      }
    }
  }
  // spread_bits: bit i -> bit f*i
  for (u32 f : {1u, 2u, 4u, 8u, 16u, 32u}) {
    for (int t = 0; t < 10000; ++t) {
      u64 x = rng() & (f == 1 ? ~0ull : ((1ull << (64 / f)) - 1));
      u64 want = 0;
      for (u32 i = 0; i < 64 / f; ++i)
        if (x >> i & 1) want |= 1ull << (f * i);
      CHECK(ha::spread_bits(x, f) == want);
    }
  }
  // mod_by against %: every n up to 2^17, then random n, with random x and
  // the edge values around multiples of n and 2^32
  for (u32 n = 1; n <= (1u << 17); ++n) {
    const u32 m = ha::mod_magic(n);
    const u32 xs[6] = {0u, n - 1, n, 0xFFFFFFFFu, 0xFFFFFFFFu - (0xFFFFFFFFu % n), (u32)rng()};
    for (u32 x : xs) CHECK(ha::mod_by(x, n, m) == x % n);
    for (int t = 0; t < 8; ++t) {
      const u32 x = (u32)rng();
      CHECK(ha::mod_by(x, n, m) == x % n);
    }
  }
  for (int t = 0; t < 2000000; ++t) {
    const u32 n = 1u + (u32)(rng() % 0xFFFFFFFFull), x = (u32)rng();
    CHECK(ha::mod_by(x, n, ha::mod_magic(n)) == x % n);
  }
  // unit fields of the large region: free / used aligned fields, widening,
  // select and lowest-n against loops over the aligned positions
  for (int t = 0; t < 400000; ++t) {
    const u32 k = 1u << (rng() % 7);
    u64 v = rng();
    if (t % 3 == 0) v &= rng();  // sparser words
    if (t % 5 == 0) v |= rng();  // denser words
    if (t % 97 == 0) v = 0;
    const u64 fm = k == 64 ? ~0ull : ((1ull << k) - 1);
    u64 want_free = 0, want_used = 0;
    for (u32 p = 0; p < 64; p += k) {
      if (((v >> p) & fm) == 0) want_free |= 1ull << p;
      else want_used |= 1ull << p;
    }
    CHECK(ha::free_fields(v, k) == want_free);
    CHECK(ha::used_fields(v, k) == want_used);
    u64 wide = 0;
    for (u32 p = 0; p < 64; p += k)
      if (want_free >> p & 1) wide |= fm << p;
    CHECK(ha::widen_fields(want_free, k) == wide);
    const u64 x = rng() & rng();
    u32 r = 0;
    for (u32 b = 0; b < 64; ++b)
      if (x >> b & 1) CHECK(ha::select_bit(x, r++) == b);
    const u32 n = (u32)(rng() % 66);
    u64 low = 0;
    u32 cnt = 0;
    for (u32 b = 0; b < 64 && cnt < n; ++b)
      if (x >> b & 1) { low |= 1ull << b; ++cnt; }
    CHECK(ha::lowest_bits(x, n) == low);
  }
  // the largest block a region's words can hand out, against a scan of every
  // aligned block of every class
  for (int t = 0; t < 3000; ++t) {
    const u32 nw = 1 + (u32)(rng() % 40), ncl = 1 + (u32)(rng() % 12);
    u64 lb[40];
    for (u32 w = 0; w < nw; ++w) {
      lb[w] = rng() % 4 == 0 ? 0ull : (rng() & rng() & rng());
      if (rng() % 6 == 0) lb[w] = ~0ull;
    }
    u64 want = 0;
    for (u32 c = 0; c < ncl; ++c) {
      const u64 k = 1ull << c;
      bool found = false;
      for (u64 u = 0; u + k <= (u64)nw * 64 && !found; u += k) {
        bool z = true;
        for (u64 x2 = u; x2 < u + k && z; ++x2) z = !((lb[x2 / 64] >> (x2 % 64)) & 1);
        found = z;
      }
      if (found) want = k;
    }
    CHECK(ha::large_max_units(lb, nw, ncl) == want);
  }
  if (fails) {
    std::printf("%d failures\n", fails);
    return 1;
  }
  std::printf("ok\n");
  return 0;
}
