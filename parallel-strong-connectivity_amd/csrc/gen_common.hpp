// Counter-based generators shared by the host (g++) and device (hipcc)
// graph builders, so both produce bit-identical graphs for a seed.
// Every random draw is a pure function of (seed, item index).
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#define PSC_HD __host__ __device__ __forceinline__
#else
#define PSC_HD inline
#endif

namespace psc {

PSC_HD uint64_t mix64(uint64_t z) {   // splitmix64 finaliser
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// i-th 64-bit draw of stream `stream` under `seed`.
PSC_HD uint64_t draw64(uint64_t seed, uint64_t stream, uint64_t i) {
  return mix64(seed * 0xD1B54A32D192ED03ull + stream * 0x9E3779B97F4A7C15ull +
               i * 0xA0761D6478BD642Full + 0x2545F4914F6CDD1Dull);
}

// Probability p as a 32-bit threshold: draw < thr happens with prob. p.
inline uint32_t prob_threshold(double p) {
  if (p <= 0) return 0;
  if (p >= 1) return 0xFFFFFFFFu;
  return (uint32_t)(p * 4294967296.0);
}

// RMAT sample i: scale levels, quadrant per level from one 32-bit draw
// (two levels per 64-bit draw).  t0 < t1 < t2 are cumulative thresholds of
// (a, a+b, a+b+c).
PSC_HD void rmat_sample(uint64_t seed, uint64_t i, int scale, uint32_t t0,
                        uint32_t t1, uint32_t t2, uint32_t *u, uint32_t *v) {
  uint32_t uu = 0, vv = 0;
  uint64_t r = 0;
  for (int lv = 0; lv < scale; ++lv) {
    uint32_t x;
    if ((lv & 1) == 0) {
      r = draw64(seed, i, (uint64_t)(lv >> 1));
      x = (uint32_t)r;
    } else {
      x = (uint32_t)(r >> 32);
    }
    const uint32_t q = x < t0 ? 0u : x < t1 ? 1u : x < t2 ? 2u : 3u;
    uu = (uu << 1) | (q >> 1);
    vv = (vv << 1) | (q & 1u);
  }
  *u = uu;
  *v = vv;
}

// Bijection of [0, 2^k) (multiply / xorshift rounds), used to scramble
// RMAT vertex ids.
struct IdPerm {
  uint64_t mask, a, b;
  int k;
  PSC_HD IdPerm(int bits, uint64_t seed) : k(bits) {
    mask = (bits >= 64) ? ~0ull : ((1ull << bits) - 1);
    a = mix64(seed ^ 0x5851F42D4C957F2Dull) & mask;
    b = mix64(seed ^ 0x14057B7EF767814Full) & mask;
  }
  PSC_HD uint64_t operator()(uint64_t x) const {
    const int sh = k / 2 + 1;
    x = (x * 0x9E3779B1ull + a) & mask;
    x ^= x >> sh;
    x = (x * 0x85EBCA77ull + b) & mask;
    x ^= x >> sh;
    x = (x * 0xC2B2AE3Dull + a) & mask;
    x ^= x >> sh;
    return x;
  }
};

// Grid pair p: returns 0 both directions, 1 a->b only, 2 b->a only.
PSC_HD int grid_pair_kind(uint64_t seed, uint64_t p, uint32_t thr_both) {
  const uint64_t r = draw64(seed, 0x6772696400000000ull, p);
  if ((uint32_t)(r >> 32) < thr_both) return 0;
  return (r & 1u) ? 2 : 1;
}

// 21-bit integer coordinate of point i along axis ax (0..2).
PSC_HD uint32_t knn_coord(uint64_t seed, uint64_t i, int ax) {
  return (uint32_t)(draw64(seed, 0x6B6E6E0000000000ull + (uint64_t)ax, i) >> 43);
}

}  // namespace psc
