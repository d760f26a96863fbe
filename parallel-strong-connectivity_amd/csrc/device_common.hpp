// Wave64 helpers shared by the gfx950 kernels: lane masks, wave scans,
// wave-aggregated appends (one atomic per wave), visited bitmaps and the
// (vertex, label) slot packing of the label table.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace psc {

constexpr int WAVE = 64;
constexpr int BLOCK = 256;
constexpr int WPB = BLOCK / WAVE;

// Internal label layout of src/scc.hpp:22-23: bit 32 marks "SCC found".
constexpr uint64_t TOP_BIT = 1ull << 32;
constexpr uint64_t VAL_MASK = 0xFFFFFFFFull;
// Empty table slot {UINT_N_MAX, UINT_N_MAX} (src/scc.hpp:88-91,395).
constexpr unsigned long long EMPTY_SLOT = ~0ull;

// 32-bit integer mixer.  Placement only; results do not depend on it.
__host__ __device__ __forceinline__ uint32_t hash32(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7feb352du;
  x ^= x >> 15;
  x *= 0x846ca68bu;
  x ^= x >> 16;
  return x;
}

__device__ __forceinline__ int lane() { return (int)__lane_id(); }

__device__ __forceinline__ uint32_t lanes_below(uint64_t m) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                   __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

__device__ __forceinline__ uint64_t ballot(bool p) {
  return (uint64_t)__ballot(p);
}

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <class T>
__device__ __forceinline__ T wave_incl_scan(T x) {
  const int l = lane();
#pragma unroll
  for (int d = 1; d < WAVE; d <<= 1) {
    T t = __shfl_up(x, d);
    if (l >= d) x += t;
  }
  return x;
}

template <class T>
__device__ __forceinline__ T wave_sum(T x) {
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) x += __shfl_xor(x, d);
  return x;
}

// Wave-aggregated append: lanes with `want` get consecutive slots from
// *ctr with one atomic for the whole wave.  Safe under partial masks.
__device__ __forceinline__ uint32_t wave_append(uint32_t *ctr, bool want) {
  uint64_t m = ballot(want);
  uint32_t base = 0;
  if (m) {
    int leader = __ffsll((unsigned long long)m) - 1;
    if (lane() == leader) base = atomicAdd(ctr, (uint32_t)__popcll(m));
    base = __shfl(base, leader);
  }
  return base + lanes_below(m);
}

__device__ __forceinline__ unsigned long long wave_append64(
    unsigned long long *ctr, bool want) {
  uint64_t m = ballot(want);
  unsigned long long base = 0;
  if (m) {
    int leader = __ffsll((unsigned long long)m) - 1;
    if (lane() == leader)
      base = atomicAdd(ctr, (unsigned long long)__popcll(m));
    base = __shfl(base, leader);
  }
  return base + lanes_below(m);
}

// Per-wave output buffer in LDS.  Appends are staged in the buffer and
// written to the global list in bulk: one atomic per flush instead of one
// per 64-lane step, and the global stores of a flush are coalesced.  `n` is
// wave-uniform.  Entries at or beyond `cap` are dropped and reported.
//
// A list whose readers never look past `cap` entries (capped = true) stops
// reserving once a flush of the wave has seen the counter pass `cap` (the
// counter only grows): later entries are only counted, and finish() adds
// their number with one atomic.  Flush reservations are device-scope
// atomics on one counter and serialise; a level that finds millions of
// vertices past the cap would otherwise make one per CAP - WAVE entries for
// nothing.
template <class T, int CAP>
struct WaveOut {
  T *buf;
  uint32_t n;
  uint32_t over;   // entries counted past the cap (capped lists)
  bool capped, past;
  __device__ __forceinline__ void init(T *b, bool cap_only = false) {
    buf = b;
    n = 0;
    over = 0;
    capped = cap_only;
    past = false;
  }
  // returns false if the global list overflowed `cap`
  template <class Ctr>
  __device__ __forceinline__ bool flush(Ctr *ctr, T *dst, unsigned long long cap) {
    if (n == 0) return true;
    if (past) {
      over += n;
      n = 0;
      return false;
    }
    wave_sync();
    Ctr base = 0;
    if (lane() == 0) base = atomicAdd(ctr, (Ctr)n);
    base = __shfl(base, 0);
    past = capped && (unsigned long long)base + n >= cap;
    for (uint32_t i = lane(); i < n; i += WAVE)
      if ((unsigned long long)base + i < cap) dst[base + i] = buf[i];
    const bool ok = (unsigned long long)base + n <= cap;
    n = 0;
    wave_sync();
    return ok;
  }
  // a capped list's entries counted past the cap join the counter
  template <class Ctr>
  __device__ __forceinline__ void finish(Ctr *ctr) {
    if (over && lane() == 0) atomicAdd(ctr, (Ctr)over);
    over = 0;
  }
  template <class Ctr>
  __device__ __forceinline__ bool put(bool want, T v, Ctr *ctr, T *dst,
                                      unsigned long long cap) {
    const uint64_t m = ballot(want);
    if (!m) return true;
    if (want) buf[n + lanes_below(m)] = v;
    n += (uint32_t)__popcll(m);
    if (n > (uint32_t)(CAP - WAVE)) return flush(ctr, dst, cap);
    return true;
  }
};

// Visited bitmap: test, and test-and-set with one atomicOr.
__device__ __forceinline__ bool bit_test(const uint32_t *bm, uint32_t v) {
  return (bm[v >> 5] >> (v & 31)) & 1u;
}
__device__ __forceinline__ bool bit_set(uint32_t *bm, uint32_t v) {
  const uint32_t b = 1u << (v & 31);
  if (bm[v >> 5] & b) return false;
  return (atomicOr(&bm[v >> 5], b) & b) == 0;
}

// Table slot / frontier pair: (vertex << 32) | label.
__device__ __forceinline__ unsigned long long make_pair(uint32_t v,
                                                        uint32_t lab) {
  return ((unsigned long long)v << 32) | lab;
}
__device__ __forceinline__ uint32_t pair_vertex(unsigned long long p) {
  return (uint32_t)(p >> 32);
}
__device__ __forceinline__ uint32_t pair_label(unsigned long long p) {
  return (uint32_t)p;
}

// Slot of a (vertex, label) pair.  The reference hashes the vertex alone
// (src/resizable_table.h:113) because it walks all labels of a vertex along
// one probe chain (iter_kv, :49-97); this path never walks per-vertex
// chains (frontiers and the partition read logs of pairs), so it hashes
// the whole pair and a vertex with many labels does not form one long
// cluster.  Placement does not affect any result.
__device__ __forceinline__ uint64_t pair_slot(unsigned long long kv) {
  kv ^= kv >> 33;
  kv *= 0xFF51AFD7ED558CCDull;
  kv ^= kv >> 33;
  kv *= 0xC4CEB9FE1A85EC53ull;
  kv ^= kv >> 33;
  return kv;
}

// Label-table insert (src/resizable_table.h:214-240): linear probing from
// the pair's slot; one 64-bit CAS claims an empty slot.  Returns 1 if (k,v)
// is new, 0 if present, -1 if the probe limit was hit (table treated as
// full: it is doubled and the search rerun).
__device__ __forceinline__ int table_insert(unsigned long long *T,
                                            uint64_t mask, uint32_t k,
                                            uint32_t v, uint32_t limit,
                                            uint32_t *probes) {
  const unsigned long long kv = make_pair(k, v);
  uint64_t h = pair_slot(kv) & mask;
  for (uint32_t i = 0; i < limit; ++i) {
    unsigned long long s = __atomic_load_n(&T[h], __ATOMIC_RELAXED);
    if (s == kv) {
      *probes += i + 1;
      return 0;
    }
    if (s == EMPTY_SLOT) {
      unsigned long long old = atomicCAS(&T[h], EMPTY_SLOT, kv);
      if (old == EMPTY_SLOT) {
        *probes += i + 1;
        return 1;
      }
      if (old == kv) {
        *probes += i + 1;
        return 0;
      }
    }
    h = (h + 1) & mask;
  }
  *probes += limit;
  return -1;
}

// contains(k, v) (src/resizable_table.h:283-294).
__device__ __forceinline__ bool table_contains(const unsigned long long *T,
                                               uint64_t mask, uint32_t k,
                                               uint32_t v, uint32_t *probes) {
  const unsigned long long kv = make_pair(k, v);
  uint64_t h = pair_slot(kv) & mask;
  for (uint64_t i = 0; i <= mask; ++i) {
    unsigned long long s = T[h];
    if (s == kv) {
      *probes += (uint32_t)i + 1;
      return true;
    }
    if (s == EMPTY_SLOT) {
      *probes += (uint32_t)i + 1;
      return false;
    }
    h = (h + 1) & mask;
  }
  return false;
}

// largest i in [0, cnt) with pre[i] <= x (pre non-decreasing, pre[0] <= x)
__device__ __forceinline__ uint32_t seg_of(const unsigned long long *pre,
                                           uint32_t cnt, uint64_t x) {
  uint32_t lo = 0, hi = cnt - 1;
  while (lo < hi) {
    const uint32_t mid = (lo + hi + 1) >> 1;
    if (pre[mid] <= x) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// Seeded bijection of [0, N) (cycle-walking a multiply/xorshift bijection
// of the enclosing power of two); stands in for parlay::random_shuffle of
// the non-trimmed vertices (src/scc.hpp:303).
struct PivotPerm {
  uint64_t N, mask, a, b;
  int k;
  __host__ __device__ PivotPerm() : N(0), mask(0), a(0), b(0), k(0) {}
  __host__ __device__ PivotPerm(uint64_t n, uint64_t seed) : N(n) {
    k = 0;
    while (k < 63 && (1ull << k) < n) k++;
    mask = (1ull << k) - 1;
    uint64_t s = seed * 0x9E3779B97F4A7C15ull + 0x632BE59BD9B4E019ull;
    a = (s ^ (s >> 31)) & mask;
    s = s * 0xD1B54A32D192ED03ull + 1;
    b = (s ^ (s >> 29)) & mask;
  }
  __host__ __device__ __forceinline__ uint64_t step(uint64_t x) const {
    const int sh = k / 2 + 1;
    x = (x * 0x9E3779B1ull + a) & mask;
    x ^= x >> sh;
    x = (x * 0x85EBCA77ull + b) & mask;
    x ^= x >> sh;
    return x;
  }
  __host__ __device__ __forceinline__ uint64_t operator()(uint64_t i) const {
    uint64_t y = step(i);
    while (y >= N) y = step(y);
    return y;
  }
  // Multiplicative inverse of an odd constant mod 2^64 (Newton iteration).
  static __host__ __device__ constexpr uint64_t inv_odd(uint64_t c) {
    uint64_t x = c;
    for (int i = 0; i < 6; ++i) x *= 2 - c * x;
    return x;
  }
  // step^-1: the xorshift is its own inverse because 2 * sh >= k, the
  // multiply-adds invert with the constants' inverses mod 2^k.
  __host__ __device__ __forceinline__ uint64_t step_inv(uint64_t y) const {
    const int sh = k / 2 + 1;
    y ^= y >> sh;
    y = ((y - b) * inv_odd(0x85EBCA77ull)) & mask;
    y ^= y >> sh;
    y = ((y - a) * inv_odd(0x9E3779B1ull)) & mask;
    return y;
  }
  // Position of y in the permutation: inverse(operator()(i)) == i.
  __host__ __device__ __forceinline__ uint64_t inverse(uint64_t y) const {
    uint64_t x = step_inv(y);
    while (x >= N) x = step_inv(x);
    return x;
  }
};

}  // namespace psc
