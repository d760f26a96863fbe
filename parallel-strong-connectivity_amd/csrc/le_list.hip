// Least-element lists (LE_LIST::le_list, src/LeList.hpp:127-245,246-328).
//
// For an order of the vertices (the reference's `centers`), the LE-list of
// a target t holds (t, s, d(s, t)) for each source s, taken in that order,
// whose distance along the out-arcs is below the distance from every
// earlier source.  Given the order the lists are unique, so the result is
// the sequential algorithm's (seq_le_list, :76-115) entry for entry.
//
// As in the reference, sources are taken in batches whose size grows by
// beta (:256-275).  A batch is one multi-source BFS over (priority, vertex)
// pairs: sub-round r follows the arcs of the pairs found in sub-round r-1
// and keeps a neighbour w only while r < dist[w], the smallest distance
// from the sources of earlier batches (:152-215); every new pair goes to
// the next frontier and to the list with r as its distance.  After the last
// batch the list is radix-sorted by (target, priority) and each target's
// run keeps the entries whose distance is below every earlier one in the
// run (:282-313).
//
// Where the reference keeps every pair of the batch in a hash set
// (:131-137,199-202), this path keeps per vertex the smallest priority of
// the batch that has reached it (bmin) and adds (p, w) only when an
// atomicMin lowers bmin[w] from above p.  A pair is then added at most once
// (after it, bmin[w] <= p), and a source stops at every vertex that an
// earlier source of its batch has reached at no larger distance.  Nothing
// the lists need is lost: if (x, p, d) is an LE entry, every vertex y on a
// shortest p -> x path has d(p, y) < d(q, y) for each earlier source q (else
// q would reach x as soon as p), so no earlier q of the batch has reached y
// by sub-round d(p, y) and p adds (p, y) exactly then; by induction p adds
// (x, p) at distance d.  Pairs that are added at a larger distance, or
// that do not beat an earlier source, are not LE entries, and an earlier
// source's entry of no larger distance precedes them in the run, so the
// filter drops them.
//
// The out-degree sum of a frontier bounds the pairs its sub-round adds; the
// expansion sums it for the pairs it adds, and the frontier and the list
// grow to fit (the reference's maybe_resize, :126-150,164-176).  Where the
// sum exceeds the buffers as they are, the arcs that pass both tests are
// counted first for a tight bound.  The expansion walks, a wave at a time,
// the concatenated arc lists of 64 frontier pairs (coalesced, balanced
// across lanes); lists over LE_HEAVY arcs are dealt over every wave of the
// grid by a second kernel; the count gives a lane its own list up to LE_LIGHT arcs
// and shares longer lists across the wave.
#include <hip/hip_runtime.h>
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <string>

#include "../../include/psc.h"
#include "device_common.hpp"
#include "engine.hpp"

namespace psc {
namespace {

constexpr uint32_t LE_INF = 0xFFFFFFFFu;
constexpr uint32_t LE_LIGHT = 32;

// device counters of one run
struct LeCtl {
  unsigned long long cnt;     // arcs passing the distance test (this sub-round)
  unsigned long long nf;      // pairs in the current frontier
  unsigned long long nnext;   // pairs added by the current sub-round
  unsigned long long deg;     // out-degree sum of the current frontier
  unsigned long long dnext;   // out-degree sum of the pairs added so far
  unsigned long long nheavy;  // frontier pairs with more than LE_HEAVY arcs
  uint32_t bad;               // input check
  uint32_t pad;
};

__global__ void k_le_fill32(uint32_t *p, uint64_t n, uint32_t v) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

// ctl->bad = 1 unless offsets run from 0 to m without falling, endpoints
// are < n and order is a permutation of 0..n-1 (seen[] starts at zero)
__global__ void k_le_check(const uint64_t *off, const uint32_t *adj, const uint32_t *order,
                           uint64_t n, uint64_t m, uint32_t *seen, LeCtl *ctl) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  const uint64_t i0 = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  for (uint64_t v = i0; v < n; v += stride) {
    if (off[v + 1] < off[v] || off[v + 1] > m) ctl->bad = 1;
    const uint32_t s = order[v];
    if (s >= n || atomicAdd(&seen[s], 1u) != 0) ctl->bad = 1;
  }
  if (i0 == 0 && (off[0] != 0 || off[n] != m)) ctl->bad = 1;
  for (uint64_t j = i0; j < m; j += stride)
    if (adj[j] >= n) ctl->bad = 1;
}

// the batch's sources: priorities [s0, s0 + k) enter the frontier and the
// list at distance 0 (:161-169); bmin[] was reset to LE_INF
__global__ void k_le_seed(const uint32_t *order, uint64_t s0, uint64_t k,
                          uint32_t *bmin, unsigned long long *front,
                          unsigned long long *lkey, uint32_t *ld, uint64_t lbase, uint32_t bw,
                          uint32_t *next_dist, const uint64_t *off, LeCtl *ctl) {
  uint64_t degs = 0;
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < k;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t p = (uint32_t)(s0 + i), v = order[p];
    degs += off[v + 1] - off[v];
    bmin[v] = p;
    front[i] = make_pair(p, v);
    lkey[lbase + i] = ((unsigned long long)v << bw) | p;
    ld[lbase + i] = 0;
    next_dist[v] = 0;
  }
  degs = wave_sum(degs);
  if (lane() == 0 && degs) atomicAdd(&ctl->deg, (unsigned long long)degs);
}

// a batch of k sources begins
__global__ void k_le_begin(LeCtl *ctl, uint64_t k) {
  ctl->nheavy = 0;
  ctl->nf = k;
  ctl->nnext = 0;
  ctl->cnt = 0;
  ctl->deg = 0;
  ctl->dnext = 0;
}

// arcs (p, v) -> w of the frontier's pairs with r < dist[w] and
// bmin[w] > p: the most pairs the sub-round can add (:178-189)
__device__ __forceinline__ bool le_passes(uint32_t p, uint32_t w, uint32_t r,
                                          const uint32_t *__restrict__ dist,
                                          const uint32_t *bmin) {
  // a plain load: a stale (larger) value only costs an extra atomicMin,
  // since bmin only falls within a batch
  return r < dist[w] && bmin[w] > p;
}

__global__ __launch_bounds__(BLOCK) void k_le_count(const unsigned long long *front,
                                                    LeCtl *ctl, const uint64_t *__restrict__ off,
                                                    const uint32_t *__restrict__ adj,
                                                    const uint32_t *__restrict__ dist,
                                                    const uint32_t *bmin, uint32_t r) {
  const uint64_t nf = ctl->nf;
  const uint64_t nwaves = (uint64_t)gridDim.x * (BLOCK / WAVE);
  unsigned long long c = 0;
  for (uint64_t base = (blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) / WAVE * WAVE;
       base < nf; base += nwaves * WAVE) {
    const uint64_t i = base + lane();
    const bool valid = i < nf;
    const unsigned long long pr = valid ? front[i] : 0;
    const uint32_t p = pair_vertex(pr), v = pair_label(pr);
    const uint64_t b = valid ? off[v] : 0, e = valid ? off[v + 1] : 0;
    const bool light = e - b <= LE_LIGHT;
    if (light)
      for (uint64_t j = b; j < e; ++j) c += le_passes(p, adj[j], r, dist, bmin);
    uint64_t hm = ballot(!light);
    while (hm) {
      const int l = __ffsll((unsigned long long)hm) - 1;
      hm &= hm - 1;
      const uint64_t bl = __shfl(b, l), el = __shfl(e, l);
      const uint32_t pl = __shfl(p, l);
      for (uint64_t j = bl + lane(); j < el; j += WAVE)
        c += le_passes(pl, adj[j], r, dist, bmin);
    }
  }
  c = wave_sum(c);
  if (lane() == 0 && c) atomicAdd(&ctl->cnt, c);
}

constexpr uint32_t LE_STAGE = 256;   // new pairs staged per wave before a flush
constexpr uint64_t LE_HEAVY = 4096;  // longer lists are spread over the grid
constexpr uint32_t LE_HSEG = 2048;   // heavy pairs per LDS prefix segment
static_assert(LE_HSEG % BLOCK == 0, "a block builds the segment prefix in equal shares");
constexpr uint64_t LE_GROW = 1ull << 28;   // entries a degree bound may add to the buffers
// The bound above, or PSC_LE_GROW from the environment: a test sets it tiny
// so that every sub-round takes the counted bound (k_le_count).
static uint64_t le_grow_limit() {
  const char *e = std::getenv("PSC_LE_GROW");
  if (e && *e) {
    const unsigned long long v = std::strtoull(e, nullptr, 10);
    if (v > 0) return v;
  }
  return LE_GROW;
}

// the staged pairs of a wave go to the next frontier and to the list at
// lbase + their frontier slot, with one atomic for all of them
__device__ __forceinline__ void le_flush(const unsigned long long *stage, uint32_t sn,
                                         unsigned long long *nxt, unsigned long long *lkey,
                                         uint32_t *ld, uint64_t lbase, uint32_t bw, uint32_t r,
                                         LeCtl *ctl) {
  wave_sync();
  unsigned long long at = 0;
  if (lane() == 0) at = atomicAdd(&ctl->nnext, (unsigned long long)sn);
  at = __shfl(at, 0);
  for (uint32_t q = lane(); q < sn; q += WAVE) {
    const unsigned long long pr = stage[q];
    nxt[at + q] = pr;
    lkey[lbase + at + q] = ((unsigned long long)pair_label(pr) << bw) | pair_vertex(pr);
    ld[lbase + at + q] = r;
  }
  wave_sync();
}

// one arc (po -> w) of a wave step: a pair (p, w) is new when an
// atomicMin lowers bmin[w] from above p (:190-215); new pairs are staged
// in the wave's LDS slice and flushed LE_STAGE at a time.  The whole wave
// calls it together.
__device__ __forceinline__ void le_take(uint32_t po, uint32_t w, bool want, uint32_t r,
                                        const uint32_t *__restrict__ dist, uint32_t *bmin,
                                        const uint64_t *__restrict__ off,
                                        unsigned long long *stage, uint32_t &sn,
                                        uint64_t &degs, unsigned long long *nxt,
                                        unsigned long long *lkey, uint32_t *ld, uint64_t lbase,
                                        uint32_t bw, uint32_t *next_dist, LeCtl *ctl) {
  bool ins = false;
  if (want && le_passes(po, w, r, dist, bmin)) ins = atomicMin(&bmin[w], po) > po;
  const uint64_t m = ballot(ins);
  if (ins) {
    stage[sn + lanes_below(m)] = make_pair(po, w);
    // every writer of this launch writes r; earlier sub-rounds wrote less
    if (next_dist[w] > r) next_dist[w] = r;
    degs += off[w + 1] - off[w];
  }
  sn += (uint32_t)__popcll(m);
  if (sn > LE_STAGE - WAVE) {
    le_flush(stage, sn, nxt, lkey, ld, lbase, bw, r, ctl);
    sn = 0;
  }
}

// A wave takes 64 frontier pairs and walks the concatenation of their arc
// lists, 64 consecutive arcs per step: lane t of a step finds the pair that
// owns arc t by a binary search over the wave's exclusive degree prefix
// (held one entry per lane), so adjacency reads are coalesced and no lane
// idles while another walks a long list.  Pairs with more than LE_HEAVY
// arcs are listed for k_le_heavy instead.  The out-degree sum of the new
// pairs bounds the next sub-round.
__global__ __launch_bounds__(BLOCK) void k_le_expand(
    const unsigned long long *front, unsigned long long *nxt, LeCtl *ctl,
    const uint64_t *__restrict__ off, const uint32_t *__restrict__ adj,
    const uint32_t *__restrict__ dist, uint32_t r, uint32_t *bmin, unsigned long long *lkey,
    uint32_t *ld, uint64_t lbase, uint32_t bw, uint32_t *next_dist, unsigned long long *heavy) {
  __shared__ unsigned long long stage_all[BLOCK / WAVE][LE_STAGE];
  unsigned long long *stage = stage_all[threadIdx.x / WAVE];
  uint32_t sn = 0;   // staged pairs (the same in every lane)
  const uint64_t nf = ctl->nf;
  const uint64_t nwaves = (uint64_t)gridDim.x * (BLOCK / WAVE);
  uint64_t degs = 0;
  for (uint64_t base = (blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) / WAVE * WAVE;
       base < nf; base += nwaves * WAVE) {
    const uint64_t i = base + lane();
    const bool valid = i < nf;
    const unsigned long long pr = valid ? front[i] : 0;
    const uint32_t p = pair_vertex(pr), v = pair_label(pr);
    const uint64_t b = valid ? off[v] : 0, full = valid ? off[v + 1] - b : 0;
    const bool big = full > LE_HEAVY;
    const unsigned long long hat = wave_append64(&ctl->nheavy, big);
    if (big) heavy[hat] = pr;
    const uint64_t len = big ? 0 : full;
    const uint64_t incl = wave_incl_scan(len), excl = incl - len;
    const uint64_t total = __shfl(incl, WAVE - 1);
    for (uint64_t t0 = 0; t0 < total; t0 += WAVE) {
      const uint64_t t = t0 + lane();
      // largest lane o with excl[o] <= t (excl is non-decreasing)
      int o = 0;
#pragma unroll
      for (int step = 32; step >= 1; step >>= 1) {
        const uint64_t ex = __shfl(excl, o + step < WAVE ? o + step : WAVE - 1);
        if (o + step < WAVE && ex <= t) o += step;
      }
      const uint32_t po = __shfl(p, o);
      const uint64_t j = __shfl(b, o) + (t - __shfl(excl, o));
      le_take(po, t < total ? adj[j] : 0, t < total, r, dist, bmin, off, stage, sn, degs, nxt,
              lkey, ld, lbase, bw, next_dist, ctl);
    }
  }
  if (sn) le_flush(stage, sn, nxt, lkey, ld, lbase, bw, r, ctl);
  degs = wave_sum(degs);
  if (lane() == 0 && degs) atomicAdd(&ctl->dnext, (unsigned long long)degs);
}

// The arcs of the listed heavy pairs, spread over every wave of the grid:
// a segment of up to LE_HSEG pairs has its degree prefix built in LDS by
// each block, and wave steps of 64 consecutive arcs are dealt round-robin
// over the grid's waves; a lane finds its arc's pair by binary search in
// the prefix.
__global__ __launch_bounds__(BLOCK) void k_le_heavy(
    const unsigned long long *heavy, unsigned long long *nxt, LeCtl *ctl,
    const uint64_t *__restrict__ off, const uint32_t *__restrict__ adj,
    const uint32_t *__restrict__ dist, uint32_t r, uint32_t *bmin, unsigned long long *lkey,
    uint32_t *ld, uint64_t lbase, uint32_t bw, uint32_t *next_dist) {
  __shared__ unsigned long long stage_all[BLOCK / WAVE][LE_STAGE];
  __shared__ unsigned long long pre[LE_HSEG + 1];
  __shared__ unsigned long long part[BLOCK / WAVE];
  unsigned long long *stage = stage_all[threadIdx.x / WAVE];
  uint32_t sn = 0;
  uint64_t degs = 0;
  const uint64_t nh = ctl->nheavy;
  const uint64_t nwaves = (uint64_t)gridDim.x * (BLOCK / WAVE);
  const uint64_t gw = (blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) / WAVE;
  constexpr uint32_t PER = LE_HSEG / BLOCK;   // prefix entries per thread
  for (uint64_t s0 = 0; s0 < nh; s0 += LE_HSEG) {
    const uint32_t cnt = (uint32_t)(nh - s0 < LE_HSEG ? nh - s0 : LE_HSEG);
    // pre[i] = arcs of the segment's pairs before pair i (block scan)
    uint64_t mine[PER];
    uint64_t sum = 0;
#pragma unroll
    for (uint32_t q = 0; q < PER; ++q) {
      const uint32_t i = threadIdx.x * PER + q;
      uint64_t len = 0;
      if (i < cnt) {
        const uint32_t v = pair_label(heavy[s0 + i]);
        len = off[v + 1] - off[v];
      }
      mine[q] = len;
      sum += len;
    }
    const uint64_t winc = wave_incl_scan(sum);
    if (lane() == WAVE - 1) part[threadIdx.x / WAVE] = winc;
    __syncthreads();
    uint64_t before = winc - sum;
    for (uint32_t wv = 0; wv < threadIdx.x / WAVE; ++wv) before += part[wv];
#pragma unroll
    for (uint32_t q = 0; q < PER; ++q) {
      pre[threadIdx.x * PER + q] = before;
      before += mine[q];
    }
    if (threadIdx.x == BLOCK - 1) pre[LE_HSEG] = before;
    __syncthreads();
    const uint64_t total = pre[LE_HSEG];
    for (uint64_t t0 = gw * WAVE; t0 < total; t0 += nwaves * WAVE) {
      const uint64_t t = t0 + lane();
      const bool want = t < total;
      // largest i < cnt with pre[i] <= t
      uint32_t lo = 0, hi = cnt - 1;
      while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) / 2;
        if (pre[mid] <= t) lo = mid;
        else hi = mid - 1;
      }
      const unsigned long long pr = heavy[s0 + lo];
      const uint32_t v = pair_label(pr);
      const uint64_t j = off[v] + (t - pre[lo]);
      le_take(pair_vertex(pr), want ? adj[j] : 0, want, r, dist, bmin, off, stage, sn, degs,
              nxt, lkey, ld, lbase, bw, next_dist, ctl);
    }
    __syncthreads();   // pre[] is rebuilt for the next segment
  }
  if (sn) le_flush(stage, sn, nxt, lkey, ld, lbase, bw, r, ctl);
  degs = wave_sum(degs);
  if (lane() == 0 && degs) atomicAdd(&ctl->dnext, (unsigned long long)degs);
}

// next sub-round: the added pairs become the frontier, their degree sum
// the bound of its arcs
__global__ void k_le_advance(LeCtl *ctl) {
  ctl->nheavy = 0;
  ctl->nf = ctl->nnext;
  ctl->nnext = 0;
  ctl->cnt = 0;
  ctl->deg = ctl->dnext;
  ctl->dnext = 0;
}

// keep[i] = 1 where the entry's distance is below every earlier entry of
// its target (:293-313); the thread at the head of a target's run walks it
__global__ void k_le_filter(const unsigned long long *key, const uint32_t *d, uint64_t cnt,
                            uint32_t bw, uint32_t *keep) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < cnt;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const unsigned long long t = key[i] >> bw;
    if (i > 0 && (key[i - 1] >> bw) == t) continue;
    uint32_t best = d[i];
    keep[i] = 1;
    for (uint64_t j = i + 1; j < cnt && (key[j] >> bw) == t; ++j) {
      const bool k = d[j] < best;
      keep[j] = k;
      if (k) best = d[j];
    }
  }
}

// kept entries to (target, source vertex, distance) rows (:320-323)
__global__ void k_le_emit(const unsigned long long *key, const uint32_t *d,
                          const uint32_t *keep, const unsigned long long *pos, uint64_t cnt,
                          uint32_t bw, const uint32_t *order, uint32_t *out) {
  const unsigned long long low = (1ull << bw) - 1;
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < cnt;
       i += (uint64_t)gridDim.x * blockDim.x) {
    if (!keep[i]) continue;
    const uint64_t o = pos[i];
    out[3 * o] = (uint32_t)(key[i] >> bw);
    out[3 * o + 1] = order[key[i] & low];
    out[3 * o + 2] = d[i];
  }
}

unsigned le_grid(uint64_t n) {
  return (unsigned)std::max<uint64_t>(1, std::min<uint64_t>((n + BLOCK - 1) / BLOCK, 65536));
}

// device buffer that grows, keeping its first `keep` bytes when asked
struct LeBuf {
  void *p = nullptr;
  size_t cap = 0;
  ~LeBuf() {
    if (p) (void)hipFree(p);
  }
  hipError_t ensure(size_t bytes, size_t keep, hipStream_t st) {
    if (bytes <= cap) return hipSuccess;
    const size_t nb = std::max(bytes, cap + cap / 2);
    void *q = nullptr;
    hipError_t e = hipMalloc(&q, nb);
    if (e != hipSuccess) return e;
    if (keep) {
      e = hipMemcpyAsync(q, p, keep, hipMemcpyDeviceToDevice, st);
      if (e == hipSuccess) e = hipStreamSynchronize(st);
      if (e != hipSuccess) {
        (void)hipFree(q);
        return e;
      }
    }
    if (p) (void)hipFree(p);
    p = q;
    cap = nb;
    return hipSuccess;
  }
  template <class T> T *as() const { return static_cast<T *>(p); }
};

#define LE_HIP(call)                                                      \
  do {                                                                    \
    hipError_t e_ = (call);                                               \
    if (e_ != hipSuccess) {                                               \
      set_error(std::string(#call) + ": " + hipGetErrorString(e_));       \
      return PSC_ERR_HIP;                                                 \
    }                                                                     \
  } while (0)

struct LeRun {
  hipStream_t st = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  LeBuf ctl, dist, next_dist, seen, bmin, heavy, front[2], lkey, ld, skey, sd, keep, pos, temp;
  LeCtl h{};
  ~LeRun() {
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (st) (void)hipStreamDestroy(st);
  }

  int read_ctl() {
    LE_HIP(hipMemcpyAsync(&h, ctl.p, sizeof(LeCtl), hipMemcpyDeviceToHost, st));
    LE_HIP(hipStreamSynchronize(st));
    return PSC_OK;
  }

  int run(const psc_graph *g, const uint32_t *order, double beta, uint32_t **out,
          uint64_t *count, double *ms) {
    const uint64_t n = g->n, m = g->m;
    const uint32_t bw = n > 1 ? 64 - __builtin_clzll(n - 1) : 1;
    LE_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    LE_HIP(hipEventCreate(&e0));
    LE_HIP(hipEventCreate(&e1));
    LE_HIP(ctl.ensure(sizeof(LeCtl), 0, st));
    LE_HIP(dist.ensure(n * 4, 0, st));
    LE_HIP(next_dist.ensure(n * 4, 0, st));
    LE_HIP(seen.ensure(n * 4, 0, st));
    LE_HIP(hipMemsetAsync(ctl.p, 0, sizeof(LeCtl), st));
    k_le_fill32<<<le_grid(n), BLOCK, 0, st>>>(seen.as<uint32_t>(), n, 0);
    k_le_check<<<le_grid(std::max(n, m)), BLOCK, 0, st>>>(g->offset, g->edges, order, n, m,
                                                          seen.as<uint32_t>(), ctl.as<LeCtl>());
    LE_HIP(hipGetLastError());
    PSC_TRY(read_ctl());
    if (h.bad) {
      set_error("psc_le_lists: offsets must run from 0 to m without falling, endpoints "
                "must be < n, and order must be a permutation of 0..n-1");
      return PSC_ERR_INVALID;
    }
    // the timed region: batches, sort, filter, rows
    int dev = 0, ncu = 1;
    LE_HIP(hipGetDevice(&dev));
    LE_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    const unsigned gwave = (unsigned)ncu * 8;
    LE_HIP(hipEventRecord(e0, st));
    k_le_fill32<<<le_grid(n), BLOCK, 0, st>>>(dist.as<uint32_t>(), n, LE_INF);
    k_le_fill32<<<le_grid(n), BLOCK, 0, st>>>(next_dist.as<uint32_t>(), n, LE_INF);
    // the list starts at 8n entries, the frontiers at 2n pairs (the
    // reference sizes them 4n ln n and n ln n, :61-67); they grow on demand
    const uint64_t lcap0 = std::max<uint64_t>(8 * n, 1024);
    LE_HIP(lkey.ensure(lcap0 * 8, 0, st));
    LE_HIP(ld.ensure(lcap0 * 4, 0, st));
    LE_HIP(front[0].ensure(std::max<uint64_t>(2 * n, 1024) * 8, 0, st));
    LE_HIP(front[1].ensure(std::max<uint64_t>(2 * n, 1024) * 8, 0, st));
    LE_HIP(bmin.ensure(n * 4, 0, st));
    uint64_t nlist = 0, finished = 0, step = 1;
    const uint64_t grow = le_grow_limit();
    while (finished < n) {
      // batch [finished, finished + k) (:264-272; the first batch is the
      // single BFS of :248-255)
      if (finished > 0) step = (uint64_t)std::ceil((double)step * beta);
      const uint64_t k = std::min(step, n - finished);
      k_le_fill32<<<le_grid(n), BLOCK, 0, st>>>(bmin.as<uint32_t>(), n, LE_INF);
      LE_HIP(front[0].ensure(k * 8, 0, st));
      LE_HIP(lkey.ensure((nlist + k) * 8, nlist * 8, st));
      LE_HIP(ld.ensure((nlist + k) * 4, nlist * 4, st));
      k_le_begin<<<1, 1, 0, st>>>(ctl.as<LeCtl>(), k);
      k_le_seed<<<le_grid(k), BLOCK, 0, st>>>(
          order, finished, k, bmin.as<uint32_t>(), front[0].as<unsigned long long>(),
          lkey.as<unsigned long long>(), ld.as<uint32_t>(), nlist, bw, next_dist.as<uint32_t>(),
          g->offset, ctl.as<LeCtl>());
      LE_HIP(hipGetLastError());
      nlist += k;
      int cf = 0;
      for (uint32_t r = 1;; ++r) {
        PSC_TRY(read_ctl());
        if (r > 1) nlist += h.nf;   // pairs the previous sub-round added
        if (h.nf == 0 || h.deg == 0) break;
        // the frontier's out-degree sum bounds the pairs this sub-round
        // adds; the buffers grow to it up to LE_GROW entries, past that
        // the arcs that pass the tests are counted for a tight bound
        uint64_t bound = h.deg;
        if (bound > grow || nlist + bound > grow + lkey.cap / 8) {
          k_le_count<<<gwave, BLOCK, 0, st>>>(front[cf].as<unsigned long long>(),
                                              ctl.as<LeCtl>(), g->offset, g->edges,
                                              dist.as<uint32_t>(), bmin.as<uint32_t>(), r);
          LE_HIP(hipGetLastError());
          PSC_TRY(read_ctl());
          bound = h.cnt;
          if (bound == 0) break;
        }
        LE_HIP(front[cf ^ 1].ensure(bound * 8, 0, st));
        LE_HIP(heavy.ensure(std::max<uint64_t>(h.nf, 1) * 8, 0, st));   // at most nf listed
        LE_HIP(lkey.ensure((nlist + bound) * 8, nlist * 8, st));
        LE_HIP(ld.ensure((nlist + bound) * 4, nlist * 4, st));
        k_le_expand<<<gwave, BLOCK, 0, st>>>(
            front[cf].as<unsigned long long>(), front[cf ^ 1].as<unsigned long long>(),
            ctl.as<LeCtl>(), g->offset, g->edges, dist.as<uint32_t>(), r, bmin.as<uint32_t>(),
            lkey.as<unsigned long long>(), ld.as<uint32_t>(), nlist, bw,
            next_dist.as<uint32_t>(), heavy.as<unsigned long long>());
        k_le_heavy<<<gwave, BLOCK, 0, st>>>(
            heavy.as<unsigned long long>(), front[cf ^ 1].as<unsigned long long>(),
            ctl.as<LeCtl>(), g->offset, g->edges, dist.as<uint32_t>(), r, bmin.as<uint32_t>(),
            lkey.as<unsigned long long>(), ld.as<uint32_t>(), nlist, bw,
            next_dist.as<uint32_t>());
        k_le_advance<<<1, 1, 0, st>>>(ctl.as<LeCtl>());
        LE_HIP(hipGetLastError());
        cf ^= 1;
      }
      LE_HIP(hipMemcpyAsync(dist.p, next_dist.p, n * 4, hipMemcpyDeviceToDevice, st));
      finished += k;
    }
    // sort by (target, priority), filter, emit
    LE_HIP(skey.ensure(nlist * 8, 0, st));
    LE_HIP(sd.ensure(nlist * 4, 0, st));
    LE_HIP(keep.ensure(nlist * 4, 0, st));
    LE_HIP(pos.ensure(nlist * 8, 0, st));
    size_t tb1 = 0, tb2 = 0;
    LE_HIP(rocprim::radix_sort_pairs(nullptr, tb1, lkey.as<unsigned long long>(),
                                     skey.as<unsigned long long>(), ld.as<uint32_t>(),
                                     sd.as<uint32_t>(), nlist, 0, 2 * bw, st));
    LE_HIP(rocprim::exclusive_scan(nullptr, tb2, keep.as<uint32_t>(),
                                   pos.as<unsigned long long>(), 0ull, nlist,
                                   rocprim::plus<unsigned long long>(), st));
    LE_HIP(temp.ensure(std::max<size_t>({tb1, tb2, 1}), 0, st));
    size_t tb = temp.cap;
    LE_HIP(rocprim::radix_sort_pairs(temp.p, tb, lkey.as<unsigned long long>(),
                                     skey.as<unsigned long long>(), ld.as<uint32_t>(),
                                     sd.as<uint32_t>(), nlist, 0, 2 * bw, st));
    k_le_filter<<<le_grid(nlist), BLOCK, 0, st>>>(skey.as<unsigned long long>(),
                                                  sd.as<uint32_t>(), nlist, bw,
                                                  keep.as<uint32_t>());
    LE_HIP(hipGetLastError());
    tb = temp.cap;
    LE_HIP(rocprim::exclusive_scan(temp.p, tb, keep.as<uint32_t>(),
                                   pos.as<unsigned long long>(), 0ull, nlist,
                                   rocprim::plus<unsigned long long>(), st));
    unsigned long long last_pos = 0;
    uint32_t last_keep = 0;
    LE_HIP(hipMemcpyAsync(&last_pos, pos.as<unsigned long long>() + (nlist - 1), 8,
                          hipMemcpyDeviceToHost, st));
    LE_HIP(hipMemcpyAsync(&last_keep, keep.as<uint32_t>() + (nlist - 1), 4,
                          hipMemcpyDeviceToHost, st));
    LE_HIP(hipStreamSynchronize(st));
    const uint64_t total = last_pos + last_keep;
    uint32_t *rows = nullptr;
    LE_HIP(hipMalloc(&rows, std::max<uint64_t>(total, 1) * 12));
    k_le_emit<<<le_grid(nlist), BLOCK, 0, st>>>(skey.as<unsigned long long>(), sd.as<uint32_t>(),
                                                keep.as<uint32_t>(),
                                                pos.as<unsigned long long>(), nlist, bw, order,
                                                rows);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipEventRecord(e1, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    float t = 0;
    if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
    if (e != hipSuccess) {
      (void)hipFree(rows);
      set_error(std::string("psc_le_lists: ") + hipGetErrorString(e));
      return PSC_ERR_HIP;
    }
    *out = rows;
    *count = total;
    if (ms) *ms = t;
    return PSC_OK;
  }
};

}  // namespace
}  // namespace psc

extern "C" {

int psc_le_lists_device(const psc_graph *dev_g, const uint32_t *d_order, double beta,
                        uint32_t **d_rows, uint64_t *count, double *ms) {
  psc::GateGuard gate;
  using namespace psc;
  if (!dev_g || !d_rows || !count || (dev_g->n && (!dev_g->offset || !d_order)) ||
      (dev_g->m && !dev_g->edges) || dev_g->n >= 0xFFFFFFFFull || !(beta >= 1.0) ||
      !(beta <= 1e6)) {
    set_error("psc_le_lists_device: bad arguments (beta must be in [1, 1e6])");
    return PSC_ERR_INVALID;
  }
  *d_rows = nullptr;
  *count = 0;
  if (ms) *ms = 0;
  if (dev_g->n == 0) return PSC_OK;
  LeRun run;
  return run.run(dev_g, d_order, beta, d_rows, count, ms);
}

int psc_le_lists(const psc_graph *g, const uint32_t *order, double beta, uint32_t **rows,
                 uint64_t *count) {
  psc::GateGuard gate;
  using namespace psc;
  if (!g || !rows || !count || (g->n && (!g->offset || !order)) || (g->m && !g->edges)) {
    set_error("psc_le_lists: bad arguments");
    return PSC_ERR_INVALID;
  }
  *rows = nullptr;
  *count = 0;
  const uint64_t n = g->n, m = g->m;
  if (n == 0) return PSC_OK;
  LeBuf off, adj, ord;
  LE_HIP(off.ensure((n + 1) * 8, 0, nullptr));
  LE_HIP(adj.ensure(std::max<uint64_t>(m, 1) * 4, 0, nullptr));
  LE_HIP(ord.ensure(n * 4, 0, nullptr));
  LE_HIP(hipMemcpy(off.p, g->offset, (n + 1) * 8, hipMemcpyHostToDevice));
  if (m) LE_HIP(hipMemcpy(adj.p, g->edges, m * 4, hipMemcpyHostToDevice));
  LE_HIP(hipMemcpy(ord.p, order, n * 4, hipMemcpyHostToDevice));
  psc_graph dg{n, m, off.as<uint64_t>(), adj.as<uint32_t>(), nullptr, nullptr};
  uint32_t *d_rows = nullptr;
  uint64_t cnt = 0;
  int rc = psc_le_lists_device(&dg, ord.as<uint32_t>(), beta, &d_rows, &cnt, nullptr);
  if (rc != PSC_OK) return rc;
  uint32_t *h = static_cast<uint32_t *>(std::malloc(std::max<uint64_t>(cnt, 1) * 12));
  if (!h) {
    (void)hipFree(d_rows);
    set_error("psc_le_lists: out of host memory");
    return PSC_ERR_NOMEM;
  }
  hipError_t e = cnt ? hipMemcpy(h, d_rows, cnt * 12, hipMemcpyDeviceToHost) : hipSuccess;
  (void)hipFree(d_rows);
  if (e != hipSuccess) {
    std::free(h);
    set_error(std::string("psc_le_lists: ") + hipGetErrorString(e));
    return PSC_ERR_HIP;
  }
  *rows = h;
  *count = cnt;
  return PSC_OK;
}

void psc_le_free(uint32_t *rows) { std::free(rows); }

int psc_le_free_device(uint32_t *d_rows) {
  psc::GateGuard gate;
  using namespace psc;
  if (d_rows) LE_HIP(hipFree(d_rows));
  return PSC_OK;
}

}  // extern "C"
