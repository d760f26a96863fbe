// MI355X (gfx950) SCC engine: the hot path of src/scc.hpp, src/reach.hpp,
// src/hash_bag.h and src/resizable_table.h, re-designed for wave64 HIP.
//
//   trim + pivot order        src/scc.hpp:284-306   stable device select
//   first SCC (fwd+bwd reach) src/scc.hpp:316-355   fused two-search levels
//     sparse push + VGC       src/reach.hpp:114-171 k_reach_push/k_reach_heavy
//     dense pull              src/reach.hpp:173-190 k_reach_pull
//     direction switch        src/reach.hpp:86-112  k_reach_judge
//   multi-search              src/scc.hpp:144-275   k_ms_seed/k_ms_expand/k_ms_heavy
//   label table               src/resizable_table.h table_insert/table_contains
//   partition                 src/scc.hpp:437-464   k_partition (atomicMax u64)
//   finalize                  src/scc.hpp:475-476   k_finalize
//
// Frontiers are append buffers (wave ballot + prefix, one atomic per wave);
// a multi-search frontier holds (vertex, label) pairs, i.e. exactly the
// table entries created in the previous sub-round, so each pair is
// propagated once.  All compactions are stable, so raw labels are a
// deterministic function of (graph, seed).
#include <hip/hip_runtime.h>

#include <cstring>
#include <rocprim/device/device_scan.hpp>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <condition_variable>
#include <mutex>
#include <vector>

#include "device_common.hpp"
#include "engine.hpp"
#include "scc_kernels.hpp"

namespace psc {

thread_local std::string g_last_error;

void set_error(const std::string &msg) { g_last_error = msg; }

// ------------------------------------------------------------ utilities
__global__ void k_fill_u64(unsigned long long *p, uint64_t n,
                           unsigned long long v) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = v;
}
__global__ void k_fill_u32(uint32_t *p, uint64_t n, uint32_t v) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

// ------------------------------------------------ stable select (3 kernels)
constexpr int SEL_ITEMS = 8;
constexpr int SEL_TILE = BLOCK * SEL_ITEMS;

// A tile is SEL_ITEMS rounds of BLOCK consecutive items; in each round
// thread t handles item round*BLOCK + t, so predicate loads are coalesced.
// Ranks come from wave ballots and a per-round block prefix, which keeps
// the output in index order (stable).
template <class F>
__global__ __launch_bounds__(BLOCK) void k_sel_count(F f, uint64_t len,
                                                     uint32_t *tile_cnt) {
  __shared__ uint32_t wc[WPB];
  const uint64_t base = (uint64_t)blockIdx.x * SEL_TILE;
  uint32_t c = 0;
#pragma unroll
  for (int j = 0; j < SEL_ITEMS; ++j) {
    const uint64_t i = base + (uint64_t)j * BLOCK + threadIdx.x;
    c += (uint32_t)__popcll(ballot(i < len && f.pred(i)));
  }
  if (lane() == 0) wc[threadIdx.x / WAVE] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t t = 0;
    for (int w = 0; w < WPB; ++w) t += wc[w];
    tile_cnt[blockIdx.x] = t;
  }
}

template <class F>
__global__ __launch_bounds__(BLOCK) void k_sel_scatter(
    F f, uint64_t len, const unsigned long long *tile_off, typename F::T *sel,
    typename F::T *rej) {
  __shared__ uint32_t wc[SEL_ITEMS][WPB];
  const uint64_t base = (uint64_t)blockIdx.x * SEL_TILE;
  const int w = threadIdx.x / WAVE;
  uint64_t r0 = tile_off[blockIdx.x];   // selected items before this tile
  bool p[SEL_ITEMS];
  uint32_t rk[SEL_ITEMS];
#pragma unroll
  for (int j = 0; j < SEL_ITEMS; ++j) {
    const uint64_t i = base + (uint64_t)j * BLOCK + threadIdx.x;
    p[j] = i < len && f.pred(i);
    const uint64_t m = ballot(p[j]);
    rk[j] = lanes_below(m);
    if (lane() == 0) wc[j][w] = (uint32_t)__popcll(m);
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < SEL_ITEMS; ++j) {
    uint32_t before = 0, round_tot = 0;
#pragma unroll
    for (int x = 0; x < WPB; ++x) {
      before += x < w ? wc[j][x] : 0u;
      round_tot += wc[j][x];
    }
    const uint64_t i = base + (uint64_t)j * BLOCK + threadIdx.x;
    if (i < len) {
      const uint64_t r = r0 + before + rk[j];   // rank among selected
      if (p[j]) sel[r] = f.val(i);
      else if (rej) rej[i - r] = f.val(i);
    }
    r0 += round_tot;
  }
}

// Exclusive scan of the per-tile counts of a select in one block (counts
// made elsewhere, e.g. by atomics while the items are placed): tile_off
// gets the prefixes, *count_dev the total, and with starts1 (a single
// graph) the graph's starts {0, total}.  For nt <= TSCAN_MAX (64 counts a
// thread); longer count arrays take the library scan.
constexpr int TSCAN_BLOCK = 1024;
constexpr uint64_t TSCAN_MAX = (uint64_t)TSCAN_BLOCK * 64;
__global__ __launch_bounds__(TSCAN_BLOCK) void k_tile_scan(const uint32_t *cnt, uint64_t nt,
                                                            unsigned long long *tile_off,
                                                            unsigned long long *count_dev,
                                                            unsigned long long *starts1) {
  __shared__ unsigned long long ws[TSCAN_BLOCK / WAVE];
  const uint64_t per = (nt + TSCAN_BLOCK - 1) / TSCAN_BLOCK;
  const uint64_t b = threadIdx.x * per, e = min(nt, b + per);
  unsigned long long sum = 0;
  for (uint64_t i = b; i < e; ++i) sum += cnt[i];
  const unsigned long long inc = wave_incl_scan(sum);
  const int w = threadIdx.x / WAVE;
  if (lane() == WAVE - 1) ws[w] = inc;
  __syncthreads();
  unsigned long long run = inc - sum, tot = 0;
  for (int x = 0; x < TSCAN_BLOCK / WAVE; ++x) {
    run += x < w ? ws[x] : 0ull;
    tot += ws[x];
  }
  for (uint64_t i = b; i < e; ++i) {
    tile_off[i] = run;
    run += cnt[i];
  }
  if (threadIdx.x == 0) {
    *count_dev = tot;
    if (starts1) {
      starts1[0] = 0;
      starts1[1] = tot;
    }
  }
}

// Zero two arrays in one launch.
__global__ void k_zero2(uint32_t *a, uint64_t na, uint32_t *b, uint64_t nb) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < na + nb; i += stride) {
    if (i < na) a[i] = 0u;
    else b[i - na] = 0u;
  }
}

// Stable select of a short list in one block of SMALL_BLOCK threads (one
// launch instead of three); each thread issues the predicates of
// SMALL_ITEMS items before ranking them.  With wpre given (the list is the
// concatenation of A windows), fstart[a] receives the number of selected
// items before window a, fstart[A] the total.
constexpr int SMALL_BLOCK = 1024, SMALL_WAVES = SMALL_BLOCK / WAVE, SMALL_ITEMS = 4;
constexpr uint64_t SMALL_SELECT = (uint64_t)SMALL_BLOCK * SMALL_ITEMS * 4;
template <class F>
__global__ __launch_bounds__(SMALL_BLOCK) void k_sel_small(
    F f, uint64_t len, typename F::T *sel, unsigned long long *count_dev,
    const unsigned long long *wpre, uint32_t A, unsigned long long *fstart) {
  __shared__ uint32_t wc[SMALL_ITEMS][SMALL_WAVES];
  const int w = threadIdx.x / WAVE;
  unsigned long long carry = 0;
  for (uint64_t b = 0; b < len; b += (uint64_t)SMALL_BLOCK * SMALL_ITEMS) {
    bool p[SMALL_ITEMS];
    uint64_t m[SMALL_ITEMS];
#pragma unroll
    for (int j = 0; j < SMALL_ITEMS; ++j) {
      const uint64_t t = b + (uint64_t)j * SMALL_BLOCK + threadIdx.x;
      p[j] = t < len && f.pred(t);
    }
#pragma unroll
    for (int j = 0; j < SMALL_ITEMS; ++j) {
      m[j] = ballot(p[j]);
      if (lane() == 0) wc[j][w] = (uint32_t)__popcll(m[j]);
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < SMALL_ITEMS; ++j) {
      uint32_t before = 0, tot = 0;
#pragma unroll
      for (int x = 0; x < SMALL_WAVES; ++x) {
        before += x < w ? wc[j][x] : 0u;
        tot += wc[j][x];
      }
      const uint64_t t = b + (uint64_t)j * SMALL_BLOCK + threadIdx.x;
      const unsigned long long r = carry + before + lanes_below(m[j]);
      if (p[j]) sel[r] = f.val(t);
      if (fstart && t < len) {
        const uint32_t a = seg_of(wpre, A, t);
        if (wpre[a] == t) fstart[a] = r;
      }
      carry += tot;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    *count_dev = carry;
    if (fstart) fstart[A] = carry;
  }
}

// pred: not yet in a found SCC, over the concatenation of per-graph windows
// of the centers list (src/scc.hpp:384-388): window a covers
// centers[wsrc[a], wsrc[a] + wpre[a+1] - wpre[a]).
struct SelWindow {
  using T = uint32_t;
  const uint32_t *centers;
  const unsigned long long *wpre, *wsrc;
  uint32_t A;
  const unsigned long long *labels;
  __device__ uint32_t val(uint64_t t) const {
    const uint32_t a = seg_of(wpre, A, t);
    return centers[wsrc[a] + (t - wpre[a])];
  }
  __device__ bool pred(uint64_t t) const { return !(labels[val(t)] & TOP_BIT); }
};
// Log-scale degree bins of the pull heads' hub selection, 8 per octave:
// degrees below 16 get bins 0..15, larger ones 8 e + (the 3 bits after the
// leading one), e = floor(log2 d) >= 4.
constexpr uint32_t HUB_BINS = 256;
__device__ __forceinline__ uint32_t hub_bin(uint32_t d) {
  if (d < 16) return d;
  const uint32_t e = 31 - __clz(d);
  return 8 * e + ((d >> (e - 3)) & 7);
}

// ------------------------------------------------------------- trimming
// Trim, pass 1 (src/scc.hpp:295-302): per SEL_TILE tile, the number of
// vertices with non-zero in- and out-degree.  The same degree reads give
// the dead-end bitmaps of the first reach: vis_f marks vertices without
// out-edges, vis_b those without in-edges, trim_bm their union (one word
// pair per bitmap, wave and 64 consecutive vertices).  They also give the
// pull heads' degree bins: lgf[v] / lgb[v] = hub_bin of the out- / in-
// degree (saturated to 32 bits) and the histograms of the non-zero ones
// (hist_f / hist_b, zeroed beforehand, one flush per block).  A block takes
// tiles with the grid's stride.
__global__ __launch_bounds__(BLOCK) void k_trim_count(
    const uint64_t *__restrict__ off, const uint64_t *__restrict__ ioff, uint64_t n,
    uint64_t words, uint32_t *tile_cnt, uint32_t *vis_f, uint32_t *vis_b,
    uint32_t *trim_bm, unsigned long long *run_zero, uint8_t *lgf, uint8_t *lgb,
    uint32_t *hist_f, uint32_t *hist_b) {
  __shared__ uint32_t wc[WPB];
  __shared__ uint32_t hf[HUB_BINS], hb[HUB_BINS];
  for (uint32_t k = threadIdx.x; k < HUB_BINS; k += BLOCK) hf[k] = hb[k] = 0;
  // the run's contains() probe counter starts here
  if (blockIdx.x == 0 && threadIdx.x == 0) *run_zero = 0;
  const uint64_t ntiles = (n + SEL_TILE - 1) / SEL_TILE;
  const int l = lane();
  // wave-uniform, so each wave's offsets past its last vertex are scalar loads
  const uint32_t w0 = __builtin_amdgcn_readfirstlane(threadIdx.x / WAVE) * WAVE;
  __syncthreads();
  for (uint64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const uint64_t base = tile * SEL_TILE;
    uint32_t c = 0;
    // every offset load of the tile first, then the tests: a lane loads its
    // vertex's start (vertex n's too, the end of vertex n - 1) and takes its
    // end from the next lane; the wave's last lane's end is a scalar load
    uint64_t o0[SEL_ITEMS], i0[SEL_ITEMS], oe[SEL_ITEMS], ie[SEL_ITEMS];
#pragma unroll
    for (int j = 0; j < SEL_ITEMS; ++j) {
      const uint64_t i = base + (uint64_t)j * BLOCK + threadIdx.x;
      const uint64_t nx = base + (uint64_t)j * BLOCK + w0 + WAVE;
      o0[j] = i0[j] = 0;
      if (i <= n) {
        o0[j] = off[i];
        i0[j] = ioff[i];
      }
      oe[j] = nx <= n ? off[nx] : 0;
      ie[j] = nx <= n ? ioff[nx] : 0;
    }
#pragma unroll
    for (int j = 0; j < SEL_ITEMS; ++j) {
      const uint64_t i = base + (uint64_t)j * BLOCK + threadIdx.x;
      // the shuffles run in every lane (an inactive source lane reads 0)
      const uint64_t on = __shfl_down((unsigned long long)o0[j], 1);
      const uint64_t in = __shfl_down((unsigned long long)i0[j], 1);
      const uint64_t o1 = l == WAVE - 1 ? oe[j] : on, i1 = l == WAVE - 1 ? ie[j] : in;
      const bool zo = i < n && o1 == o0[j];
      const bool zi = i < n && i1 == i0[j];
      c += (uint32_t)__popcll(ballot(i < n && !zo && !zi));
      const uint64_t mf = ballot(zo), mb = ballot(zi), mt = mf | mb;
      const uint64_t wd = (i - l) / 32;   // first word of the wave's 64 vertices
      if (l == 0 && wd < words) {
        vis_f[wd] = (uint32_t)mf;
        vis_b[wd] = (uint32_t)mb;
        trim_bm[wd] = (uint32_t)mt;
      }
      if (l == 32 && wd + 1 < words) {
        vis_f[wd + 1] = (uint32_t)(mf >> 32);
        vis_b[wd + 1] = (uint32_t)(mb >> 32);
        trim_bm[wd + 1] = (uint32_t)(mt >> 32);
      }
      if (i < n) {
        const uint32_t df = (uint32_t)min(o1 - o0[j], (uint64_t)0xFFFFFFFFu);
        const uint32_t db = (uint32_t)min(i1 - i0[j], (uint64_t)0xFFFFFFFFu);
        const uint32_t bf = hub_bin(df), bb = hub_bin(db);
        lgf[i] = (uint8_t)bf;
        lgb[i] = (uint8_t)bb;
        if (df) atomicAdd(&hf[bf], 1u);
        if (db) atomicAdd(&hb[bb], 1u);
      }
    }
    if (l == 0) wc[threadIdx.x / WAVE] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
      uint32_t t = 0;
      for (int w = 0; w < WPB; ++w) t += wc[w];
      tile_cnt[tile] = t;
    }
    __syncthreads();   // wc is written again by the next tile
  }
  for (uint32_t k = threadIdx.x; k < HUB_BINS; k += BLOCK) {
    if (hf[k]) atomicAdd(&hist_f[k], hf[k]);
    if (hb[k]) atomicAdd(&hist_b[k], hb[k]);
  }
}

// A run's first step: both visited bitmaps start as their search's dead-end
// set (dead holds the forward one, then the backward one), and the run's
// contains() probe counter starts at zero.
__global__ void k_run_start(const uint32_t *__restrict__ dead, uint64_t words,
                            uint32_t *vis_f, uint32_t *vis_b,
                            unsigned long long *run_zero) {
  if (blockIdx.x == 0 && threadIdx.x == 0) *run_zero = 0;
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < words;
       i += (uint64_t)gridDim.x * blockDim.x) {
    vis_f[i] = dead[i];
    vis_b[i] = dead[words + i];
  }
}

// Per-graph starts: nstart[g] = non-trimmed vertices before voff[g] (the
// tile prefix plus a recount of the tile's head, one wave per graph) and
// zstart[g] = voff[g] - nstart[g], for g in [0, G].
__global__ void k_graph_starts(const uint64_t *__restrict__ off,
                               const uint64_t *__restrict__ ioff, uint64_t n,
                               const unsigned long long *voff, uint32_t G,
                               const unsigned long long *tile_off,
                               const unsigned long long *total,
                               unsigned long long *nstart,
                               unsigned long long *zstart) {
  const int l = lane();
  for (uint64_t g = (blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) / WAVE; g <= G;
       g += (uint64_t)gridDim.x * blockDim.x / WAVE) {
    const uint64_t v0 = voff[g];
    unsigned long long c;
    if (v0 >= n) {
      c = *total;
    } else {
      const uint64_t t = v0 / SEL_TILE;
      c = tile_off[t];
      for (uint64_t b = t * SEL_TILE; b < v0; b += WAVE) {
        const uint64_t i = b + l;
        const bool keep = i < v0 && off[i + 1] != off[i] && ioff[i + 1] != ioff[i];
        c += (unsigned long long)__popcll(ballot(keep));
      }
    }
    if (l == 0) {
      nstart[g] = c;
      zstart[g] = v0 - c;
    }
  }
}

// ------------------------------------------------------- per-graph helpers
// A handle holds G graphs as one disjoint union: graph g owns the vertex id
// range [voff[g], voff[g+1]).  Per-graph lists (trimmed vertices, pivot
// order, SCC centers, round fronts) are stored graph by graph.

// start[a] = first index i of `list` (grouped by graph, graphs ascending)
// whose vertex belongs to graph gid[a] or a later one; start[A] = len.
__global__ void k_segment_starts(const uint32_t *list, uint64_t len,
                                 const unsigned long long *len_ptr,
                                 const unsigned long long *voff, uint32_t G,
                                 const uint32_t *gid, uint32_t A,
                                 unsigned long long *start) {
  if (len_ptr) len = *len_ptr;
  for (uint64_t a = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; a <= A;
       a += (uint64_t)gridDim.x * blockDim.x) {
    if (a == A) {
      start[a] = len;
      continue;
    }
    const uint32_t g = gid ? gid[a] : (uint32_t)a;
    uint64_t lo = 0, hi = len;
    while (lo < hi) {
      const uint64_t mid = (lo + hi) >> 1;
      if (seg_of(voff, G, list[mid]) < g) lo = mid + 1; else hi = mid;
    }
    start[a] = lo;
  }
}

// ----------------------------------------------------------- reach kernels
// Dead ends are pre-marked visited: in a search that pushes along CSR
// `off`, a vertex with no edges in `off` (a sink of the forward search, a
// source of the backward one) is a trimmed vertex that cannot lead anywhere,
// so it is neither appended nor scanned.  It also never appears as a pull
// neighbour (pull follows the opposite CSR, where it would need an edge in
// `off`), so "visited" keeps meaning "reached" wherever a pull step reads
// it.  First-round labels only concern the non-trimmed vertices
// (src/scc.hpp:341-347), whose reach sets this does not change.  The
// bitmaps start as these dead-end sets (written by k_trim_count);
// k_reach_reset clears the level control of both searches.
__global__ void k_reach_reset(ReachCtl *C, int ns) {
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    for (int s = 0; s < ns; ++s) {
      ReachSlot z{};
      for (int j = 0; j < 3; ++j) C->slot[s][j] = z;
      C->done[s] = 0;
      C->levels[s] = 0;
      for (int c = 0; c < NRC; ++c) C->cnt[s][c] = 0;
    }
  }
}

// Level 0: the first pivot P_g[0] of every graph with non-trimmed vertices
// (src/scc.hpp:317); the graphs are disjoint, so one multi-source reach
// gives each graph's single reach from its own pivot.
// The first pivot of every graph, P_g[0] (src/scc.hpp:303,316): the
// vertex of rank nstart[g] + perm_g(0) among the non-trimmed vertices in
// vertex order, found by one wave per graph without listing them: a 64-way
// search over the trim's tile prefix (tile_off, SEL_TILE vertices a tile)
// picks the tile, a scan of its 64 bitmap words the word, then the bit.
static_assert(SEL_TILE == 32 * WAVE, "one bitmap word per lane and tile");
__global__ void k_reach_sources(ReachSearch *S, ReachCtl *C, int ns,
                                const uint32_t *__restrict__ trim_bm, uint64_t words,
                                const unsigned long long *__restrict__ tile_off,
                                uint64_t ntiles, const unsigned long long *nstart,
                                uint32_t G, uint64_t seed, uint64_t m) {
  const int l = lane();
  // the non-trimmed count (and the mean degree of a frontier vertex) of
  // the direction rule come from the trim on the device
  if (blockIdx.x == 0 && threadIdx.x < (unsigned)ns) {
    const unsigned long long nz = nstart[G];
    S[threadIdx.x].nnz = nz;
    S[threadIdx.x].avg_deg = nz ? (double)m / (double)nz : 0.0;
  }
  for (uint64_t g = (blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) / WAVE; g < G;
       g += (uint64_t)gridDim.x * blockDim.x / WAVE) {
    const uint64_t b = nstart[g];
    if (nstart[g + 1] == b) continue;
    const uint64_t k = b + PivotPerm(nstart[g + 1] - b, seed)(0);   // its rank
    // last tile t with tile_off[t] <= k (tile_off[0] = 0)
    uint64_t lo = 0, hi = ntiles;
    while (hi - lo > 1) {
      const uint64_t step = (hi - lo + WAVE - 1) / WAVE;
      const uint64_t p = lo + (uint64_t)l * step;
      const uint64_t ok = ballot(p < hi && tile_off[p] <= k);   // lanes 0..x
      const uint64_t nlo = lo + (uint64_t)(63 - __clzll((unsigned long long)ok)) * step;
      hi = min(hi, nlo + step);
      lo = nlo;
    }
    const uint64_t w = lo * WAVE + l;
    const uint32_t bits = w < words ? ~trim_bm[w] : 0u;
    const uint32_t cnt = (uint32_t)__popc(bits);
    const uint64_t incl = wave_incl_scan((uint64_t)cnt);
    const uint64_t r = k - tile_off[lo];   // rank inside the tile
    // bits past n read as non-trimmed but rank after every vertex < n
    const bool mine = incl - cnt <= r && r < incl;
    uint32_t src = 0;
    if (mine) {
      uint32_t x = bits;
      for (uint64_t t = incl - cnt; t < r; ++t) x &= x - 1;
      src = (uint32_t)(w * 32 + (uint64_t)(__ffs(x) - 1));
    }
    src = __shfl(src, __ffsll((unsigned long long)ballot(mine)) - 1);
    if (l == 0) {
      for (int s = 0; s < ns; ++s) {
        atomicOr(&S[s].vis[src >> 5], 1u << (src & 31));
        atomicAdd(&C->slot[s][0].cum_vis, 1ull);
        const uint32_t idx = atomicAdd(&C->slot[s][0].qsize, 1u);
        S[s].q[0][idx] = src;
        atomicAdd(&C->slot[s][0].degsum,
                  (unsigned long long)(S[s].off[src + 1] - S[s].off[src]));
      }
    }
  }
}

// ------------------------------------------------- single-pivot rounds
// A round whose windows hold one position each has at most one pivot per
// graph.  Its multi-searches then insert exactly the pairs (v, label_g) of
// the vertices reachable from the pivot inside its subproblem (edges to
// unfinished vertices of the same label, src/scc.hpp:177-219), so the
// round runs as plain fused reaches over visited bitmaps: a vertex is
// included when it carries its graph's pivot label, excluded vertices are
// never claimed (push) nor scanned (pull), and the partition becomes a
// bitmap test per included vertex.  Labels are the same as the multi-
// search's: label_off_g | TOP_BIT in both reaches, label_off_g in one.

// Pivot label of every graph with a pivot this round (plab: EMPTY_SLOT
// elsewhere); resets the included counts.
__global__ void k_rr_plab(const uint32_t *front, const unsigned long long *nfront_ptr,
                          const unsigned long long *fstart, uint32_t A,
                          const uint32_t *wgid, const unsigned long long *labels,
                          unsigned long long *plab, ReachSearch *S) {
  const uint64_t nf = *nfront_ptr;
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < nf;
       i += (uint64_t)gridDim.x * blockDim.x)
    plab[wgid[seg_of(fstart, A, i)]] = labels[front[i]];
  if (blockIdx.x == 0 && threadIdx.x < 2) S[threadIdx.x].nnz = 0;
}

// Remainder positions [0, n_rem) grouped by graph (cstart): vertices with
// their graph's pivot label leave the exclusion bitmap; their count is the
// nnz of both searches.
__global__ __launch_bounds__(BLOCK) void k_rr_include(
    const uint32_t *centers, uint64_t n_rem, const unsigned long long *cstart, uint32_t G,
    const unsigned long long *plab, const unsigned long long *labels, uint32_t *excl,
    ReachSearch *S) {
  __shared__ unsigned long long red[WPB];
  unsigned long long c = 0;
  for (uint64_t j = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; j < n_rem;
       j += (uint64_t)gridDim.x * blockDim.x) {
    const unsigned long long pl = plab[G == 1 ? 0u : seg_of(cstart, G, j)];
    if (pl == EMPTY_SLOT) continue;
    const uint32_t v = centers[j];
    if (labels[v] == pl) {
      atomicAnd(&excl[v >> 5], ~(1u << (v & 31)));
      ++c;
    }
  }
  c = wave_sum(c);
  if (lane() == 0) red[threadIdx.x / WAVE] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long t = 0;
    for (int w = 0; w < WPB; ++w) t += red[w];
    if (t) {
      atomicAdd((unsigned long long *)&S[0].nnz, t);
      atomicAdd((unsigned long long *)&S[1].nnz, t);
    }
  }
}

// Level 0 of a single-pivot round: the round's pivots.
__global__ void k_reach_sources_front(ReachSearch *S, ReachCtl *C, const uint32_t *front,
                                      const unsigned long long *nfront_ptr) {
  const uint64_t nf = *nfront_ptr;
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < nf;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t src = front[i];
    for (int s = 0; s < 2; ++s) {
      atomicOr(&S[s].vis[src >> 5], 1u << (src & 31));
      S[s].q[0][i] = src;
      atomicAdd(&C->slot[s][0].degsum,
                (unsigned long long)(S[s].off[src + 1] - S[s].off[src]));
    }
  }
  if (blockIdx.x == 0 && threadIdx.x < 2) {
    C->slot[threadIdx.x][0].qsize = (uint32_t)nf;
    C->slot[threadIdx.x][0].cum_vis = nf;
  }
}

// Partition of a single-pivot round (src/scc.hpp:437-464 for one label per
// graph): included vertices reached both ways take label_off_g | TOP_BIT,
// one way label_off_g (write_max); the rest keep their labels.
__global__ __launch_bounds__(BLOCK) void k_rr_mark(
    const uint32_t *centers, uint64_t n_rem, const unsigned long long *cstart, uint32_t G,
    const unsigned long long *plab, const uint32_t *excl, const uint32_t *vf,
    const uint32_t *vb, const unsigned long long *lab_off, unsigned long long *labels) {
  for (uint64_t j = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; j < n_rem;
       j += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t g = G == 1 ? 0u : seg_of(cstart, G, j);
    if (plab[g] == EMPTY_SLOT) continue;
    const uint32_t v = centers[j];
    if (bit_test(excl, v)) continue;
    const bool a = bit_test(vf, v), b = bit_test(vb, v);
    if (!(a || b)) continue;
    const unsigned long long ell = (uint32_t)lab_off[g];
    const unsigned long long val = (a && b) ? (ell | TOP_BIT) : ell;
    if (labels[v] < val) labels[v] = val;
  }
}

// label_offset_g += pivots of g in this round (src/scc.hpp:425).
__global__ void k_rr_advance(const unsigned long long *fstart, uint32_t A,
                             const uint32_t *wgid, unsigned long long *lab_off) {
  for (uint32_t a = threadIdx.x; a < A; a += blockDim.x)
    lab_off[wgid[a]] += fstart[a + 1] - fstart[a];
}

// Work split of the fused forward/backward levels by XCD.  Workgroups are
// dispatched round-robin over the 8 XCDs (blockIdx % 8); with two searches,
// XCDs 0-3 run the forward search and XCDs 4-7 the backward one, so each
// XCD's L2 holds one visited bitmap instead of both.  gw / nw / total_waves
// become the wave's index and count within its half.
__device__ __forceinline__ void xcd_split(int ns, int w, int *s_lo, int *s_hi,
                                          uint64_t *gw, uint64_t *nw,
                                          uint32_t *total_waves) {
  if (ns == 2 && (gridDim.x & 7) == 0) {
    const uint32_t xcd = blockIdx.x & 7;
    *s_lo = (int)(xcd >> 2);
    *s_hi = *s_lo + 1;
    const uint64_t hb = (uint64_t)(blockIdx.x >> 3) * 4 + (xcd & 3);
    *gw = hb * WPB + w;
    *nw = (uint64_t)(gridDim.x / 2) * WPB;
    *total_waves = (uint32_t)*nw;
  } else {
    *s_lo = 0;
    *s_hi = ns;
    *gw = blockIdx.x * (uint64_t)WPB + w;
    *nw = (uint64_t)gridDim.x * WPB;
  }
}

// Direction choice of judge() (src/reach.hpp:86-112): sparse iff
// n_front + out_edges < m / dense_div.  out_edges is exact after a push
// level (summed at append time) and estimated as n_front * m / nnz after a
// pull level, where the reference samples (dense_sample, :192-206).  A
// pull level here only visits the unvisited non-trimmed vertices (the
// bitmap walk skips the rest), so a level also pulls when those are few
// enough to cost less than the push would; the direction never changes
// the reach sets.  Every input is final before the level starts.
__device__ __forceinline__ int reach_mode(const ReachSlot *c,
                                          const ReachSearch &R) {
  if (c->deferred) return MODE_PULL;
  const uint32_t nf = c->qsize;
  if (nf == 0) return MODE_NONE;
  if (c->from_pull == PULL_UNSTORED) return MODE_PULL;
  unsigned long long ds = c->degsum;
  if (c->from_pull) ds = (unsigned long long)((double)nf * R.avg_deg);
  const uint64_t push = (uint64_t)nf + ds;
  if (push >= R.thresh) return MODE_PULL;
  const uint64_t unv = R.nnz > c->cum_vis ? R.nnz - c->cum_vis : 0;
  return (unv * PULL_VERTEX_COST + R.n / PULL_WALK_DIV < push) ? MODE_PULL : MODE_PUSH;
}

// A dense pull level (the reference's criterion) right after push levels
// tests heads only and leaves every head miss to the next level, which then
// pulls completely: most of those vertices are found there by their head,
// and the rest would be scanned there again anyway.  Reached sets do not
// depend on the level a vertex is found at; the search cannot end while
// misses are deferred.  Such a level runs in k_reach_step (head_level).
__device__ __forceinline__ bool pull_defers(const ReachSlot *c, const ReachSearch &R) {
  return !c->from_pull && !c->deferred && (uint64_t)c->qsize + c->degsum >= R.thresh;
}

// Frontier entries a pull level stores.  Once a pull level has found
// pull_store_cap vertices, the next level pulls whatever else it finds
// (reach_mode: nf + nf * avg_deg >= thresh, with one to spare for the
// rounding of the degree estimate), and a pull level never reads its
// frontier: the entries past the cap are counted but not stored.  The
// complete pull level after a head-only one stores none if it starts with
// more unvisited vertices than the cap (from_pull = PULL_UNSTORED: the
// next level pulls unless nothing was found): most of them are found
// there, by their heads, so the next level would pull anyway.  Only that
// level: a level forced to pull that way stores its frontier, so a search
// that finds few vertices per level returns to push levels.
__device__ __forceinline__ uint64_t pull_store_cap(const ReachSearch &R) {
  const double t = ceil((double)R.thresh / (1.0 + R.avg_deg)) + 1.0;
  return t < (double)R.n ? (uint64_t)t : (uint64_t)R.n;
}

// K counters summed over the block, one atomic per counter and block.
// Every thread of the block must call it (it synchronises the block).
template <int K>
__device__ __forceinline__ void block_add(unsigned long long *const (&ptr)[K],
                                          const unsigned long long (&val)[K]) {
  __shared__ unsigned long long red[WPB][K];
  const int w = threadIdx.x / WAVE;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const unsigned long long x = wave_sum(val[k]);
    if (lane() == 0) red[w][k] = x;
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < K; ++k) {
    if (threadIdx.x == (unsigned)k) {
      unsigned long long t = 0;
      for (int x = 0; x < WPB; ++x) t += red[x][k];
      if (t) atomicAdd(ptr[k], t);
    }
  }
  __syncthreads();
}

__device__ __forceinline__ void add_cnt(unsigned long long *c, unsigned long long v) {
  v = wave_sum(v);
  if (lane() == 0 && v) atomicAdd(c, v);
}

// Push-side claim of v: outside the searched subproblems it is never
// taken; otherwise the visited bit is set once.
__device__ __forceinline__ bool reach_claim(const ReachSearch &R, uint32_t v) {
  if (R.excl && bit_test(R.excl, v)) return false;
  return bit_set(R.vis, v);
}

// Heavy-list append shared by push and pull levels: vertex u with `deg`
// edges still to scan becomes ceil(deg / CHUNK) work items; one packed
// atomic per wave reserves both the list slots and the chunk numbers.
__device__ __forceinline__ void heavy_append(const ReachSearch &R, ReachSlot *cur,
                                             bool heavy, uint32_t u, uint64_t deg) {
  const uint64_t hm = ballot(heavy);
  if (!hm) return;
  const int l = lane();
  const unsigned long long nch = heavy ? (deg + CHUNK - 1) / CHUNK : 0;
  const unsigned long long cinc = wave_incl_scan(nch);
  const unsigned long long ctot = __shfl(cinc, WAVE - 1);
  unsigned long long old = 0;
  if (l == 0)
    old = atomicAdd(&cur->heavy,
                    ((unsigned long long)__popcll(hm) << H_CNT_SHIFT) | ctot);
  old = __shfl(old, 0);
  if (heavy) {
    const uint64_t idx = (old >> H_CNT_SHIFT) + lanes_below(hm);
    R.hv[idx] = u;
    R.hc[idx] = (old & H_CHUNK_MASK) + cinc - nch;
  }
}

// Head-only pull level of one search (pull_defers): every unvisited
// non-trimmed vertex tests its head (one in-neighbour of large degree, see
// k_pull_heads) and is found if the head is visited; a miss is left to the
// next level, which pulls completely.  That level is forced (`deferred`)
// even without misses, so it never reads this level's frontier: the found
// vertices are only counted (one atomic per block), not stored -- a
// frontier flush is a device-scope atomic on one counter, and a level that
// finds tens of millions of vertices would serialise on tens of thousands
// of them.  A wave walks 2048-vertex chunks (one (visited | skipped) word
// per lane, the next chunk's loaded ahead) and takes the chunk's
// non-trimmed vertices in rank order, 64 consecutive ranks per step, so the
// heads (stored by rank) are one coalesced load; each lane first lists its
// word's non-trimmed vertices at their ranks in an LDS map (offset in the
// chunk, and whether the vertex is tested), so lane k of a step reads its
// vertex from the map, and vertices already visited load nothing.
// HEAD_STEPS steps keep their loads in flight together.  Found bits gather
// in LDS and go out with one atomic per word.  The map's two halves
// (1024 entries each) and `aux` (WAVE words) are the wave's own LDS.  Every
// thread of the block calls it (block-wide counter reduction).
constexpr int HEAD_STEPS = 6;
constexpr int PUSH_WIN = 4;   // push windows whose loads are in flight together
constexpr uint64_t RING_NO_OFF = ~0ull;
static_assert(VGC_MAX_DEG < 16, "a queued degree fits 4 bits");
constexpr int CLAIM_CACHE = 256;            // per-wave claim cache entries
constexpr uint32_t CLAIM_EMPTY = ~0u;       // no vertex has this id
__device__ __forceinline__ uint32_t claim_slot(uint32_t v) {
  return (v * 0x9E3779B1u) >> (32 - 8);
}
static_assert(CLAIM_CACHE == 256, "claim_slot yields 8 bits");

__device__ __forceinline__ void head_level(const ReachSearch &R, ReachSlot *nxt,
                                           uint16_t *map_lo, uint16_t *map_hi,
                                           uint32_t *aux, uint64_t gw, uint64_t nw,
                                           unsigned long long *cnt_s) {
  const int l = lane();
  if (gw == 0 && l == 0) {
    nxt->from_pull = 1;
    nxt->deferred = 1;
  }
  const uint64_t words = ((uint64_t)R.n + 31) / 32;
  const uint64_t chunks = (words + WAVE - 1) / WAVE;
  const uint32_t tail_bits = R.n & 31;
  const bool own_trim = R.tstat == R.trim;
  uint32_t *fw = aux;   // found bits of the chunk's words
  uint32_t n_test = 0, n_vis = 0, n_chunks = 0;   // per wave: < 2^32
  // this lane's (visited | skipped) word of chunk c, and its trimmed word
  auto load = [&](uint64_t c, uint32_t *ts) -> uint32_t {
    const uint64_t wi = c * WAVE + l;
    uint32_t sk = ~0u;
    *ts = ~0u;
    if (c < chunks && wi < words) {
      const uint32_t tr = R.trim[wi];
      *ts = own_trim ? tr : R.tstat[wi];
      sk = R.vis[wi] | tr;
      if (wi == words - 1 && tail_bits) sk |= ~((1u << tail_bits) - 1);
    }
    return sk;
  };
  // the chunk's map: rank in the chunk -> offset in the chunk, bit 15 set
  // if that vertex is tested (2048 entries over two halves)
  auto map_at = [&](uint32_t k) -> uint32_t {
    return k < 1024 ? map_lo[k] : map_hi[k - 1024];
  };
  uint32_t ts_next;
  uint32_t sk_next = load(gw, &ts_next);
  unsigned long long r0_next = gw < chunks ? R.nzoff[gw] : 0ull;
  for (uint64_t c = gw; c < chunks; c += nw) {
    const uint32_t sk = sk_next, ts = ts_next;
    const unsigned long long r0 = r0_next;
    sk_next = load(c + nw, &ts_next);
    r0_next = c + nw < chunks ? R.nzoff[c + nw] : 0ull;
    ++n_chunks;
    if (!ballot(sk != ~0u)) continue;
    const uint32_t nz = (uint32_t)__popc(~ts);
    const uint32_t inc = wave_incl_scan(nz);
    const uint32_t total = __shfl(inc, WAVE - 1);
    wave_sync();   // the previous chunk's readers are done
    {
      // each lane lists its word's non-trimmed vertices at their ranks
      uint32_t at = inc - nz;
      for (uint32_t m = ~ts; m; m &= m - 1) {
        const uint32_t b = (uint32_t)(__ffs(m) - 1);
        const uint16_t e = (uint16_t)((l * 32 + b) | (((sk >> b) & 1u) ? 0u : 0x8000u));
        if (at < 1024) map_lo[at] = e; else map_hi[at - 1024] = e;
        ++at;
      }
    }
    fw[l] = 0;
    wave_sync();
    // pos: the vertex's offset in the chunk, bit 31 set if it is tested;
    // the heads of the next HEAD_STEPS steps are loaded while the visited
    // words of the current ones are (software pipelined)
    uint32_t hd[HEAD_STEPS], pos[HEAD_STEPS];
    auto issue = [&](uint32_t base) {
#pragma unroll
      for (int j = 0; j < HEAD_STEPS; ++j) {
        const uint32_t k = base + j * WAVE + l;
        pos[j] = 0;
        if (k < total) {
          const uint32_t e = map_at(k);
          pos[j] = (e & 2047u) | ((e & 0x8000u) << 16);
        }
        hd[j] = (pos[j] >> 31) ? R.rhead[r0 + k] : 0u;
      }
    };
    issue(0);
    for (uint32_t base = 0; base < total; base += WAVE * HEAD_STEPS) {
      uint32_t hw[HEAD_STEPS], ph[HEAD_STEPS], pp[HEAD_STEPS];
#pragma unroll
      for (int j = 0; j < HEAD_STEPS; ++j) {
        hw[j] = (pos[j] >> 31) ? R.vis[hd[j] >> 5] : 0u;
        ph[j] = hd[j];
        pp[j] = pos[j];
      }
      if (base + WAVE * HEAD_STEPS < total) issue(base + WAVE * HEAD_STEPS);
#pragma unroll
      for (int j = 0; j < HEAD_STEPS; ++j) {
        // v's own bit is clear, so an empty list's head (v) misses
        const bool todo = pp[j] >> 31;
        const uint32_t o = pp[j] & 2047u;
        const bool hit = todo && ((hw[j] >> (ph[j] & 31)) & 1u);
        n_test += todo;
        n_vis += hit;
        if (hit) atomicOr(&fw[o >> 5], 1u << (o & 31));
      }
    }
    wave_sync();
    // an atomic, not a plain store: the bit reaches the other XCDs' waves
    // within this level (a plain store stays in this XCD's L2 until the
    // kernel ends), so later head tests there can hit it
    if (fw[l]) atomicOr(&R.vis[c * WAVE + l], fw[l]);
  }
  const uint32_t wv = wave_sum(n_vis);   // the next level's frontier size
  if (l == 0 && wv) atomicAdd(&nxt->qsize, wv);
  unsigned long long *const cp[4] = {&cnt_s[R_HEAD_TESTS], &cnt_s[R_HEAD_VIS], &nxt->cum_vis,
                                     &cnt_s[R_HEAD_UNITS]};
  const unsigned long long cv[4] = {n_test, n_vis, n_vis, l == 0 ? n_chunks : 0ull};
  block_add<4>(cp, cv);
}

// One BFS level of both searches.  Push levels (src/reach.hpp:114-171):
// wave w takes a contiguous share of the frontier, S seeds at a time, keeps
// a queue of up to REACH_RING vertices in LDS and expands it with
// load-balanced edge scans (lane e of a 64-edge window finds its owner by
// binary search over the wave's inclusive degree scan).  Newly visited
// sparse vertices stay in the queue while the VGC budget (tau per seed)
// lasts and the ring has room; the rest are staged in LDS and spilled to
// the next frontier in bulk.  A local search is a chain of dependent
// steps, so each step keeps few memory round trips: a claim's offsets are
// loaded with its atomic, and a queued vertex keeps them in the ring.  Vertices with more than HEAVY_DEG edges go to
// the heavy list (k_reach_heavy).
// Pull levels (src/reach.hpp:173-190) run in k_reach_pull, except the
// head-only ones (pull_defers), which run here (head_level): their waves
// need little LDS and few registers, so this kernel's occupancy suits them.
__global__ __launch_bounds__(BLOCK, 3) void k_reach_step(
    const ReachSearch *__restrict__ S, ReachCtl *C, int ns, uint32_t L, uint32_t tau,
    int vgc, uint32_t total_waves) {
  __shared__ uint32_t ring_s[WPB][REACH_RING];
  // per ring entry: (first edge << 4) | degree of a vertex queued by a
  // claim, RING_NO_OFF for one taken from the frontier
  __shared__ uint64_t ringo_s[WPB][REACH_RING];
  __shared__ uint32_t out_s[WPB][OUT_CAP];
  __shared__ uint32_t sc_s[WPB][WAVE];
  __shared__ uint64_t st_s[WPB][WAVE];
  __shared__ uint32_t cc_s[WPB][CLAIM_CACHE];   // the wave's recent claims
  const int w = threadIdx.x / WAVE, l = lane();
  uint32_t *ring = ring_s[w];
  uint64_t *ringo = ringo_s[w];
  uint32_t *cc = cc_s[w];
  uint32_t *sc = sc_s[w];
  uint64_t *st = st_s[w];
  const int ci = L % 3, ni = (L + 1) % 3, zi = (L + 2) % 3;
  const int p = L & 1;
  int s_lo, s_hi;
  uint64_t gw, nw;
  xcd_split(ns, w, &s_lo, &s_hi, &gw, &nw, &total_waves);
  if (blockIdx.x == 0 && threadIdx.x < (unsigned)ns) {
    const int s = threadIdx.x;
    ReachSlot z{};
    C->slot[s][zi] = z;
    if (C->slot[s][ci].qsize == 0 && C->slot[s][ci].deferred == 0 && !C->done[s]) {
      C->done[s] = 1;
      C->levels[s] = L;
    }
  }
  WaveOut<uint32_t, OUT_CAP> out;
  out.init(out_s[w]);
  for (int s = s_lo; s < s_hi; ++s) {
    ReachSlot *cur = &C->slot[s][ci];
    ReachSlot *nxt = &C->slot[s][ni];
    const ReachSearch R = S[s];
    const int mode = reach_mode(cur, R);
    const uint32_t nf = cur->qsize;
    // carry the visited count into the next level (once per search)
    if (gw == 0 && l == 0 && cur->cum_vis) atomicAdd(&nxt->cum_vis, cur->cum_vis);
    uint32_t *qn = (p ? R.q[0] : R.q[1]);
    if (mode == MODE_PUSH) {
      for (int i = l; i < CLAIM_CACHE; i += WAVE) cc[i] = CLAIM_EMPTY;
      wave_sync();
      // contiguous share of the frontier, S seeds at a time
      const uint64_t per = (nf + (uint64_t)total_waves - 1) / total_waves;
      const uint32_t sd = (uint32_t)min((uint64_t)64, per);
      uint64_t g = gw * per;
      const uint64_t gend = min((uint64_t)nf, g + per);
      unsigned long long n_edges = 0, n_exp = 0, n_vis = 0, n_app = 0, dsum = 0;
      uint64_t wbud = vgc ? (uint64_t)tau * VGC_WAVE_MULT : 0;   // per launch
      uint32_t nbatch = 0;
      uint32_t sp_t = 1, sp_u = 1;   // the last pass's targets: tested, unvisited
      while (g < gend) {
        const uint32_t cnt = (uint32_t)min((uint64_t)sd, gend - g);
        if ((uint32_t)l < cnt) {
          ring[l] = (p ? R.q[1] : R.q[0])[g + l];
          ringo[l] = RING_NO_OFF;
        }
        g += cnt;
        uint32_t head = 0, tail = cnt;
        uint64_t budget = min((uint64_t)tau * cnt, wbud);
        const uint64_t budget0 = budget;
        uint64_t ebudget = budget * VGC_EDGE_MULT;
        wave_sync();
        while (head < tail) {
          // a wave's serial batches per launch are capped: past the cap its
          // new vertices go to the next frontier, where every wave shares them
          if (++nbatch > (uint32_t)vgc) budget = 0;
          const uint32_t k = min(64u, tail - head);
          const bool valid = (uint32_t)l < k;
          const uint32_t slot = (head + l) & (REACH_RING - 1);
          const uint32_t u = valid ? ring[slot] : 0u;
          const uint64_t ro = valid ? ringo[slot] : 0ull;
          head += k;
          uint64_t b = 0, deg = 0;
          if (valid) {
            if (ro == RING_NO_OFF) {
              b = R.off[u];
              deg = R.off[u + 1] - b;
            } else {
              b = ro >> 4;
              deg = ro & 15u;
            }
            ++n_exp;
          }
          const bool heavy = deg > HEAVY_DEG;
          heavy_append(R, cur, heavy, u, deg);
          if (heavy) deg = 0;
          const uint32_t inc = wave_incl_scan((uint32_t)deg);
          const uint32_t total = __shfl(inc, WAVE - 1);
          sc[l] = inc;
          st[l] = b;
          ebudget = ebudget > total ? ebudget - total : 0;
          if (!ebudget) budget = 0;
          wave_sync();
          // PUSH_WIN 64-edge windows at a time: their edge loads, visited
          // tests, claims and degree loads are each in flight together (a
          // local search is a chain of such steps, so each step's latency
          // counts)
          for (uint32_t base = 0; base < total; base += WAVE * PUSH_WIN) {
            uint32_t v[PUSH_WIN], w[PUSH_WIN];
            bool found[PUSH_WIN];
#pragma unroll
            for (int x = 0; x < PUSH_WIN; ++x) {
              const uint32_t e = base + x * WAVE + l;
              found[x] = e < total;
              v[x] = 0;
              if (found[x]) {
                int lo = 0, hi = (int)k - 1;   // lanes >= k hold no list
                while (lo < hi) {
                  const int mid = (lo + hi) >> 1;
                  if (sc[mid] > e) hi = mid; else lo = mid + 1;
                }
                const uint32_t ex = lo ? sc[lo - 1] : 0u;
                v[x] = R.adj[st[lo] + (e - ex)];
                ++n_edges;
              }
            }
            // claim (reach_claim): outside the searched subproblems never,
            // else the visited bit once.  During a local search whose last
            // pass found at least half of its targets unvisited, every
            // target's offsets are loaded with its visited word (one step
            // less in the chain, a wasted load per visited target); else an
            // unvisited target's after the test.
            const bool spec = budget > 0 && 2 * sp_u >= sp_t;
            uint64_t ob[PUSH_WIN], dv[PUSH_WIN];
            sp_t = sp_u = 0;
#pragma unroll
            for (int x = 0; x < PUSH_WIN; ++x) {
              w[x] = found[x] ? R.vis[v[x] >> 5] : 0u;
              if (found[x] && R.excl) w[x] |= R.excl[v[x] >> 5];
              ob[x] = spec && found[x] ? R.off[v[x]] : 0;
              dv[x] = spec && found[x] ? R.off[v[x] + 1] : 0;
            }
#pragma unroll
            for (int x = 0; x < PUSH_WIN; ++x) {
              sp_t += (uint32_t)__popcll(ballot(found[x]));
              found[x] = found[x] && !((w[x] >> (v[x] & 31)) & 1u);
              sp_u += (uint32_t)__popcll(ballot(found[x]));
              if (!spec) {
                ob[x] = found[x] ? R.off[v[x]] : 0;
                dv[x] = found[x] ? R.off[v[x] + 1] : 0;
              }
            }
            // the wave's claims of this launch by exact vertex id: a target
            // that another lane or an earlier step of the wave took is
            // dropped (an evicted entry only costs a repeated expansion)
#pragma unroll
            for (int x = 0; x < PUSH_WIN; ++x)
              if (found[x]) found[x] = atomicExch(&cc[claim_slot(v[x])], v[x]) != v[x];
#pragma unroll
            for (int x = 0; x < PUSH_WIN; ++x) dv[x] = found[x] ? dv[x] - ob[x] : 0;
#pragma unroll
            for (int x = 0; x < PUSH_WIN; ++x) {
              const uint64_t fm = ballot(found[x]);
              if (!fm) continue;
              // VGC (src/reach.hpp:125-154): sparse new vertices stay in the
              // wave's queue while the budget lasts and the ring has room
              const bool want = found[x] && vgc && dv[x] <= VGC_MAX_DEG;
              const uint64_t lm = ballot(want);
              const uint32_t nwant = (uint32_t)__popcll(lm);
              const uint32_t room = REACH_RING - (tail - head);
              const uint32_t take = (uint32_t)min((uint64_t)min(nwant, room), budget);
              const bool loc = want && lanes_below(lm) < take;
              const uint32_t bit = 1u << (v[x] & 31);
              // a local claim sets the bit without waiting for the old word
              // (two waves may both expand a vertex they raced for); a claim
              // for the next frontier is exact, so its entries stay unique
              if (loc) {
                atomicOr(&R.vis[v[x] >> 5], bit);
                const uint32_t at = (tail + lanes_below(lm)) & (REACH_RING - 1);
                ring[at] = v[x];
                ringo[at] = (ob[x] << 4) | dv[x];
              }
              tail += take;
              budget -= take;
              bool glob = found[x] && !loc;
              if (glob) glob = !(atomicOr(&R.vis[v[x] >> 5], bit) & bit);
              n_vis += loc || glob;
              if (glob) {
                dsum += dv[x];
                ++n_app;
              }
              out.put(glob, v[x], &nxt->qsize, qn, R.n);
            }
          }
          wave_sync();
        }
        wbud -= budget0 - budget;
      }
      out.flush(&nxt->qsize, qn, R.n);
      unsigned long long *const cp[6] = {&nxt->degsum, &C->cnt[s][R_PUSH_EDGES],
                                         &C->cnt[s][R_PUSH_EXP], &C->cnt[s][R_PUSH_VIS],
                                         &nxt->cum_vis, &C->cnt[s][R_PUSH_APP]};
      const unsigned long long cv[6] = {dsum, n_edges, n_exp, n_vis, n_vis, n_app};
      block_add<6>(cp, cv);
    } else if (mode == MODE_PULL && pull_defers(cur, R)) {
      head_level(R, nxt, reinterpret_cast<uint16_t *>(ring),
                 reinterpret_cast<uint16_t *>(out_s[w]), reinterpret_cast<uint32_t *>(st), gw,
                 nw, C->cnt[s]);
    }
  }
}

// ------------------------------------------------------------ pull heads
// Each vertex gets a head: one neighbour in the pull direction of large
// degree in the search's own direction.  A vertex of large degree is
// reached early, so most pull candidates are decided by their head, and the
// heads concentrate on few vertices whose visited words stay cached.  The
// head is the neighbour of largest degree among the graph's top hubs (at
// most HUB_MAX vertices of largest degree) found in the first HUB_SCAN1
// entries of the list; a list without a hub there takes the same maximum
// over entries [HUB_SCAN1, HUB_SCAN); a list without a hub in its first
// HUB_SCAN entries takes the neighbour of largest degree among its first
// HUB_FALLBACK entries (earliest on ties; a trimmed v gets no head, since
// heads are stored by non-trimmed rank).  Hub membership is an LDS hash
// probe, so the only random global reads are the few fallback degrees.
// On a skewed graph most lists meet a hub within HUB_SCAN1 entries, so the
// second scan reads few lists.
#ifndef PSC_HUB_SCAN1
#define PSC_HUB_SCAN1 16
#endif
#ifndef PSC_HUB_FALLBACK
#define PSC_HUB_FALLBACK 4
#endif
constexpr uint32_t HUB_SCAN1 = PSC_HUB_SCAN1;
constexpr uint32_t HUB_SCAN = 64;
constexpr uint32_t HUB_FALLBACK = PSC_HUB_FALLBACK;
constexpr uint32_t HUB_MAX = 2048;
constexpr uint32_t HUB_SLOTS = 2 * HUB_MAX;   // open addressing, load <= 1/2
constexpr uint32_t HUB_EMPTY = 0xFFFFFFFFu;
constexpr int HEAD_BLOCK = 1024;
// layout of the hub scratch (u32): histogram, threshold, keys, degrees
constexpr uint32_t HUB_WORDS = HUB_BINS + 2 + 2 * HUB_SLOTS;

__device__ __forceinline__ uint32_t hub_bin_floor(uint32_t b) {   // smallest d in bin b
  if (b < 16) return b;
  const uint32_t e = b / 8;
  return (8 + (b & 7)) << (e - 3);
}
__device__ __forceinline__ uint32_t hub_hash(uint32_t v) {
  return (v * 0x9E3779B1u) >> (32 - 12);   // log2(HUB_SLOTS) bits
}
static_assert(HUB_SLOTS == 1u << 12, "hub_hash yields 12 bits");

// Both searches are set up in the same launches: blockIdx.y selects the
// search, whose own CSR ranks the heads and whose opposite CSR is pulled.
struct HeadDir {
  const uint64_t *own_off;    // degrees in the search's direction
  const uint64_t *pull_off;   // the pull direction's CSR
  const uint32_t *pull_adj;
  uint8_t *lg;                // n log-scale degree bins (hub_bin, k_trim_count), scratch
  uint32_t *hub;              // HUB_WORDS: histogram, threshold, table
  uint32_t *head;             // n heads (out)
};
struct HeadDirs {
  HeadDir d[2];
  const uint32_t *trim;                // trimmed vertices
  const unsigned long long *nzoff;     // non-trimmed vertices before each tile
};

// One block: the hub threshold T = the floor of the lowest bin such that at
// most HUB_MAX vertices have degree >= T (no vertex qualifies when the top
// bin alone holds more), and an empty hub table.
__global__ void k_hub_select(HeadDirs hd) {
  uint32_t *hub = hd.d[blockIdx.y].hub;
  uint32_t *hist = hub, *thr = hub + HUB_BINS, *keys = hub + HUB_BINS + 2;
  for (uint32_t i = threadIdx.x; i < HUB_SLOTS; i += blockDim.x) keys[i] = HUB_EMPTY;
  if (threadIdx.x == 0) {
    uint32_t acc = 0, t = 0xFFFFFFFFu;
    for (int b = HUB_BINS - 1; b >= 0; --b) {
      if (acc + hist[b] > HUB_MAX) break;
      acc += hist[b];
      if (hist[b]) t = hub_bin_floor((uint32_t)b);
    }
    thr[0] = t;
    thr[1] = acc;
  }
}

// Vertices of degree >= T enter the hub table (key = id, value = degree).
// T is the floor of a bin and bins grow with the degree, so d >= T exactly
// when d's bin is at least T's: the scan reads the one-byte bins, 16 per
// thread, and only the selected vertices' degrees.
__global__ __launch_bounds__(BLOCK) void k_hub_insert(HeadDirs hd, uint64_t n) {
  const HeadDir &D = hd.d[blockIdx.y];
  const uint64_t *__restrict__ off = D.own_off;
  const uint8_t *__restrict__ lg = D.lg;
  uint32_t *hub = D.hub;
  // the degrees >= t fill exactly the selected bins: at most HUB_MAX
  // vertices, so every insert finds an empty slot
  const uint32_t t = hub[HUB_BINS];
  if (hub[HUB_BINS + 1] == 0) return;
  const uint32_t tb = hub_bin(t);   // >= 1: the histogram counts no zero degree
  uint32_t *keys = hub + HUB_BINS + 2, *vals = keys + HUB_SLOTS;
  auto insert = [&](uint64_t v) {
    uint32_t s = hub_hash((uint32_t)v);
    while (atomicCAS(&keys[s], HUB_EMPTY, (uint32_t)v) != HUB_EMPTY) s = (s + 1) & (HUB_SLOTS - 1);
    vals[s] = (uint32_t)min(off[v + 1] - off[v], (uint64_t)0xFFFFFFFFu);
  };
  const uint64_t stride = (uint64_t)gridDim.x * BLOCK, nq = n / 16;
  const uint64_t g = blockIdx.x * (uint64_t)BLOCK + threadIdx.x;
  for (uint64_t q = g; q < nq; q += stride) {
    const uint4 x = reinterpret_cast<const uint4 *>(lg)[q];
    const uint32_t wd[4] = {x.x, x.y, x.z, x.w};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      if (!wd[k]) continue;
#pragma unroll
      for (int b = 0; b < 4; ++b)
        if (((wd[k] >> (8 * b)) & 0xFFu) >= tb) insert(q * 16 + 4 * k + b);
    }
  }
  for (uint64_t v = nq * 16 + g; v < n; v += stride)
    if (lg[v] >= tb) insert(v);
}

// head[v] for every vertex (see the rule above).  `off` / `adj` is the pull
// direction's CSR, `deg` the degrees in the search's own direction (zero
// for a vertex that cannot lie on a path of this search).  A wave takes 64
// consecutive vertices and walks the concatenation of their scanned entries
// in 64-entry windows, HEAD_WIN windows' loads in flight (lane e finds its
// vertex by binary search over the wave's degree scan), so lists of very
// different lengths cost their own entries only.  Hub hits raise an LDS
// maximum per vertex of (hub degree, earliest position, neighbour); the
// vertices left without a hub then walk their first HUB_FALLBACK entries the
// same way, ranked by the neighbours' log-scale degree bins (one byte per
// vertex: a quarter of the degree array's lines, so the random lookups
// mostly hit the caches).
constexpr int HEAD_WIN = 4;

__device__ __forceinline__ unsigned long long head_key(uint32_t d, uint32_t pos, uint32_t u) {
  // degree saturated to 24 bits; ties go to the earlier position
  return ((unsigned long long)min(d, 0xFFFFFFu) << 40) |
         ((unsigned long long)(255u - pos) << 32) | u;
}

__device__ __forceinline__ int head_owner(const uint32_t *sc, uint32_t e) {
  int lo = 0, hi = WAVE - 1;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (sc[mid] > e) hi = mid; else lo = mid + 1;
  }
  return lo;
}

__global__ __launch_bounds__(HEAD_BLOCK) __attribute__((amdgpu_waves_per_eu(8)))
void k_pull_heads(HeadDirs hd, uint64_t n) {
  const HeadDir &D = hd.d[blockIdx.y];
  const uint64_t *__restrict__ off = D.pull_off;
  const uint32_t *__restrict__ adj = D.pull_adj;
  const uint32_t *__restrict__ hub = D.hub;
  uint32_t *head = D.head;
  constexpr int HW = HEAD_BLOCK / WAVE;
  __shared__ uint32_t keys[HUB_SLOTS], vals[HUB_SLOTS];
  __shared__ uint32_t sc_s[HW][WAVE];
  __shared__ uint64_t st_s[HW][WAVE];
  __shared__ unsigned long long best_s[HW][WAVE];
  __shared__ uint32_t fbu_s[HW][WAVE * HUB_FALLBACK];   // first entries of each list
  const uint32_t *gk = hub + HUB_BINS + 2;
  for (uint32_t i = threadIdx.x; i < HUB_SLOTS; i += HEAD_BLOCK) {
    keys[i] = gk[i];
    vals[i] = gk[HUB_SLOTS + i];
  }
  __syncthreads();
  // wave-uniform in a scalar register, so the group's trim words and tile
  // prefix are scalar loads
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x / WAVE), l = lane();
  uint32_t *sc = sc_s[w];
  uint64_t *st = st_s[w];
  unsigned long long *best = best_s[w];
  uint32_t *fbu = fbu_s[w];
  const uint32_t *__restrict__ trim = hd.trim;
  const uint64_t stride = (uint64_t)gridDim.x * HW * WAVE;
  const uint64_t nwords = (n + 31) / 32;
  // a wave's groups of 64 consecutive vertices follow each other with a
  // stride; a group's list starts and its two trim words (scalar loads) are
  // loaded one group ahead (lane 63 also loads the next group's first
  // start, the end of its own list)
  uint64_t v0 = (blockIdx.x * (uint64_t)HW + w) * WAVE;
  uint64_t nb = 0, ne = 0;
  uint32_t nta = ~0u, ntb = ~0u;
  auto load_starts = [&](uint64_t g0) {
    if (g0 >= n) return;
    const uint64_t gw = g0 >> 5;
    nta = trim[gw];
    ntb = gw + 1 < nwords ? trim[gw + 1] : ~0u;
    if (g0 + l <= n) nb = off[g0 + l];
    if (l == WAVE - 1 && g0 + WAVE <= n) ne = off[g0 + WAVE];
  };
  load_starts(v0);
  for (; v0 < n; v0 += stride) {
    const uint64_t v = v0 + l;
    // the tile's trim words before the group's own two (all ones past them,
    // so they count nothing) and the tile's non-trimmed prefix, for the
    // ranks: in flight while the lists are scanned
    const uint64_t tile = v0 / SEL_TILE, gw = v0 >> 5, wl = tile * (SEL_TILE / 32) + l;
    const uint32_t tw = wl < gw ? trim[wl] : ~0u;
    const unsigned long long nz = hd.nzoff[tile];
    const uint32_t tw_a = nta, tw_b = ntb;
    // a vertex without edges in the search's own direction, or without a
    // pull list, is trimmed: it is never a pull candidate
    const uint32_t mine = l < 32 ? tw_a : tw_b;
    const bool cand = v < n && !((mine >> (l & 31)) & 1u);
    // the shuffle runs in every lane: a lane that reads an inactive lane
    // gets 0, not its neighbour's start
    const uint64_t nx = __shfl_down((unsigned long long)nb, 1);
    const uint64_t end = l == WAVE - 1 ? ne : nx;
    const uint32_t d = cand ? (uint32_t)min(end - nb, (uint64_t)HUB_SCAN) : 0u;
    st[l] = nb;
    load_starts(v0 + stride);
    best[l] = 0;
    // hub hits among entries [p0, p0 + c) of each list (c per lane), as
    // 64-entry windows over the concatenated lists (lane e finds its list
    // by binary search over the wave's count scan), HEAD_WIN windows'
    // loads in flight
    auto scan_hubs = [&](uint32_t p0, uint32_t c) {
      const uint32_t inc = wave_incl_scan(c);
      const uint32_t total = __shfl(inc, WAVE - 1);
      wave_sync();
      sc[l] = inc;
      wave_sync();
      for (uint32_t base = 0; base < total; base += HEAD_WIN * WAVE) {
        uint32_t u[HEAD_WIN], pos[HEAD_WIN];
        int own[HEAD_WIN];
#pragma unroll
        for (int j = 0; j < HEAD_WIN; ++j) {
          const uint32_t e = base + j * WAVE + l;
          own[j] = -1;
          u[j] = HUB_EMPTY;
          pos[j] = 0;
          if (e < total) {
            own[j] = head_owner(sc, e);
            pos[j] = e - (own[j] ? sc[own[j] - 1] : 0u);
            u[j] = adj[st[own[j]] + p0 + pos[j]];
          }
        }
#pragma unroll
        for (int j = 0; j < HEAD_WIN; ++j) {
          if (own[j] < 0) continue;
          uint32_t s = hub_hash(u[j]), key;
          while ((key = keys[s]) != HUB_EMPTY && key != u[j]) s = (s + 1) & (HUB_SLOTS - 1);
          if (key == u[j]) atomicMax(&best[own[j]], head_key(vals[s], p0 + pos[j], u[j]));
          // the first pass keeps each list's first entries for the fallback
          if (p0 == 0 && pos[j] < HUB_FALLBACK) fbu[own[j] * HUB_FALLBACK + pos[j]] = u[j];
        }
      }
      wave_sync();
    };
    scan_hubs(0, min(d, HUB_SCAN1));
    // fallback of a vertex without a hub in its first HUB_SCAN1 entries:
    // the largest log-scale degree bin among its first HUB_FALLBACK entries
    // (kept in LDS by the first pass), loaded before the second pass and
    // taken if that pass finds no hub either
    const bool nohub = best[l] == 0;
    const uint32_t f = nohub ? min(d, HUB_FALLBACK) : 0u;
    uint32_t fd[HUB_FALLBACK];
#pragma unroll
    for (uint32_t k = 0; k < HUB_FALLBACK; ++k)
      fd[k] = k < f && HUB_FALLBACK > 1 ? D.lg[fbu[l * HUB_FALLBACK + k]] : 0u;
    const uint32_t d2 = !nohub || d <= HUB_SCAN1 ? 0u : d - HUB_SCAN1;
    if (ballot(d2 > 0)) scan_hubs(HUB_SCAN1, d2);
    unsigned long long hk = best[l];
    if (!hk) {
      // the second pass leaves the kept entries alone
#pragma unroll
      for (uint32_t k = 0; k < HUB_FALLBACK; ++k)
        if (k < f) hk = max(hk, head_key(fd[k], k, fbu[l * HUB_FALLBACK + k]));
    }
    // the non-trimmed rank of v: the tile's prefix, the clear bits of the
    // tile's words before this wave's two, then the lanes below in them
    const uint32_t before = (uint32_t)__popc(~tw);
    const uint32_t below = l < 32 ? (uint32_t)__popc(~tw_a & ((1u << l) - 1u))
                                  : (uint32_t)__popc(~tw_a) +
                                        (uint32_t)__popc(~tw_b & ((1u << (l - 32)) - 1u));
    const uint64_t rank = nz + wave_sum(before) + below;
    if (cand) head[rank] = (uint32_t)hk;
    wave_sync();
  }
}

// Pull level of both searches (src/reach.hpp:173-190): every unvisited,
// non-trimmed vertex scans its in-edges until it meets a visited vertex;
// testing "visited" instead of "in the previous frontier" keeps the
// reachable set identical.  Trimmed vertices are skipped: they never join
// a first-round reach set (src/scc.hpp:341-347) and no pull probe reads
// them as visited.
// The vertex range is walked by bitmap: a wave loads 64 words of
// (visited | trimmed) with one coalesced load (2048 vertices) and lists the
// chunk's candidates (clear bits) in LDS; chunks with few candidates are
// listed together (up to PULL_CHUNKS).  The words of the wave's next
// PULL_PF chunks are in flight while the current ones are listed.
// Phase 1: every candidate tests its head (one in-neighbour of large degree,
// from the per-vertex head array, coalesced), PULL_HEADS x 64 at a time.
// Phase 2: the misses gather in a per-wave LDS list across chunks and are
// taken PULL_PAIRS x 64 at a time (the rest once the wave's range is
// walked), with their loads in flight together: in-offsets, the first
// PULL_PROBE in-edges, their visited words.  Vertices found in phase 1 set
// their bits in LDS copies of their chunks' words, flushed with one atomic
// per word; phase 2 and the drains set theirs directly.  Vertices left
// undecided queue their remaining in-edges in LDS; a full batch of 64
// queued vertices is scanned as load-balanced 64-edge windows.  Remainders
// longer than PULL_HEAVY go to the heavy list (k_reach_heavy) instead.
// Head-only levels (pull_defers) are k_reach_step's (head_level).
constexpr int PULL_PAIRS = 2;
constexpr int PULL_CHUNKS = 8;  // chunks whose candidates one pass may hold
constexpr int PULL_OUT_CAP = 512;   // per-wave staged frontier entries
#ifndef PSC_PULL_HEADS
#define PSC_PULL_HEADS 4
#endif
constexpr int PULL_HEADS = PSC_PULL_HEADS;   // head tests per lane in flight
constexpr int PULL_QCAP = WAVE * (PULL_PAIRS + 1);
// miss list: fewer than one phase-2 batch left over plus one head batch
constexpr int PULL_MCAP = WAVE * (PULL_PAIRS + PULL_HEADS);
#ifndef PSC_PULL_PF
#define PSC_PULL_PF 4
#endif
constexpr int PULL_PF = PSC_PULL_PF;   // chunks whose words are loaded ahead

struct PullQueue {   // per-wave LDS queue of undecided vertices
  uint32_t *v;
  uint64_t *b;
  uint32_t *deg;
  uint32_t n;
};

// Scan the first min(64, q.n) queued vertices' remaining in-edges as
// 64-edge windows; a vertex meeting a visited in-neighbour is marked and
// appended to the next frontier, or, when `resid` is set, each vertex that
// met none is appended to the residual list (counted in *rctr).
template <class Out>
__device__ __forceinline__ void pull_drain(const ReachSearch &R, PullQueue &q,
                                           uint32_t *sc, uint64_t *st, uint32_t *fl,
                                           Out &out, uint32_t *qsize, uint32_t *qn,
                                           uint64_t qcap, bool resid,
                                           unsigned long long *rctr,
                                           unsigned long long &n_edges,
                                           unsigned long long &n_vis) {
  const int l = lane();
  const uint32_t k = min(q.n, (uint32_t)WAVE);
  const bool own = (uint32_t)l < k;
  const uint32_t d = own ? q.deg[l] : 0u;
  const uint64_t b = own ? q.b[l] : 0ull;
  const uint32_t v = own ? q.v[l] : 0u;
  wave_sync();
  const uint32_t inc = wave_incl_scan(d);
  const uint32_t total = __shfl(inc, WAVE - 1);
  sc[l] = inc;
  st[l] = b;
  fl[l] = 0;
  wave_sync();
  for (uint32_t base = 0; base < total; base += WAVE) {
    const uint32_t e = base + l;
    if (e < total) {
      int lo = 0, hi = (int)k - 1;   // lanes >= k hold no list
      while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (sc[mid] > e) hi = mid; else lo = mid + 1;
      }
      const uint32_t ex = lo ? sc[lo - 1] : 0u;
      if (!fl[lo]) {
        ++n_edges;
        if (bit_test(R.vis, R.radj[st[lo] + (e - ex)])) fl[lo] = 1;
      }
    }
    wave_sync();
    if (!ballot(d > 0 && !fl[l])) break;
  }
  bool found = false;
  if (own && d > 0 && fl[l]) {
    const uint32_t bit = 1u << (v & 31);
    found = !(atomicOr(&R.vis[v >> 5], bit) & bit);
  }
  n_vis += found;
  // move the unprocessed tail (q.n > 64) to the front
  const uint32_t rest = q.n - k;
  wave_sync();
  for (uint32_t i = l; i < rest; i += WAVE) {
    q.v[i] = q.v[k + i];
    q.b[i] = q.b[k + i];
    q.deg[i] = q.deg[k + i];
  }
  q.n = rest;
  wave_sync();
  if (resid) out.put(own && d > 0 && !fl[l], v, rctr, qn, (uint64_t)R.n);
  else out.put(found, v, qsize, qn, qcap);
}

__global__ __launch_bounds__(BLOCK) void k_reach_pull(
    const ReachSearch *__restrict__ S, ReachCtl *C, int ns, uint32_t L, uint32_t total_waves) {
  __shared__ uint32_t out_s[WPB][PULL_OUT_CAP];
  __shared__ uint32_t sc_s[WPB][WAVE];
  __shared__ uint64_t st_s[WPB][WAVE];
  __shared__ uint32_t fl_s[WPB][WAVE];
  __shared__ uint32_t qv_s[WPB][PULL_QCAP];
  __shared__ uint64_t qb_s[WPB][PULL_QCAP];
  __shared__ uint32_t qd_s[WPB][PULL_QCAP];
  // listed candidates: chunk slot << 11 | offset in the chunk
  __shared__ uint16_t cand_s[WPB][WAVE * 32];
  __shared__ uint32_t miss_s[WPB][PULL_MCAP];          // phase-2 vertices
  __shared__ uint32_t fw_s[WPB][PULL_CHUNKS * WAVE];   // found bits per listed chunk
  __shared__ uint64_t ck_s[WPB][PULL_CHUNKS];          // the listed chunks
  // per listed chunk: its trimmed words, the non-trimmed vertices before
  // each word inside the chunk, and before the chunk (head ranks)
  __shared__ uint32_t tw_s[WPB][PULL_CHUNKS * WAVE];
  __shared__ uint16_t np_s[WPB][PULL_CHUNKS * WAVE];
  __shared__ unsigned long long nb_s[WPB][PULL_CHUNKS];
  // the group of PULL_PF chunks being listed: skip and trimmed words, and
  // the non-trimmed vertices before each chunk
  __shared__ uint32_t gs_s[WPB][PULL_PF * WAVE];
  __shared__ uint32_t gt_s[WPB][PULL_PF * WAVE];
  __shared__ unsigned long long gn_s[WPB][PULL_PF];
  const int w = threadIdx.x / WAVE, l = lane();
  uint32_t *sc = sc_s[w], *fl = fl_s[w];
  uint64_t *st = st_s[w];
  uint16_t *cand = cand_s[w];
  uint32_t *ml = miss_s[w];
  uint32_t *gs = gs_s[w], *gt = gt_s[w];
  unsigned long long *gn = gn_s[w];
  uint32_t *fw = fw_s[w];
  uint64_t *ck = ck_s[w];
  uint32_t *tw = tw_s[w];
  uint16_t *np = np_s[w];
  unsigned long long *nb = nb_s[w];
  const int ci = L % 3, ni = (L + 1) % 3;
  const int p = L & 1;
  int s_lo, s_hi;
  uint64_t gw, nw;
  xcd_split(ns, w, &s_lo, &s_hi, &gw, &nw, &total_waves);
  WaveOut<uint32_t, PULL_OUT_CAP> out;
  PullQueue q{qv_s[w], qb_s[w], qd_s[w], 0};
  for (int s = s_lo; s < s_hi; ++s) {
    ReachSlot *cur = &C->slot[s][ci];
    ReachSlot *nxt = &C->slot[s][ni];
    const ReachSearch R = S[s];
    // head-only levels run in k_reach_step
    if (reach_mode(cur, R) != MODE_PULL || pull_defers(cur, R)) continue;
    uint32_t *qn = (p ? R.q[0] : R.q[1]);
    const uint64_t qcap = pull_store_cap(R);
    const uint64_t unv = R.nnz > cur->cum_vis ? R.nnz - cur->cum_vis : 0;
    const bool unstored = cur->deferred && unv > qcap;
    // the level after an unstored one pulls the previous level's residual
    // vertices (its frontier buffer) instead of walking the bitmap
    const bool listed = cur->from_pull == PULL_UNSTORED;
    if (gw == 0 && l == 0) nxt->from_pull = unstored ? PULL_UNSTORED : 1;   // first wave
    // an unstored level's staged entries are its residual vertices (every
    // candidate left unvisited or handed to the heavy list), counted in the
    // next slot's degsum; its found vertices are only counted
    unsigned long long *const rctr = &nxt->degsum;
    out.init(out_s[w], !unstored);   // frontier: read only up to pull_store_cap
    const uint64_t words = ((uint64_t)R.n + 31) / 32;
    const uint64_t chunks = (words + WAVE - 1) / WAVE;
    const uint32_t tail_bits = R.n & 31;
    const bool own_trim = R.tstat == R.trim;
    unsigned long long n_edges = 0, n_scan = 0, n_vis = 0, n_chunks = 0;
    uint32_t ncand = 0, nck = 0, nmiss = 0;
    // phase 2 over the first cnt (at most WAVE * PULL_PAIRS) misses of the
    // list: in-offsets, then the first PULL_PROBE in-edges, then their
    // visited words; the rest of the list moves to the front
    auto probe = [&](uint32_t cnt) __attribute__((always_inline)) {
      uint64_t v[PULL_PAIRS], b[PULL_PAIRS], deg[PULL_PAIRS];
      bool todo[PULL_PAIRS], rs[PULL_PAIRS];
#pragma unroll
      for (int j = 0; j < PULL_PAIRS; ++j) {
        const uint32_t idx = j * WAVE + l;
        todo[j] = idx < cnt;
        v[j] = todo[j] ? ml[idx] : 0u;
        b[j] = todo[j] ? R.roff[v[j]] : 0ull;
        deg[j] = todo[j] ? R.roff[v[j] + 1] : 0ull;
      }
      uint32_t nv[PULL_PAIRS][PULL_PROBE];
#pragma unroll
      for (int j = 0; j < PULL_PAIRS; ++j) {
        deg[j] -= b[j];
        n_scan += todo[j];
#pragma unroll
        for (uint32_t t = 0; t < PULL_PROBE; ++t)
          nv[j][t] = t < deg[j] ? R.radj[b[j] + t] : 0u;
      }
      uint32_t wd[PULL_PAIRS][PULL_PROBE];
#pragma unroll
      for (int j = 0; j < PULL_PAIRS; ++j)
#pragma unroll
        for (uint32_t t = 0; t < PULL_PROBE; ++t)
          wd[j][t] = t < deg[j] ? R.vis[nv[j][t] >> 5] : 0u;
#pragma unroll
      for (int j = 0; j < PULL_PAIRS; ++j) {
        bool found = false;
#pragma unroll
        for (uint32_t t = 0; t < PULL_PROBE; ++t)
          found |= t < deg[j] && ((wd[j][t] >> (nv[j][t] & 31)) & 1u);
        const uint64_t k = min(deg[j], (uint64_t)PULL_PROBE);
        n_edges += k;
        const uint64_t rem = found ? 0 : deg[j] - k;
        // long remainders go to the heavy list (k_reach_heavy, chunked),
        // the rest to the wave's queue
        const bool heavy = rem > PULL_HEAVY;
        heavy_append(R, cur, heavy, (uint32_t)v[j], rem);
        rs[j] = todo[j] && (heavy || (!found && rem == 0));
        const bool qd = rem > 0 && !heavy;
        const uint64_t qm = ballot(qd);
        if (qd) {
          const uint32_t at = q.n + lanes_below(qm);
          q.v[at] = (uint32_t)v[j];
          q.b[at] = b[j] + k;
          q.deg[at] = (uint32_t)rem;
        }
        q.n += (uint32_t)__popcll(qm);
        // the wave owns v in this level: no other lane sets its bit
        if (found) atomicOr(&R.vis[v[j] >> 5], 1u << (v[j] & 31));
        n_vis += found;
        if (!unstored) out.put(found, (uint32_t)v[j], &nxt->qsize, qn, qcap);
      }
      if (unstored)
#pragma unroll
        for (int j = 0; j < PULL_PAIRS; ++j) out.put(rs[j], (uint32_t)v[j], rctr, qn, (uint64_t)R.n);
      wave_sync();
      const uint32_t rest = nmiss - cnt;
      for (uint32_t i = l; i < rest; i += WAVE) ml[i] = ml[cnt + i];
      nmiss = rest;
      wave_sync();
      while (q.n >= WAVE)
        pull_drain(R, q, sc, st, fl, out, &nxt->qsize, qn, qcap, unstored, rctr, n_edges,
                   n_vis);
    };
    // Phase 1 over the listed candidates (of up to PULL_CHUNKS chunks), the
    // misses to the miss list, then the found bits of the chunks' words go
    // out, one atomic per word.
    auto process = [&]() __attribute__((always_inline)) {
      // every candidate tests its head, PULL_HEADS per lane at a time (two
      // dependent loads); a full batch of misses goes through phase 2
      for (uint32_t cb = 0; cb < ncand; cb += WAVE * PULL_HEADS) {
        uint32_t e[PULL_HEADS], hd[PULL_HEADS], hw[PULL_HEADS];
        uint64_t vx[PULL_HEADS];
        bool todo[PULL_HEADS];
#pragma unroll
        for (int j = 0; j < PULL_HEADS; ++j) {
          const uint32_t idx = cb + j * WAVE + l;
          todo[j] = idx < ncand;
          e[j] = todo[j] ? cand[idx] : 0u;
          const uint32_t sl = e[j] >> 11, wo = (e[j] & 2047u) >> 5;
          vx[j] = ck[sl] * (uint64_t)(WAVE * 32) + (e[j] & 2047u);
          const uint64_t rank = nb[sl] + np[sl * WAVE + wo] +
                                (uint32_t)__popc(~tw[sl * WAVE + wo] & ((1u << (e[j] & 31)) - 1u));
          hd[j] = todo[j] ? R.rhead[rank] : 0u;
        }
#pragma unroll
        for (int j = 0; j < PULL_HEADS; ++j) hw[j] = todo[j] ? R.vis[hd[j] >> 5] : 0u;
#pragma unroll
        for (int j = 0; j < PULL_HEADS; ++j) {
          // v's own bit is clear, so an empty list's head (v) misses
          const bool hit = todo[j] && ((hw[j] >> (hd[j] & 31)) & 1u);
          const bool miss = todo[j] && !hit;
          n_edges += todo[j];
          const uint64_t mm = ballot(miss);
          if (miss) ml[nmiss + lanes_below(mm)] = (uint32_t)vx[j];
          nmiss += (uint32_t)__popcll(mm);
          // found bits gather in the chunks' LDS words
          if (hit) atomicOr(&fw[(e[j] >> 11) * WAVE + ((e[j] & 2047u) >> 5)], 1u << (e[j] & 31));
          n_vis += hit;
          if (!unstored) out.put(hit, (uint32_t)vx[j], &nxt->qsize, qn, qcap);
        }
        wave_sync();
        while (nmiss >= WAVE * PULL_PAIRS) probe(WAVE * PULL_PAIRS);
      }
      wave_sync();
      for (uint32_t k = 0; k < nck; ++k) {
        // atomics make the bits visible to the other XCDs within the level
        const uint32_t f = fw[k * WAVE + l];
        if (f) atomicOr(&R.vis[ck[k] * WAVE + l], f);
      }
      ncand = 0;
      nck = 0;
      wave_sync();
    };
    // This lane's (visited | skipped) word of chunk c, its trimmed word,
    // and the non-trimmed vertices before the chunk.  The words of the next
    // PULL_PF chunks are loaded before a group of chunks is listed (in a
    // pull level a vertex's bit is set only by its own processing, so an
    // early load reads the same word).  The skip set is the trimmed set in
    // a first-reach level, another set in a single-pivot round.
    auto skip_word = [&](uint64_t c, uint32_t &ts, unsigned long long &nz) -> uint32_t {
      const uint64_t wi = c * WAVE + l;
      uint32_t sk = ~0u;
      ts = ~0u;
      nz = 0;
      if (c < chunks) {
        nz = R.nzoff[c];
        if (wi < words) {
          const uint32_t tr = R.trim[wi];
          ts = own_trim ? tr : R.tstat[wi];
          sk = R.vis[wi] | tr;
          if (wi == words - 1 && tail_bits) sk |= ~((1u << tail_bits) - 1);
        }
      }
      return sk;
    };
    if (listed) {
      // the residual vertices, 64 a wave at a time: those still unvisited
      // go straight to the miss list (their heads missed last level)
      const uint32_t *res = p ? R.q[1] : R.q[0];
      const uint64_t nres = cur->degsum;
      for (uint64_t i0 = gw * WAVE; i0 < nres; i0 += nw * WAVE) {
        const uint64_t i = i0 + l;
        uint32_t v = 0;
        bool c = false;
        if (i < nres) {
          v = res[i];
          c = !bit_test(R.vis, v);
        }
        const uint64_t mm = ballot(c);
        if (c) ml[nmiss + lanes_below(mm)] = v;
        nmiss += (uint32_t)__popcll(mm);
        wave_sync();
        while (nmiss >= WAVE * PULL_PAIRS) probe(WAVE * PULL_PAIRS);
      }
    }
    uint32_t pf_sk[PULL_PF], pf_ts[PULL_PF];
    unsigned long long pf_nz[PULL_PF];
#pragma unroll
    for (int i = 0; i < PULL_PF; ++i)
      pf_sk[i] = skip_word(listed ? chunks : gw + (uint64_t)i * nw, pf_ts[i], pf_nz[i]);
    for (uint64_t c0 = listed ? chunks : gw; c0 < chunks; c0 += (uint64_t)PULL_PF * nw) {
      // this group's words to LDS, the next group's loads issued
#pragma unroll
      for (int i = 0; i < PULL_PF; ++i) {
        gs[i * WAVE + l] = pf_sk[i];
        gt[i * WAVE + l] = pf_ts[i];
        if (l == 0) gn[i] = pf_nz[i];
      }
#pragma unroll
      for (int i = 0; i < PULL_PF; ++i)
        pf_sk[i] = skip_word(c0 + (uint64_t)(PULL_PF + i) * nw, pf_ts[i], pf_nz[i]);
      wave_sync();
      for (int i = 0; i < PULL_PF; ++i) {
        const uint64_t c = c0 + (uint64_t)i * nw;
        if (c >= chunks) break;
        const uint32_t sk = gs[i * WAVE + l], ts = gt[i * WAVE + l];
        ++n_chunks;
        // the chunk's candidates (clear bits), appended to the list in
        // order; chunks with few candidates share one pass of phase 1
        const uint32_t todo_w = ~sk;
        const uint32_t cnt = (uint32_t)__popc(todo_w);
        const uint32_t inc = wave_incl_scan(cnt);
        const uint32_t total = __shfl(inc, WAVE - 1);
        if (total == 0) continue;
        if (ncand + total > WAVE * 32 || nck == PULL_CHUNKS) process();
        wave_sync();
        {
          uint32_t at = ncand + inc - cnt;
          for (uint32_t tt = todo_w; tt; tt &= tt - 1)
            cand[at++] = (uint16_t)((nck << 11) | (l * 32 + (__ffs(tt) - 1)));
        }
        fw[nck * WAVE + l] = 0;
        tw[nck * WAVE + l] = ts;
        {
          const uint32_t nz = (uint32_t)__popc(~ts);
          np[nck * WAVE + l] = (uint16_t)(wave_incl_scan(nz) - nz);
        }
        if (l == 0) {
          ck[nck] = c;
          nb[nck] = gn[i];
        }
        ncand += total;
        ++nck;
        wave_sync();
        if (ncand >= WAVE * PULL_HEADS) process();
      }
      wave_sync();
    }
    if (ncand) process();
    if (nmiss) probe(nmiss);
    while (q.n) pull_drain(R, q, sc, st, fl, out, &nxt->qsize, qn, qcap, unstored, rctr, n_edges,
                   n_vis);
    if (unstored) {
      out.flush(rctr, qn, (uint64_t)R.n);
      const unsigned long long nv = wave_sum(n_vis);   // the next level's frontier size
      if (l == 0 && nv) atomicAdd(&nxt->qsize, (uint32_t)nv);
    } else {
      out.flush(&nxt->qsize, qn, qcap);
      out.finish(&nxt->qsize);
    }
    unsigned long long *const cp[5] = {&C->cnt[s][R_PULL_EDGES], &C->cnt[s][R_PULL_SCANS],
                                       &C->cnt[s][R_PULL_VIS], &nxt->cum_vis,
                                       &C->cnt[s][R_PULL_UNITS]};
    const unsigned long long cv[5] = {n_edges, n_scan, n_vis, n_vis,
                                      l == 0 ? n_chunks : 0ull};
    block_add<5>(cp, cv);
  }
}

// Heavy vertices of a level, CHUNK-edge work items; wave w takes a
// contiguous range of items and walks the owner list forward.
// Push level: edge-parallel expansion (src/reach.hpp:155-167, granularity
// 2048).  Pull level: the in-edge remainder of an unvisited vertex with a
// long in-list; the first chunk to meet a visited in-neighbour sets the
// vertex's bit and appends it, the other chunks of that vertex stop early.
__global__ __launch_bounds__(BLOCK) void k_reach_heavy(const ReachSearch *__restrict__ S,
                                                       ReachCtl *C, int ns,
                                                       uint32_t L,
                                                       uint32_t total_waves) {
  __shared__ uint32_t out_s[WPB][OUT_CAP];
  const int l = lane();
  const int ci = L % 3, ni = (L + 1) % 3;
  const int p = L & 1;
  // XCDs are split between the searches in proportion to their chunk counts
  int s_lo = 0, s_hi = ns;
  uint64_t gw = blockIdx.x * (uint64_t)WPB + threadIdx.x / WAVE;
  if (ns == 2 && (gridDim.x & 7) == 0) {
    const uint64_t c0 = C->slot[0][ci].heavy & H_CHUNK_MASK;
    const uint64_t c1 = C->slot[1][ci].heavy & H_CHUNK_MASK;
    if (c0 && c1) {
      const uint32_t k = (uint32_t)min<uint64_t>(7, max<uint64_t>(1, (8 * c0 + (c0 + c1) / 2) / (c0 + c1)));
      const uint32_t xcd = blockIdx.x & 7;
      const bool second = xcd >= k;
      const uint32_t kx = second ? 8 - k : k;
      s_lo = second ? 1 : 0;
      s_hi = s_lo + 1;
      gw = ((uint64_t)(blockIdx.x >> 3) * kx + (second ? xcd - k : xcd)) * WPB +
           threadIdx.x / WAVE;
      total_waves = (gridDim.x >> 3) * kx * WPB;
    }
  }
  WaveOut<uint32_t, OUT_CAP> out;
  out.init(out_s[threadIdx.x / WAVE]);
  for (int s = s_lo; s < s_hi; ++s) {
    ReachSlot *cur = &C->slot[s][ci];
    ReachSlot *nxt = &C->slot[s][ni];
    const unsigned long long hw = cur->heavy;
    const uint64_t H = hw >> H_CNT_SHIFT, Ct = hw & H_CHUNK_MASK;
    if (H == 0) continue;
    const ReachSearch R = S[s];
    uint32_t *qn = (p ? R.q[0] : R.q[1]);
    const bool pull = reach_mode(cur, R) == MODE_PULL;
    // an unstored pull level's frontier buffer holds its residual list
    const uint64_t qcap = !pull ? R.n : nxt->from_pull == PULL_UNSTORED ? 0 : pull_store_cap(R);
    const uint64_t *off = pull ? R.roff : R.off;
    const uint32_t *adj = pull ? R.radj : R.adj;
    // a pull remainder starts after the PULL_PROBE probed edges
    const uint64_t skip = pull ? PULL_PROBE : 0;
    unsigned long long n_edges = 0, n_vis = 0, n_app = 0, n_items = 0, dsum = 0;
    const uint64_t per = (Ct + total_waves - 1) / total_waves;
    uint64_t c = gw * per;
    const uint64_t cend = min(Ct, c + per);
    uint64_t own = 0;
    if (c < cend) {   // owner of the first item: last idx with hc[idx] <= c
      uint64_t lo = 0, hi = H - 1;
      while (lo < hi) {
        const uint64_t mid = (lo + hi + 1) >> 1;
        if (R.hc[mid] <= c) lo = mid; else hi = mid - 1;
      }
      own = lo;
    }
    for (; c < cend; ++c) {
      while (own + 1 < H && R.hc[own + 1] <= c) ++own;
      if (l == 0) ++n_items;
      const uint32_t u = R.hv[own];
      const uint64_t k = c - R.hc[own];
      const uint64_t eb = off[u] + skip + k * CHUNK;
      const uint64_t ee = min(off[u + 1], eb + CHUNK);
      if (!pull) {
        for (uint64_t j0 = eb; j0 < ee; j0 += WAVE) {
          const uint64_t j = j0 + l;
          bool found = false;
          uint32_t v = 0;
          if (j < ee) {
            v = adj[j];
            found = reach_claim(R, v);
            ++n_edges;
          }
          if (found) {
            dsum += R.off[v + 1] - R.off[v];
            ++n_vis;
            ++n_app;
          }
          out.put(found, v, &nxt->qsize, qn, R.n);
        }
      } else if (!bit_test(R.vis, u)) {
        bool hit = false;
        for (uint64_t j0 = eb; j0 < ee && !hit; j0 += WAVE) {
          const uint64_t j = j0 + l;
          bool h = false;
          if (j < ee) {
            ++n_edges;
            h = bit_test(R.vis, adj[j]);
          }
          hit = ballot(h) != 0;
        }
        bool mine = false;
        if (hit && l == 0) {
          const uint32_t bit = 1u << (u & 31);
          mine = !(atomicOr(&R.vis[u >> 5], bit) & bit);
          if (mine) {
            ++n_vis;
            ++n_app;
          }
        }
        out.put(mine, u, &nxt->qsize, qn, qcap);
      }
    }
    out.flush(&nxt->qsize, qn, qcap);
    unsigned long long *const cp[6] = {&nxt->degsum, &C->cnt[s][R_HEAVY_EDGES],
                                       &C->cnt[s][R_HEAVY_VIS], &nxt->cum_vis,
                                       &C->cnt[s][R_HEAVY_APP], &C->cnt[s][R_HEAVY_ITEMS]};
    const unsigned long long cv[6] = {dsum, n_edges, n_vis, n_vis, n_app, n_items};
    block_add<6>(cp, cv);
  }
}

// Every label, once, in vertex order (whole lines), after the first reach:
// trimmed vertices get 1 + (rank among their graph's trimmed | TOP_BIT)
// (src/scc.hpp:304-305); a non-trimmed vertex reached both ways is in the
// first SCC (label | TOP_BIT), one way gets the new subproblem label, none
// keeps 0 (src/scc.hpp:340-349).  A non-trimmed vertex not in the first
// SCC remains: it is placed at its position in its graph's pivot order
// P_g = NZ_g permuted (src/scc.hpp:303), pos = nstart[g] +
// perm_g^-1(j - nstart[g]) with j its rank among the non-trimmed vertices
// (NZ index): posv[pos] = v and bit pos of pbits set, so compacting pbits
// in order yields the remaining vertices in P order without materialising
// or sorting P; each placement also counts in its SEL_TILE tile of
// positions (pos_tiles), so the compaction needs no counting pass.  Ranks
// come from the trim's tile prefix and ballots.
__global__ __launch_bounds__(BLOCK) void k_first_label(
    const uint32_t *__restrict__ trim_bm, uint64_t n, const unsigned long long *tile_off,
    const unsigned long long *voff, uint32_t G, const unsigned long long *zstart,
    const unsigned long long *nstart, uint64_t seed, const uint32_t *vf,
    const uint32_t *vb, unsigned long long *labels, const unsigned long long *lab1,
    uint32_t *posv, uint32_t *pbits, int direct, uint32_t *first_trim,
    uint32_t *pos_tiles, const ReachCtl *gate, uint32_t gate_level,
    unsigned long long lab_v, unsigned long long next_v, unsigned long long *lab_out,
    unsigned long long *next_out) {
  // enqueued ahead of the poll that shows the reach over: if it is not
  // over at gate_level, the pass is enqueued again later and this one
  // does nothing
  if (gate) {
    const int gi = (int)(gate_level % 3);
    for (int s = 0; s < 2; ++s)
      if (!gate->done[s] && (gate->slot[s][gi].qsize != 0 || gate->slot[s][gi].deferred != 0))
        return;
  }
  // each wave labels its own SEG = SEL_TILE / WPB consecutive vertices of
  // the tile, 64 at a time, with no barrier: its first rank is the tile's
  // prefix plus the non-trimmed vertices of the waves before it (their
  // trimmed words, one per lane)
  constexpr uint32_t SEG = SEL_TILE / WPB;
  static_assert(SEG == SEL_ITEMS * WAVE && SEG / 32 * (WPB - 1) <= WAVE, "segment layout");
  const int w = threadIdx.x / WAVE, l = lane();
  const uint64_t base = (uint64_t)blockIdx.x * SEL_TILE + (uint64_t)w * SEG;
  // every load the segment needs first: the tile's prefix, the words of
  // the earlier segments, the per-graph constants of a single graph, and
  // the bitmap words (trimmed, both reaches)
  uint64_t r0 = tile_off[blockIdx.x];
  uint32_t pw = ~0u;
  {
    const uint64_t wd = (uint64_t)blockIdx.x * (SEL_TILE / 32) + l;
    if ((uint32_t)l < (uint32_t)w * (SEG / 32) && wd * 32 < n) {
      pw = trim_bm[wd];
      const uint64_t hi = n - wd * 32;   // positions of the word inside the graph
      if (hi < 32) pw |= ~((1u << hi) - 1u);
    }
  }
  // a single graph's label offsets come by value (lab1 null) and are
  // stored here for the later kernels
  const unsigned long long z0 = zstart[0], lab_one = lab1 ? lab1[0] : lab_v;
  if (!lab1 && blockIdx.x == 0 && threadIdx.x == 0) {
    lab_out[0] = lab_v;
    next_out[0] = next_v;
  }
  uint32_t tw[SEL_ITEMS], fw[SEL_ITEMS], bw[SEL_ITEMS];
#pragma unroll
  for (int j = 0; j < SEL_ITEMS; ++j) {
    const uint64_t i = base + (uint64_t)j * WAVE + l;
    tw[j] = ~0u;
    fw[j] = bw[j] = 0;
    if (i < n) {
      tw[j] = trim_bm[i >> 5];
      fw[j] = vf[i >> 5];
      bw[j] = vb[i >> 5];
    }
  }
  r0 += wave_sum((uint32_t)__popc(~pw));
  uint32_t place = 0;            // items left to place by pivot position
  uint64_t rr[SEL_ITEMS];
  bool p[SEL_ITEMS];
#pragma unroll
  for (int j = 0; j < SEL_ITEMS; ++j) {
    const uint64_t i = base + (uint64_t)j * WAVE + l;
    p[j] = i < n && !((tw[j] >> (i & 31)) & 1u);
    const uint64_t m = ballot(p[j]);
    const uint64_t r = r0 + lanes_below(m);
    rr[j] = r;
    r0 += (uint32_t)__popcll(m);
    if (i >= n) continue;
    const uint32_t g = G == 1 ? 0u : seg_of(voff, G, i);
    if (!p[j]) {
      const unsigned long long zs = G == 1 ? z0 : zstart[g];
      const unsigned long long rank = i - r - zs;
      if (direct && rank) {
        labels[i] = rank;   // already final: (1 + rank) - 1
      } else {
        labels[i] = 1 + (rank | TOP_BIT);
        if (direct) first_trim[g] = (uint32_t)i;   // the one final value 0
      }
      continue;
    }
    const bool a = (fw[j] >> (i & 31)) & 1u, b = (bw[j] >> (i & 31)) & 1u;
    const unsigned long long label = G == 1 ? lab_one : lab1[g];
    const unsigned long long done = direct ? label - 1 : (label | TOP_BIT);
    labels[i] = (a && b) ? done : ((a || b) ? label : 0ull);
    if (!(a && b)) place |= 1u << j;
  }
  // placement of the few remaining vertices, out of the unrolled loop
#pragma unroll 1
  for (uint32_t mm = place; mm; mm &= mm - 1) {
    const int j = __ffs(mm) - 1;
    const uint64_t i = base + (uint64_t)j * WAVE + l;
    const uint32_t g = G == 1 ? 0u : seg_of(voff, G, i);
    const uint64_t b0 = nstart[g];
    const uint64_t pos = b0 + PivotPerm(nstart[g + 1] - b0, seed).inverse(rr[j] - b0);
    posv[pos] = (uint32_t)i;
    atomicOr(&pbits[pos >> 5], 1u << (pos & 31));
    atomicAdd(&pos_tiles[pos / SEL_TILE], 1u);   // the compaction's tile counts
  }
}

// Remaining vertices in pivot order: the set bits of pbits, in order.
// The pivot-order compaction: sel[tile_off[t] + k] = posv of the k-th set
// bit of position tile t.  One wave per tile, a bitmap word per lane (the
// bits are sparse: only the vertices left after the first SCC are set).
__global__ __launch_bounds__(BLOCK) void k_sel_bits(const uint32_t *__restrict__ pbits,
                                                    const uint32_t *__restrict__ posv,
                                                    uint64_t len,
                                                    const unsigned long long *tile_off,
                                                    uint32_t *sel) {
  static_assert(SEL_TILE == 32 * WAVE, "one bitmap word per lane and tile");
  const uint64_t t = blockIdx.x * (uint64_t)WPB + threadIdx.x / WAVE;
  if (t * SEL_TILE >= len) return;
  const int l = lane();
  const uint64_t wd = t * WAVE + l;
  uint32_t bits = 0;
  if (wd * 32 < len) {
    bits = pbits[wd];
    const uint64_t hi = len - wd * 32;   // positions of the word below len
    if (hi < 32) bits &= (1u << hi) - 1u;
  }
  const uint32_t c = (uint32_t)__popc(bits);
  uint64_t r = tile_off[t] + (wave_incl_scan(c) - c);
  for (uint32_t m = bits; m; m &= m - 1) sel[r++] = posv[wd * 32 + (uint32_t)(__ffs(m) - 1)];
}

// ----------------------------------------------------- multi-search kernels
// Start of a multi-search attempt for the searches in `want` (bit s):
// empty their tables (slots of EMPTY_SLOT, src/resizable_table.h:44-50)
// and reset their level control; a search not in `want` keeps its table and
// log from the previous attempt (the other one overflowed) and is inactive.
__global__ void k_ms_clear(MSCtl *C, uint32_t want, unsigned long long *t0,
                           uint64_t s0, unsigned long long *t1, uint64_t s1) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  const uint64_t i0 = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  if (want & 1)
    for (uint64_t i = i0; i < s0; i += stride) t0[i] = EMPTY_SLOT;
  if (want & 2)
    for (uint64_t i = i0; i < s1; i += stride) t1[i] = EMPTY_SLOT;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    for (int s = 0; s < 2; ++s) {
      const bool w = (want >> s) & 1;
      C->active[s] = w;
      if (!w) continue;
      MSSlot z{};
      for (int j = 0; j < 3; ++j) C->slot[s][j] = z;
      C->log_n[s] = 0;
      C->full[s] = 0;
      for (int c = 0; c < NMC; ++c) C->cnt[s][c] = 0;
    }
  }
}

// Seed (pivot_i, label_offset_g + i) into table, log and frontier of both
// searches (src/scc.hpp:147-150); i counts the pivots of graph g in this
// round, whose front segment starts at fstart[a] (a = the graph's window,
// g = wgid[a]).  The front size is read on the device.
__global__ void k_ms_begin(const MSSearch *S, MSCtl *C, uint32_t want,
                           const uint32_t *front,
                           const unsigned long long *nfront_ptr,
                           const unsigned long long *fstart, uint32_t A,
                           const uint32_t *wgid,
                           const unsigned long long *lab_off) {
  const uint64_t nfront = *nfront_ptr;
  for (int s = 0; s < 2; ++s) {
    if (!((want >> s) & 1)) continue;
    const MSSearch M = S[s];
    uint32_t probes = 0;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < nfront;
         i += (uint64_t)gridDim.x * blockDim.x) {
      const uint32_t a = seg_of(fstart, A, i);
      const uint32_t v = front[i];
      const uint32_t lab = (uint32_t)(lab_off[wgid[a]] + (i - fstart[a]));
      const int r = table_insert(M.table, M.mask, v, lab, M.probe_limit, &probes);
      if (r < 0) C->full[s] = 1;
      M.log[i] = make_pair(v, lab);
      M.q[0][i] = make_pair(v, lab);
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      C->slot[s][0].qsize = nfront;
      C->log_n[s] = nfront;
      C->cnt[s][M_INS] = nfront;
    }
  }
}

// Staged log / frontier appends of one wave; an overflow of either list
// (the table's load limit) marks the search full.
struct MSOut {
  WaveOut<unsigned long long, OUT_CAP64> log, q;
  __device__ __forceinline__ void put(const MSSearch &M, MSCtl *C, int s,
                                      unsigned long long *qn, MSSlot *nxt,
                                      bool logit, bool spill,
                                      unsigned long long pr) {
    if (!log.put(logit, pr, &C->log_n[s], M.log, M.log_cap)) C->full[s] = 1;
    if (!q.put(spill, pr, &nxt->qsize, qn, M.q_cap)) C->full[s] = 1;
  }
  __device__ __forceinline__ void flush(const MSSearch &M, MSCtl *C, int s,
                                        unsigned long long *qn, MSSlot *nxt) {
    if (!log.flush(&C->log_n[s], M.log, M.log_cap)) C->full[s] = 1;
    if (!q.flush(&nxt->qsize, qn, M.q_cap)) C->full[s] = 1;
  }
};

// One sub-round of multi-search (src/scc.hpp:160-252) for both directions.
// Frontier items are (v, l) pairs; for each edge v->w with
// labels[w] == labels[v] and w not in a found SCC, insert (w, l); a new
// pair is logged and either expanded by the same wave (VGC, budget tau per
// seed, src/scc.hpp:177-219) or spilled to the next frontier.  Wave w takes
// a contiguous share of the frontier; log and frontier appends are staged
// in LDS.
__global__ __launch_bounds__(BLOCK) void k_ms_expand(
    const MSSearch *S, MSCtl *C, int ns, uint32_t L,
    const unsigned long long *__restrict__ labels, uint32_t tau, int vgc,
    uint32_t total_waves) {
  __shared__ unsigned long long ring_s[WPB][MS_RING];
  __shared__ unsigned long long olog_s[WPB][OUT_CAP64];
  __shared__ unsigned long long oq_s[WPB][OUT_CAP64];
  __shared__ uint32_t sc_s[WPB][WAVE];
  __shared__ uint64_t st_s[WPB][WAVE];
  __shared__ unsigned long long lab_s[WPB][WAVE];
  __shared__ uint32_t el_s[WPB][WAVE];
  const int w = threadIdx.x / WAVE, l = lane();
  unsigned long long *ring = ring_s[w];
  uint32_t *sc = sc_s[w];
  uint64_t *st = st_s[w];
  unsigned long long *lab = lab_s[w];
  uint32_t *el = el_s[w];
  const int ci = L % 3, ni = (L + 1) % 3, zi = (L + 2) % 3;
  if (blockIdx.x == 0 && threadIdx.x < (unsigned)ns) {
    MSSlot z{};
    C->slot[threadIdx.x][zi] = z;
  }
  const int pq = L & 1;
  const uint64_t gw = (blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) / WAVE;
  MSOut out;
  out.log.init(olog_s[w]);
  out.q.init(oq_s[w]);
  uint32_t acc0[5] = {}, acc1[5] = {};   // per-thread counters of search 0 / 1
  for (int s = 0; s < ns; ++s) {
    if (!C->active[s]) continue;
    MSSlot *cur = &C->slot[s][ci];
    MSSlot *nxt = &C->slot[s][ni];
    const unsigned long long nf = cur->qsize;
    if (nf == 0 || C->full[s]) continue;
    const MSSearch M = S[s];
    unsigned long long *qn = (pq ? M.q[0] : M.q[1]);
    const uint64_t per = (nf + total_waves - 1) / total_waves;
    const uint32_t sd = (uint32_t)min((uint64_t)64, per);
    uint64_t g = gw * per;
    const uint64_t gend = min((uint64_t)nf, g + per);
    unsigned long long n_edges = 0, n_exp = 0, n_ins = 0, n_app = 0;
    uint32_t probes = 0;
    uint64_t wbud = vgc ? (uint64_t)tau * VGC_WAVE_MULT : 0;   // per launch
    uint32_t nbatch = 0;
    while (g < gend) {
      if (__atomic_load_n(&C->full[s], __ATOMIC_RELAXED)) break;
      const uint32_t cnt = (uint32_t)min((uint64_t)sd, gend - g);
      if ((uint32_t)l < cnt) ring[l] = (pq ? M.q[1] : M.q[0])[g + l];
      g += cnt;
      uint32_t head = 0, tail = cnt;
      uint64_t budget = min((uint64_t)tau * cnt, wbud);
      const uint64_t budget0 = budget;
      uint64_t ebudget = budget * VGC_EDGE_MULT;
      wave_sync();
      while (head < tail) {
        if (++nbatch > (uint32_t)vgc) budget = 0;   // as in k_reach_step
        const uint32_t k = min(64u, tail - head);
        const bool valid = (uint32_t)l < k;
        const unsigned long long pr = valid ? ring[(head + l) & (MS_RING - 1)] : 0ull;
        head += k;
        const uint32_t v = pair_vertex(pr);
        uint64_t b = 0, deg = 0;
        unsigned long long lv = 0;
        if (valid) {
          b = M.off[v];
          deg = M.off[v + 1] - b;
          lv = labels[v];
          ++n_exp;
        }
        const bool heavy = deg > HEAVY_DEG;
        const uint64_t hm = ballot(heavy);
        if (hm) {
          const unsigned long long nch = heavy ? (deg + CHUNK - 1) / CHUNK : 0;
          const unsigned long long cinc = wave_incl_scan(nch);
          const unsigned long long ctot = __shfl(cinc, WAVE - 1);
          unsigned long long old = 0;
          if (l == 0)
            old = atomicAdd(&cur->heavy,
                            ((unsigned long long)__popcll(hm) << H_CNT_SHIFT) | ctot);
          old = __shfl(old, 0);
          if (heavy) {
            const uint64_t idx = (old >> H_CNT_SHIFT) + lanes_below(hm);
            M.hp[idx] = pr;
            M.hc[idx] = (old & H_CHUNK_MASK) + cinc - nch;
            deg = 0;
          }
        }
        const uint32_t inc = wave_incl_scan((uint32_t)deg);
        const uint32_t total = __shfl(inc, WAVE - 1);
        sc[l] = inc;
        st[l] = b;
        lab[l] = lv;
        el[l] = pair_label(pr);
        ebudget = ebudget > total ? ebudget - total : 0;
        if (!ebudget) budget = 0;
        wave_sync();
        for (uint32_t base = 0; base < total; base += WAVE) {
          const uint32_t e = base + l;
          bool ins = false;
          unsigned long long np = 0;
          if (e < total) {
            int lo = 0, hi = (int)k - 1;   // lanes >= k hold no list
            while (lo < hi) {
              const int mid = (lo + hi) >> 1;
              if (sc[mid] > e) hi = mid; else lo = mid + 1;
            }
            const uint32_t ex = lo ? sc[lo - 1] : 0u;
            const uint32_t x = M.adj[st[lo] + (e - ex)];
            ++n_edges;
            const unsigned long long lx = labels[x];
            if (!(lx & TOP_BIT) && lx == lab[lo]) {
              const int r = table_insert(M.table, M.mask, x, el[lo],
                                         M.probe_limit, &probes);
              if (r < 0) C->full[s] = 1;
              ins = r > 0;
              np = make_pair(x, el[lo]);
            }
          }
          const uint64_t fm = ballot(ins);
          if (!fm) continue;
          n_ins += ins;
          // VGC (src/scc.hpp:177-219): sparse new pairs stay in the wave's
          // queue while the budget lasts and the ring has room
          bool want = false;
          if (ins && vgc) {
            const uint32_t x = pair_vertex(np);
            want = M.off[x + 1] - M.off[x] <= VGC_MAX_DEG;
          }
          const uint64_t lm = ballot(want);
          const uint32_t nwant = (uint32_t)__popcll(lm);
          const uint32_t room = MS_RING - (tail - head);
          const uint32_t take = (uint32_t)min((uint64_t)min(nwant, room), budget);
          const bool loc = want && lanes_below(lm) < take;
          if (loc) ring[(tail + lanes_below(lm)) & (MS_RING - 1)] = np;
          tail += take;
          budget -= take;
          const bool glob = ins && !loc;
          n_app += glob;
          out.put(M, C, s, qn, nxt, ins, glob, np);
        }
        wave_sync();
      }
      wbud -= budget0 - budget;
    }
    out.flush(M, C, s, qn, nxt);
    const uint32_t v5[5] = {(uint32_t)n_edges, (uint32_t)n_exp, (uint32_t)n_ins,
                            (uint32_t)n_app, probes};
#pragma unroll
    for (int k = 0; k < 5; ++k) {
      if (s) acc1[k] += v5[k]; else acc0[k] += v5[k];
    }
  }
  // counters: one atomic per block and counter (every thread gets here)
  unsigned long long *const cp[10] = {
      &C->cnt[0][M_EDGES], &C->cnt[0][M_EXP], &C->cnt[0][M_INS], &C->cnt[0][M_APP],
      &C->cnt[0][M_PROBES], &C->cnt[1][M_EDGES], &C->cnt[1][M_EXP], &C->cnt[1][M_INS],
      &C->cnt[1][M_APP], &C->cnt[1][M_PROBES]};
  const unsigned long long cv[10] = {acc0[0], acc0[1], acc0[2], acc0[3], acc0[4],
                                     acc1[0], acc1[1], acc1[2], acc1[3], acc1[4]};
  block_add<10>(cp, cv);
}

// Heavy pairs of this sub-round: CHUNK-edge items, a contiguous range per
// wave.
__global__ __launch_bounds__(BLOCK) void k_ms_heavy(
    const MSSearch *S, MSCtl *C, int ns, uint32_t L,
    const unsigned long long *__restrict__ labels, uint32_t total_waves) {
  __shared__ unsigned long long olog_s[WPB][OUT_CAP64];
  __shared__ unsigned long long oq_s[WPB][OUT_CAP64];
  const int l = lane();
  const int ci = L % 3, ni = (L + 1) % 3;
  const int pq = L & 1;
  const uint64_t gw = (blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) / WAVE;
  MSOut out;
  out.log.init(olog_s[threadIdx.x / WAVE]);
  out.q.init(oq_s[threadIdx.x / WAVE]);
  uint32_t acc0[5] = {}, acc1[5] = {};   // per-thread counters of search 0 / 1
  for (int s = 0; s < ns; ++s) {
    if (!C->active[s]) continue;
    MSSlot *cur = &C->slot[s][ci];
    MSSlot *nxt = &C->slot[s][ni];
    const unsigned long long hw = cur->heavy;
    const uint64_t H = hw >> H_CNT_SHIFT, Ct = hw & H_CHUNK_MASK;
    if (H == 0 || C->full[s]) continue;
    const MSSearch M = S[s];
    unsigned long long *qn = (pq ? M.q[0] : M.q[1]);
    unsigned long long n_edges = 0, n_ins = 0, n_app = 0, n_items = 0;
    uint32_t probes = 0;
    const uint64_t per = (Ct + total_waves - 1) / total_waves;
    uint64_t c = gw * per;
    const uint64_t cend = min(Ct, c + per);
    uint64_t own = 0;
    if (c < cend) {
      uint64_t lo = 0, hi = H - 1;
      while (lo < hi) {
        const uint64_t mid = (lo + hi + 1) >> 1;
        if (M.hc[mid] <= c) lo = mid; else hi = mid - 1;
      }
      own = lo;
    }
    for (; c < cend; ++c) {
      while (own + 1 < H && M.hc[own + 1] <= c) ++own;
      if (l == 0) ++n_items;
      const unsigned long long pr = M.hp[own];
      const uint32_t v = pair_vertex(pr), ell = pair_label(pr);
      const unsigned long long lv = labels[v];
      const uint64_t k = c - M.hc[own];
      const uint64_t eb = M.off[v] + k * CHUNK;
      const uint64_t ee = min(M.off[v + 1], eb + CHUNK);
      for (uint64_t j0 = eb; j0 < ee; j0 += WAVE) {
        const uint64_t j = j0 + l;
        bool ins = false;
        unsigned long long np = 0;
        if (j < ee) {
          const uint32_t x = M.adj[j];
          ++n_edges;
          const unsigned long long lx = labels[x];
          if (!(lx & TOP_BIT) && lx == lv) {
            const int r = table_insert(M.table, M.mask, x, ell, M.probe_limit,
                                       &probes);
            if (r < 0) C->full[s] = 1;
            ins = r > 0;
            np = make_pair(x, ell);
          }
        }
        n_ins += ins;
        n_app += ins;
        out.put(M, C, s, qn, nxt, ins, ins, np);
      }
    }
    out.flush(M, C, s, qn, nxt);
    const uint32_t v5[5] = {(uint32_t)n_edges, (uint32_t)n_ins, (uint32_t)n_app, probes,
                            (uint32_t)n_items};
#pragma unroll
    for (int k = 0; k < 5; ++k) {
      if (s) acc1[k] += v5[k]; else acc0[k] += v5[k];
    }
  }
  // counters: one atomic per block and counter (every thread gets here)
  unsigned long long *const cp[10] = {
      &C->cnt[0][M_HEAVY_EDGES], &C->cnt[0][M_HEAVY_INS], &C->cnt[0][M_HEAVY_APP],
      &C->cnt[0][M_HEAVY_PROBES], &C->cnt[0][M_HEAVY_ITEMS], &C->cnt[1][M_HEAVY_EDGES],
      &C->cnt[1][M_HEAVY_INS], &C->cnt[1][M_HEAVY_APP], &C->cnt[1][M_HEAVY_PROBES],
      &C->cnt[1][M_HEAVY_ITEMS]};
  const unsigned long long cv[10] = {acc0[0], acc0[1], acc0[2], acc0[3], acc0[4],
                                     acc1[0], acc1[1], acc1[2], acc1[3], acc1[4]};
  block_add<10>(cp, cv);
}

// Partition (src/scc.hpp:437-464): pairs of the smaller table get
// label|TOP_BIT when the larger table holds them too, else label; pairs of
// the larger table get label.  write_max is commutative, so both logs are
// applied in one launch.
__global__ __launch_bounds__(BLOCK) void k_partition(
    const MSSearch *S, const MSCtl *C, int small, unsigned long long *labels,
    unsigned long long *probe_ctr, const unsigned long long *fstart, uint32_t nwin,
    const uint32_t *wgid, unsigned long long *lab_off) {
  // label_offset_g += pivots of g in this round (src/scc.hpp:425)
  if (blockIdx.x == 0)
    for (uint32_t a = threadIdx.x; a < nwin; a += blockDim.x)
      lab_off[wgid[a]] += fstart[a + 1] - fstart[a];
  const int large = small ^ 1;
  const MSSearch A = S[small], B = S[large];
  const uint64_t na = C->log_n[small], nb = C->log_n[large];
  uint32_t probes = 0;
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
       i < na + nb; i += (uint64_t)gridDim.x * blockDim.x) {
    unsigned long long pr, val;
    if (i < na) {
      pr = A.log[i];
      const uint32_t v = pair_vertex(pr), ell = pair_label(pr);
      val = table_contains(B.table, B.mask, v, ell, &probes)
                ? ((unsigned long long)ell | TOP_BIT)
                : (unsigned long long)ell;
    } else {
      pr = B.log[i - na];
      val = pair_label(pr);
    }
    unsigned long long *dst = &labels[pair_vertex(pr)];
    if (*dst < val) atomicMax(dst, val);
  }
  unsigned long long *const cp[1] = {probe_ctr};
  const unsigned long long cv[1] = {probes};
  block_add<1>(cp, cv);   // one atomic per block
}

// -------------------------------------------------------- finalize, status
// `go` (may be null): skip unless *go is nonzero (a finalize enqueued
// before the host knows that every graph is finished).
__global__ void k_finalize(unsigned long long *labels, uint64_t n,
                           const unsigned long long *go) {
  if (go && *go == 0) return;
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    labels[i] = (labels[i] & VAL_MASK) - 1;   // src/scc.hpp:475-476
}
// Finalize when the first labelling wrote final values directly: only the
// remaining vertices (centers) and each graph's trimmed vertex of rank 0
// still hold internal labels.
__global__ void k_finalize_rest(unsigned long long *labels, const uint32_t *centers,
                                uint64_t n_rem, const uint32_t *first_trim,
                                const unsigned long long *zstart, uint32_t G,
                                const unsigned long long *go) {
  if (go && *go == 0) return;
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n_rem + G;
       i += (uint64_t)gridDim.x * blockDim.x) {
    uint64_t v;
    if (i < n_rem) {
      v = centers[i];
    } else {
      const uint64_t g = i - n_rem;
      if (zstart[g + 1] == zstart[g]) continue;   // no trimmed vertex
      v = first_trim[g];
    }
    labels[v] = (labels[v] & VAL_MASK) - 1;
  }
}
// Canonical labels: minimum local vertex id per (graph, label); labels
// are in [0, n_g], so (voff[g] + g + label) indexes a table of n + G slots.
__global__ void k_canon_min(const unsigned long long *labels, uint64_t n,
                            const unsigned long long *voff, uint32_t G,
                            uint32_t *minv) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t g = seg_of(voff, G, i);
    atomicMin(&minv[voff[g] + g + labels[i]], (uint32_t)(i - voff[g]));
  }
}
__global__ void k_canon_apply(unsigned long long *labels, uint64_t n,
                              const unsigned long long *voff, uint32_t G,
                              const uint32_t *minv) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t g = seg_of(voff, G, i);
    labels[i] = minv[voff[g] + g + labels[i]];
  }
}
__global__ void k_count_labels(const unsigned long long *labels, uint64_t n,
                               uint32_t *cnt, unsigned long long *bad) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const unsigned long long x = labels[i];
    if (x <= n) atomicAdd(&cnt[x], 1u);
    else atomicAdd(bad, 1ull);
  }
}
__global__ void k_status_reduce(const uint32_t *cnt, uint64_t len,
                                unsigned long long *nscc, uint32_t *largest) {
  unsigned long long c = 0;
  uint32_t mx = 0;
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < len;
       i += (uint64_t)gridDim.x * blockDim.x) {
    c += cnt[i] != 0;
    mx = max(mx, cnt[i]);
  }
  c = wave_sum(c);
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) mx = max(mx, (uint32_t)__shfl_xor(mx, d));
  if (lane() == 0) {
    if (c) atomicAdd(nscc, c);
    atomicMax(largest, mx);
  }
}

// ======================================================================
//                              host side
// ======================================================================

static int grid_for(uint64_t n, int per_block = BLOCK, int cap = 65536) {
  uint64_t g = (n + per_block - 1) / per_block;
  if (g < 1) g = 1;
  if (g > (uint64_t)cap) g = cap;
  return (int)g;
}

// ---------------------------------------------------------- capture gate
namespace {
std::mutex g_gate_mu;
std::condition_variable g_gate_cv;
int g_gate_readers = 0;
bool g_gate_writer = false;
thread_local int t_gate_depth = 0;
}  // namespace

void gate_enter() {
  if (t_gate_depth++ > 0) return;
  std::unique_lock<std::mutex> l(g_gate_mu);
  g_gate_cv.wait(l, [] { return !g_gate_writer; });
  ++g_gate_readers;
}

void gate_leave() {
  if (--t_gate_depth > 0) return;
  std::lock_guard<std::mutex> l(g_gate_mu);
  if (--g_gate_readers == 0) g_gate_cv.notify_all();
}

CaptureGuard::CaptureGuard() {
  std::unique_lock<std::mutex> l(g_gate_mu);
  g_gate_cv.wait(l, [] { return !g_gate_writer; });
  g_gate_writer = true;   // new entries wait from here on
  g_gate_cv.wait(l, [] { return g_gate_readers == 0; });
}

CaptureGuard::~CaptureGuard() {
  std::lock_guard<std::mutex> l(g_gate_mu);
  g_gate_writer = false;
  g_gate_cv.notify_all();
}

static std::once_flag g_pool_once;
static hipStream_t g_free_stream = nullptr;

void pool_setup() {
  GateGuard gate;
  std::call_once(g_pool_once, [] {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return;
    hipMemPool_t pool;
    if (hipDeviceGetDefaultMemPool(&pool, dev) == hipSuccess) {
      uint64_t thr = ~0ull;   // keep freed memory in the pool for reuse
      (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &thr);
    }
    (void)hipStreamCreateWithFlags(&g_free_stream, hipStreamNonBlocking);
  });
}

hipStream_t pool_free_stream() {
  pool_setup();
  return g_free_stream;
}

void *Arena::take(size_t bytes) {
  const size_t at = (used + 255) & ~(size_t)255;
  if (at > cap || bytes > cap - at) return nullptr;
  used = at + bytes;
  peak = std::max(peak, used);
  return base + at;
}

int DevBuf::ensure(size_t bytes) {
  if (bytes <= cap) return PSC_OK;
  if (arena && *arena) {
    Arena *a = *arena;
    void *p = a->take(bytes);
    if (!p) {
      set_error("scratch exhausted: a buffer of " + std::to_string(bytes) + " B does not fit (" +
                std::to_string(a->used) + " of " + std::to_string(a->cap) + " B in use)");
      return PSC_ERR_NOMEM;
    }
    ptr = p;
    cap = bytes;
    return PSC_OK;
  }
  pool_setup();
  GateGuard gate;
  hipStream_t st = owner ? *owner : nullptr;
  if (ptr) {
    // a piece of a set-up block stays with the block
    hipError_t e = carved ? hipSuccess : hipFreeAsync(ptr, st);
    carved = false;
    ptr = nullptr;
    // a buffer that has to grow again gets headroom
    bytes = std::max(bytes, cap + cap / 2);
    cap = 0;
    if (e != hipSuccess) {
      set_error(std::string("hipFreeAsync: ") + hipGetErrorString(e));
      return PSC_ERR_HIP;
    }
  }
  if (bytes == 0) return PSC_OK;
  hipError_t e = hipMallocAsync(&ptr, bytes, st);
  if (e != hipSuccess) {
    ptr = nullptr;
    set_error("hipMallocAsync(" + std::to_string(bytes) + " B): " +
              hipGetErrorString(e));
    return PSC_ERR_NOMEM;
  }
  cap = bytes;
  return PSC_OK;
}

void DevBuf::release() {
  if (ptr && !carved && !(arena && *arena)) {
    GateGuard gate;
    (void)hipFreeAsync(ptr, owner ? *owner : nullptr);
  }
  carved = false;
  ptr = nullptr;
  cap = 0;
}

Engine::Engine() {
  for (DevBuf *b : all_bufs()) {
    b->owner = &stream;
    b->arena = &arena;
  }
  for (DevBuf *b : {&g_off, &g_adj, &g_ioff, &g_iadj}) {
    b->owner = &stream;
    b->arena = &arena;
  }
}

Engine::~Engine() {
  GateGuard gate;
  if (ready) (void)hipStreamSynchronize(stream);
  for (auto &g : gcache) (void)hipGraphExecDestroy(g.second);
  gcache.clear();
  for (DevBuf *b : all_bufs()) b->release();
  if (owned_graph)
    for (DevBuf *b : {&g_off, &g_adj, &g_ioff, &g_iadj}) b->release();
  if (host_static) return;   // stream and host staging belong to the caller / the engine
  if (lane_arena.base) {
    if (stream) (void)hipStreamSynchronize(stream);
    (void)hipFree(lane_arena.base);
  }
  for (void *b : setup_blocks) (void)hipFreeAsync(b, stream);
  if (stream) (void)hipStreamSynchronize(stream);
  for (auto &e : pev) (void)hipEventDestroy(e);
  // the stream, its events and the pinned staging buffers go back to the
  // process's cache for the next handle on this device
  StreamKit k{device, stream, {}, h_pinned, h_gpin, h_gpin_cap, tail_pin, tail_pin_cap};
  for (int i = 0; i < 8; ++i) k.ev[i] = ev[i];
  if (!stream || !kit_put(k)) {
    if (h_gpin) (void)hipHostFree(h_gpin);
    if (tail_pin) (void)hipHostFree(tail_pin);
    if (!stream) return;
    if (h_pinned) (void)hipHostFree(h_pinned);
    for (auto &e : ev) if (e) (void)hipEventDestroy(e);
    (void)hipStreamDestroy(stream);
  }
}

// Streams, their timing events and the pinned staging buffers of destroyed
// handles, kept for reuse: creating a stream or pinning host memory costs
// the runtime up to milliseconds, which a new handle (a serving lane's, a
// batch group's) would otherwise pay on every graph.
namespace {
std::mutex g_kit_mu;
std::vector<StreamKit> g_kits;
constexpr size_t KIT_CACHE = 32;
}  // namespace

bool kit_put(const StreamKit &k) {
  std::lock_guard<std::mutex> g(g_kit_mu);
  if (g_kits.size() >= KIT_CACHE) return false;
  g_kits.push_back(k);
  return true;
}

bool kit_take(int device, StreamKit *k) {
  std::lock_guard<std::mutex> g(g_kit_mu);
  for (size_t i = g_kits.size(); i-- > 0;)
    if (g_kits[i].device == device) {
      *k = g_kits[i];
      g_kits.erase(g_kits.begin() + i);
      return true;
    }
  return false;
}

int Engine::ensure_all(const std::vector<std::pair<DevBuf *, size_t>> &req) {
  size_t total = 0, fresh = 0;
  bool all_empty = true;
  for (const auto &r : req)
    if (r.second > r.first->cap) {
      if (r.first->ptr) all_empty = false;
      total += (r.second + 255) & ~(size_t)255;
      ++fresh;
    }
  if (arena || !all_empty || fresh < 2) {
    for (const auto &r : req) PSC_TRY(r.first->ensure(r.second));
    return PSC_OK;
  }
  pool_setup();
  GateGuard gate;
  void *blk = nullptr;
  hipError_t e = hipMallocAsync(&blk, total, stream);
  if (e != hipSuccess) {
    set_error("hipMallocAsync(" + std::to_string(total) + " B): " + hipGetErrorString(e));
    return PSC_ERR_NOMEM;
  }
  setup_blocks.push_back(blk);
  setup_bytes += total;
  char *at = static_cast<char *>(blk);
  for (const auto &r : req)
    if (r.second > r.first->cap) {
      r.first->ptr = at;
      r.first->cap = r.second;
      r.first->carved = true;
      at += (r.second + 255) & ~(size_t)255;
    }
  return PSC_OK;
}

std::vector<DevBuf *> Engine::all_bufs() {
  return {&labels, &vis[0], &vis[1], &rq[0][0], &rq[0][1], &rq[1][0],
          &rq[1][1], &hv[0], &hv[1], &hc[0], &hc[1], &pbits,
          &sort_tmp, &centers,
          &front, &trim_bm, &tiles, &tile_off, &trim_off, &reach_desc, &reach_ctl, &ms_desc,
          &ms_ctl, &misc, &table[0], &table[1], &log[0], &log[1],
          &mq[0][0], &mq[0][1], &mq[1][0], &mq[1][1], &mhp[0], &mhp[1],
          &mhc[0], &mhc[1], &scratch32, &voff_d, &gA, &gB, &gC, &gD,
          &gid_d, &rnd, &rr_plab, &first_trim, &rhead[0], &rhead[1], &hubs, &trim_static, &dead, &tail_cnt, &tail_lab, &tail_off, &tail_adj, &tail_table,
          &tail_lists, &tail_meta, &tail_spec, &tail_trace};
}

int Engine::download_labels(uint64_t *dst) {
  hipError_t he = hipMemcpyAsync(dst, labels.ptr, n * 8, hipMemcpyDeviceToHost, stream);
  if (he == hipSuccess) he = hipStreamSynchronize(stream);
  if (he != hipSuccess) {
    psc::set_error(std::string("labels download: ") + hipGetErrorString(he));
    return PSC_ERR_HIP;
  }
  return PSC_OK;
}

int Engine::setup() {
  if (ready) return PSC_OK;
  GateGuard gate;
  PSC_HIP(hipGetDevice(&device));
  hipDeviceProp_t prop;
  PSC_HIP(hipGetDeviceProperties(&prop, device));
  num_cu = prop.multiProcessorCount;
  if (host_static) {
    // host staging in the engine itself: polls and descriptors [0, 4096),
    // per-graph regions [4096, 7168), remainder descriptors [7168, 8192)
    h_pinned = host_stage;
    h_gpin = reinterpret_cast<unsigned long long *>(host_stage + 4096);
    h_gpin_cap = 3072 / 8;
    tail_pin = host_stage + 7168;
    tail_pin_cap = 1024;
  } else {
    pool_setup();
    StreamKit k;
    if (kit_take(device, &k)) {
      stream = k.stream;
      for (int i = 0; i < 8; ++i) ev[i] = k.ev[i];
      h_pinned = k.pinned;
      h_gpin = static_cast<unsigned long long *>(k.gpin);
      h_gpin_cap = k.gpin_cap;
      tail_pin = k.tpin;
      tail_pin_cap = k.tpin_cap;
    } else {
      PSC_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
      for (auto &e : ev) PSC_HIP(hipEventCreate(&e));
      PSC_HIP(hipHostMalloc(&h_pinned, 4096));
    }
  }
  int occ = 0;
  PSC_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k_reach_step,
                                                       BLOCK, 0));
  grid_step = std::max(1, occ) * num_cu;
  PSC_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k_ms_expand,
                                                       BLOCK, 0));
  grid_ms = std::max(1, occ) * num_cu;
  PSC_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k_reach_heavy,
                                                       BLOCK, 0));
  grid_heavy = std::max(1, occ) * num_cu;
  PSC_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k_reach_pull, BLOCK, 0));
  grid_pull = std::max(1, std::min(occ, 2)) * num_cu;
  ready = true;
  return PSC_OK;
}

// The device buffers a binding of n vertices, m edges and G graphs needs
// before its first run (the run itself sizes the round buffers: pivot
// bitmap, tables, logs, remainder copy).
std::vector<std::pair<DevBuf *, size_t>> Engine::init_requests(uint64_t n_, uint64_t m_,
                                                               uint64_t G_) {
  const uint64_t nn = std::max<uint64_t>(n_, 1);
  const uint64_t words = (nn + 31) / 32;
  std::vector<std::pair<DevBuf *, size_t>> req;
  req.emplace_back(&labels, nn * 8);
  const uint64_t hcap = std::min<uint64_t>(nn, m_ / std::min(HEAVY_DEG, PULL_HEAVY) + 1);
  for (int s = 0; s < 2; ++s) {
    req.emplace_back(&vis[s], words * 4);
    req.emplace_back(&rq[s][0], nn * 4);
    req.emplace_back(&rq[s][1], nn * 4);
    req.emplace_back(&hv[s], hcap * 4);
    req.emplace_back(&hc[s], hcap * 8);
    req.emplace_back(&rhead[s], nn * 4);
  }
  req.emplace_back(&centers, nn * 4);
  req.emplace_back(&front, nn * 4);
  const uint64_t ntiles = (nn + SEL_TILE - 1) / SEL_TILE;
  req.emplace_back(&tiles, ntiles * 4);
  req.emplace_back(&trim_bm, words * 4);
  req.emplace_back(&trim_static, words * 4);
  req.emplace_back(&dead, words * 8);
  req.emplace_back(&tile_off, ntiles * 8);
  req.emplace_back(&trim_off, ntiles * 8);
  req.emplace_back(&reach_desc, sizeof(ReachSearch) * 2);
  req.emplace_back(&reach_ctl, sizeof(ReachCtl));
  req.emplace_back(&ms_desc, sizeof(MSSearch) * 2);
  req.emplace_back(&ms_ctl, sizeof(MSCtl));
  req.emplace_back(&misc, 256);
  req.emplace_back(&scratch32, (nn + G_) * 4);
  for (DevBuf *b : {&voff_d, &gB, &gC, &gD, &rr_plab}) req.emplace_back(b, (G_ + 1) * 8);
  req.emplace_back(&gA, 2 * (G_ + 1) * 8);   // nstart, then zstart
  req.emplace_back(&gid_d, (G_ + 1) * 4);
  req.emplace_back(&first_trim, (G_ + 1) * 4);
  req.emplace_back(&hubs, 2 * HUB_WORDS * 4);
  return req;
}

int Engine::init(const psc_graph *dg, bool owned, uint32_t ngraphs,
                 const uint64_t *voff) {
  n = dg->n;
  m = dg->m;
  G = ngraphs ? ngraphs : 1;
  voff_h.assign(G + 1, 0);
  if (voff) {
    for (uint32_t g = 0; g <= G; ++g) voff_h[g] = voff[g];
  } else {
    voff_h[G] = n;
  }
  owned_graph = owned;
  d_off = dg->offset;
  d_adj = dg->edges;
  d_ioff = dg->in_offset;
  d_iadj = dg->in_edges;
  // a handle may be re-bound to new graphs (serving lanes): the stream,
  // events, pinned staging, launch shapes and device buffers persist and
  // buffers only grow
  PSC_TRY(setup());
  ms_dev_valid = false;   // descriptors may move

  const uint64_t nn = std::max<uint64_t>(n, 1);
  const uint64_t words = (nn + 31) / 32;
  if (arena && labels_ext) {   // the caller's output array holds the labels
    labels.ptr = labels_ext;
    labels.cap = nn * 8;
  }
  PSC_TRY(ensure_all(init_requests(n, m, G)));
  PSC_TRY(gpin(2 * (G + 1)));   // per-graph starts (init), windows (runs)
  PSC_TRY(upload(voff_d, voff_h.data(), G + 1, 1));
  // per-graph state read by every pull level: the first neighbour of each
  // vertex in the pull direction, gathered once here so that a pull probe of
  // it is one coalesced 4-byte load instead of an offset pair and a line of
  // the adjacency array
  {
    // a pull of search s reads the other CSR; its candidates' in-neighbours
    // are reached along the search's own CSR, whose degree ranks them.  The
    // round front and the centers (n entries each, unused until the first
    // labelling) hold the degrees.
    HeadDirs hd;
    for (int s = 0; s < 2; ++s) {
      hd.d[s].own_off = s == 0 ? d_off : d_ioff;
      hd.d[s].pull_off = s == 0 ? d_ioff : d_off;
      hd.d[s].pull_adj = s == 0 ? d_iadj : d_adj;
      hd.d[s].lg = rq[s][0].as<uint8_t>();   // frontier buffer, free until the runs
      hd.d[s].hub = hubs.as<uint32_t>() + s * HUB_WORDS;
      hd.d[s].head = rhead[s].as<uint32_t>();
    }
    hd.trim = trim_static.as<uint32_t>();
    hd.nzoff = trim_off.as<unsigned long long>();
    // Trimming (src/scc.hpp:295-306) depends on the graph alone, so the
    // binding does it once: the trimmed vertices (zero in- or out-degree),
    // the dead ends of each search, the non-trimmed count before each tile
    // and before each graph.  Heads are stored by non-trimmed rank, so the
    // pull levels read them densely.
    const uint64_t ntiles = (nn + SEL_TILE - 1) / SEL_TILE;
    unsigned long long *nstart = gA.as<unsigned long long>();
    starts_h.assign(2 * (G + 1), 0);
    if (n) {
      unsigned long long *tot = misc.as<unsigned long long>();
      for (int s = 0; s < 2; ++s)
        k_fill_u32<<<1, BLOCK, 0, stream>>>(hd.d[s].hub, HUB_BINS, 0u);
      k_trim_count<<<(unsigned)std::min<uint64_t>(ntiles, 5 * (uint64_t)num_cu), BLOCK, 0, stream>>>(
          d_off, d_ioff, n, words, tiles.as<uint32_t>(), dead.as<uint32_t>(),
          dead.as<uint32_t>() + words, trim_static.as<uint32_t>(), tot + 2, hd.d[0].lg,
          hd.d[1].lg, hd.d[0].hub, hd.d[1].hub);
      PSC_TRY(scan_tiles(ntiles, tot, trim_off.as<unsigned long long>()));
      k_graph_starts<<<grid_for((G + 1) * (uint64_t)WAVE), BLOCK, 0, stream>>>(
          d_off, d_ioff, n, voff_d.as<unsigned long long>(), G,
          trim_off.as<unsigned long long>(), tot, nstart, nstart + (G + 1));
      PSC_TRY(gpin(2 * (G + 1)));
      PSC_HIP(hipMemcpyAsync(region(2), nstart, 16 * (G + 1), hipMemcpyDeviceToHost, stream));
    }
    const unsigned blocks = (unsigned)std::min<uint64_t>((nn + BLOCK - 1) / BLOCK, 2048);
    const unsigned hblocks = (unsigned)std::max<uint64_t>(
        1, std::min<uint64_t>((nn + HEAD_BLOCK - 1) / HEAD_BLOCK, (uint64_t)num_cu));
    if (!n)
      for (int s = 0; s < 2; ++s)
        k_fill_u32<<<1, BLOCK, 0, stream>>>(hd.d[s].hub, HUB_BINS, 0u);
    k_hub_select<<<dim3(1, 2), BLOCK, 0, stream>>>(hd);
    k_hub_insert<<<dim3(blocks, 2), BLOCK, 0, stream>>>(hd, n);
    k_pull_heads<<<dim3(hblocks, 2), HEAD_BLOCK, 0, stream>>>(hd, n);
    PSC_HIP(hipGetLastError());
  }

  ReachSearch rs[2];
  for (int s = 0; s < 2; ++s) {
    rs[s].off = s == 0 ? d_off : d_ioff;
    rs[s].adj = s == 0 ? d_adj : d_iadj;
    rs[s].roff = s == 0 ? d_ioff : d_off;
    rs[s].radj = s == 0 ? d_iadj : d_adj;
    rs[s].rhead = rhead[s].as<uint32_t>();
    rs[s].vis = vis[s].as<uint32_t>();
    rs[s].trim = trim_static.as<uint32_t>();
    rs[s].tstat = trim_static.as<uint32_t>();
    rs[s].nzoff = trim_off.as<unsigned long long>();
    rs[s].q[0] = rq[s][0].as<uint32_t>();
    rs[s].q[1] = rq[s][1].as<uint32_t>();
    rs[s].hv = hv[s].as<uint32_t>();
    rs[s].hc = hc[s].as<unsigned long long>();
    rs[s].thresh = 0;
    rs[s].n = (uint32_t)n;
    rs[s].pad = 0;
  }
  std::memcpy(h_reach, rs, sizeof(rs));
  PSC_HIP(hipStreamSynchronize(stream));
  if (n) std::memcpy(starts_h.data(), region(2), 16 * (G + 1));
  return PSC_OK;
}

// Stable select over [0, len) with functor f; count -> host.
template <class F>
int Engine::select(F f, uint64_t len, typename F::T *sel, typename F::T *rej,
                   uint64_t *count, unsigned long long *count_dev) {
  if (len == 0) {
    if (count) *count = 0;
    if (count_dev)
      k_fill_u64<<<1, 64, 0, stream>>>(count_dev, 1, 0ull);
    return PSC_OK;
  }
  const uint64_t nt = (len + SEL_TILE - 1) / SEL_TILE;
  unsigned long long *tot = count_dev ? count_dev : misc.as<unsigned long long>();
  prof_begin(KK_SELECT);
  k_sel_count<F><<<(unsigned)nt, BLOCK, 0, stream>>>(f, len, tiles.as<uint32_t>());
  PSC_TRY(scan_tiles(nt, tot));
  k_sel_scatter<F><<<(unsigned)nt, BLOCK, 0, stream>>>(
      f, len, tile_off.as<unsigned long long>(), sel, rej);
  prof_end();
  PSC_HIP(hipGetLastError());
  if (count_dev) return PSC_OK;   // count stays on the device, no sync
  return poll(count, tot, 8);
}

__global__ void k_tile_total(const uint32_t *cnt, const unsigned long long *off,
                             uint64_t nt, unsigned long long *total) {
  if (blockIdx.x == 0 && threadIdx.x == 0) *total = off[nt - 1] + cnt[nt - 1];
}

// Exclusive scan of the per-tile counts (tiles -> tile_off) and their
// total (-> *total), on the device.
int Engine::scan_tiles(uint64_t nt, unsigned long long *total, unsigned long long *out) {
  size_t tb = 0;
  const uint32_t *in = tiles.as<uint32_t>();
  if (!out) out = tile_off.as<unsigned long long>();
  PSC_HIP(rocprim::exclusive_scan(nullptr, tb, in, out, 0ull, nt,
                                  rocprim::plus<unsigned long long>(), stream));
  PSC_TRY(sort_tmp.ensure(std::max<size_t>(tb, 8)));
  tb = sort_tmp.cap;
  PSC_HIP(rocprim::exclusive_scan(sort_tmp.ptr, tb, in, out, 0ull, nt,
                                  rocprim::plus<unsigned long long>(), stream));
  k_tile_total<<<1, 64, 0, stream>>>(in, out, nt, total);
  return PSC_OK;
}

// Pinned staging for per-graph arrays: NREG regions of (G + 2) u64.
// Region 0 serves polls; each uploaded array of a round has its own region.
// Every round polls the device after its uploads, so a region is never
// rewritten while its previous copy is still pending.
// Empty the arena for a new binding.  The fixed-size control blocks are
// carved first, so they keep their addresses from binding to binding (the
// keys of captured level batches hold them).
void Engine::arena_reset() {
  arena->used = 0;
  for (DevBuf *b : all_bufs()) b->ptr = nullptr, b->cap = 0, b->carved = false;
  for (DevBuf *b : {&g_off, &g_adj, &g_ioff, &g_iadj}) b->ptr = nullptr, b->cap = 0;
  (void)reach_desc.ensure(sizeof(ReachSearch) * 2);
  (void)reach_ctl.ensure(sizeof(ReachCtl));
  (void)ms_desc.ensure(sizeof(MSSearch) * 2);
  (void)ms_ctl.ensure(sizeof(MSCtl));
  (void)misc.ensure(256);
}

int Engine::gpin(size_t count) {
  const size_t want = count * NREG;
  if (want <= h_gpin_cap) return PSC_OK;
  if (host_static) {
    set_error("allocation-free run: per-graph staging holds " +
              std::to_string(h_gpin_cap / NREG) + " entries, " + std::to_string(count) +
              " needed");
    return PSC_ERR_INVALID;
  }
  GateGuard gate;
  PSC_HIP(hipStreamSynchronize(stream));
  if (h_gpin) (void)hipHostFree(h_gpin);
  h_gpin = nullptr;
  h_gpin_cap = 0;
  PSC_HIP(hipHostMalloc((void **)&h_gpin, want * 8));
  h_gpin_cap = want;
  return PSC_OK;
}

unsigned long long *Engine::region(int r) {
  return h_gpin + (size_t)r * (h_gpin_cap / NREG);
}

int Engine::upload(DevBuf &dst, const uint64_t *src, size_t count, int reg) {
  PSC_TRY(dst.ensure(std::max<size_t>(count, 1) * 8));
  PSC_TRY(gpin(count));
  std::memcpy(region(reg), src, count * 8);
  PSC_HIP(hipMemcpyAsync(dst.ptr, region(reg), count * 8, hipMemcpyHostToDevice,
                         stream));
  return PSC_OK;
}

int Engine::upload32(DevBuf &dst, const uint32_t *src, size_t count, int reg) {
  PSC_TRY(dst.ensure(std::max<size_t>(count, 1) * 4));
  PSC_TRY(gpin((count + 1) / 2));
  std::memcpy(region(reg), src, count * 4);
  PSC_HIP(hipMemcpyAsync(dst.ptr, region(reg), count * 4, hipMemcpyHostToDevice,
                         stream));
  return PSC_OK;
}

// The windows of a round in one copy: prefix (A+1) u64, sources (A) u64,
// graph ids (A) u32, staged in pinned regions 1..3 (3 (G + 2) u64 >= what
// A <= G windows need).
int Engine::upload_round(const std::vector<uint64_t> &wpre,
                         const std::vector<uint64_t> &wsrc,
                         const std::vector<uint32_t> &wgid) {
  const size_t A = wgid.size();
  const size_t words = (A + 1) + A + (A + 1) / 2;
  PSC_TRY(rnd.ensure(std::max<size_t>(words, 1) * 8));
  PSC_TRY(gpin(G + 2));
  unsigned long long *h = region(1);
  std::memcpy(h, wpre.data(), (A + 1) * 8);
  std::memcpy(h + A + 1, wsrc.data(), A * 8);
  std::memcpy(h + 2 * A + 1, wgid.data(), A * 4);
  PSC_HIP(hipMemcpyAsync(rnd.ptr, h, words * 8, hipMemcpyHostToDevice, stream));
  return PSC_OK;
}

// Replays the captured level batch `k`, capturing it first from
// `enqueue` (which launches the batch's kernels on the stream).  At most
// 32 batches are kept; the oldest is dropped once the stream is idle.
template <class F>
int Engine::replay(const BatchKey &k, F enqueue) {
  hipGraphExec_t ex = nullptr;
  for (auto &g : gcache)
    if (g.first == k) {
      ex = g.second;
      break;
    }
  if (!ex) {
    CaptureGuard capture;   // no allocation or free in any thread meanwhile
    hipGraph_t graph = nullptr;
    PSC_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeRelaxed));
    enqueue();
    const hipError_t le = hipGetLastError();
    const hipError_t ee = hipStreamEndCapture(stream, &graph);
    if (le != hipSuccess || ee != hipSuccess || !graph) {
      if (graph) (void)hipGraphDestroy(graph);
      set_error(std::string("level batch capture: ") +
                hipGetErrorString(le != hipSuccess ? le : ee));
      return PSC_ERR_HIP;
    }
    const hipError_t ie = hipGraphInstantiate(&ex, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (ie != hipSuccess) {
      set_error(std::string("hipGraphInstantiate: ") + hipGetErrorString(ie));
      return PSC_ERR_HIP;
    }
    if (gcache.size() >= 32) {
      PSC_HIP(hipStreamSynchronize(stream));
      (void)hipGraphExecDestroy(gcache.front().second);
      gcache.erase(gcache.begin());
    }
    gcache.push_back({k, ex});
    ++graph_captures;
  }
  PSC_HIP(hipGraphLaunch(ex, stream));
  ++graph_replays;
  return PSC_OK;
}

int Engine::poll(void *dst, const void *src, size_t bytes) {
  char *buf = (char *)h_pinned + PIN_POLL;
  if (bytes > PIN_REACH - PIN_POLL) {
    PSC_TRY(gpin((bytes + 7) / 8));
    buf = (char *)region(0);
  }
  PSC_HIP(hipMemcpyAsync(buf, src, bytes, hipMemcpyDeviceToHost, stream));
  PSC_HIP(hipStreamSynchronize(stream));
  ++syncs;
  std::memcpy(dst, buf, bytes);
  return PSC_OK;
}

// Serial 64-vertex batches a wave may run in one launch of a VGC local
// search (psc_params.vgc_batches, default VGC_MAX_BATCHES).
static int vgc_batch_cap(const psc_params &p) {
  const uint64_t b = p.vgc_batches ? p.vgc_batches : VGC_MAX_BATCHES;
  return (int)std::min<uint64_t>(b, 1u << 30);
}

// First round: fused forward + backward reach from P[0] (src/scc.hpp:316-339).
// Levels are enqueued in batches of K without host round trips; kernels of
// levels past the end exit at once.  The host reads the done flags after
// each batch.
int Engine::first_reach(const psc_params &p, psc_stats *st, ReachRound *rr,
                        const std::function<int()> &after_levels, bool *after_ran) {
  if (after_ran) *after_ran = false;
  ReachSearch rs[2];
  static_assert(2 * sizeof(ReachSearch) <= sizeof(h_reach), "reach descriptors");
  std::memcpy(rs, h_reach, sizeof(rs));
  const uint64_t thresh = m / (p.dense_div ? p.dense_div : 20);
  for (int s = 0; s < 2; ++s) {
    rs[s].thresh = thresh;
    // frontier vertices are non-trimmed: their mean degree is m / nnz
    rs[s].avg_deg = nnz_total ? (double)m / (double)nnz_total : 0.0;
    rs[s].nnz = nnz_total;   // a single-pivot round counts it on the device
    // a single-pivot round skips its exclusions (trim_bm); the first reach
    // skips the trimmed vertices; head ranks always count the trimmed set
    rs[s].excl = rr ? trim_bm.as<uint32_t>() : nullptr;
    rs[s].trim = rr ? trim_bm.as<uint32_t>() : trim_static.as<uint32_t>();
    rs[s].tstat = trim_static.as<uint32_t>();
  }
  char *pin = (char *)h_pinned;
  std::memcpy(pin + PIN_REACH, rs, sizeof(rs));
  PSC_HIP(hipMemcpyAsync(reach_desc.ptr, pin + PIN_REACH, sizeof(rs),
                         hipMemcpyHostToDevice, stream));
  ReachSearch *S = reach_desc.as<ReachSearch>();
  ReachCtl *C = reach_ctl.as<ReachCtl>();
  prof_begin(KK_MISC);
  k_reach_reset<<<1, 64, 0, stream>>>(C, 2);
  if (!rr) {
    const uint64_t words = (std::max<uint64_t>(n, 1) + 31) / 32;
    k_reach_sources<<<grid_for((uint64_t)G * WAVE), BLOCK, 0, stream>>>(
        S, C, 2, trim_static.as<uint32_t>(), words, trim_off.as<unsigned long long>(),
        (std::max<uint64_t>(n, 1) + SEL_TILE - 1) / SEL_TILE, gA.as<unsigned long long>(),
        G, p.seed, m);
  } else {
    // visited bitmaps empty, every vertex excluded but those carrying
    // their graph's pivot label (the exclusion bitmap is also the pull
    // walk's skip set)
    const uint64_t words = (std::max<uint64_t>(n, 1) + 31) / 32;
    k_fill_u32<<<grid_for(words, BLOCK, 16384), BLOCK, 0, stream>>>(
        vis[0].as<uint32_t>(), words, 0u);
    k_fill_u32<<<grid_for(words, BLOCK, 16384), BLOCK, 0, stream>>>(
        vis[1].as<uint32_t>(), words, 0u);
    k_fill_u32<<<grid_for(words, BLOCK, 16384), BLOCK, 0, stream>>>(
        trim_bm.as<uint32_t>(), words, ~0u);
    k_fill_u64<<<grid_for(G), BLOCK, 0, stream>>>(rr_plab.as<unsigned long long>(), G,
                                                  EMPTY_SLOT);
    k_rr_plab<<<grid_for(rr->A), BLOCK, 0, stream>>>(
        rr->front, rr->nfront, gC.as<unsigned long long>(), rr->A, rr->wgid,
        labels.as<unsigned long long>(), rr_plab.as<unsigned long long>(), S);
    // few blocks: each adds its count to both searches' nnz with one atomic
    k_rr_include<<<grid_for(rr->n_rem, BLOCK, 1024), BLOCK, 0, stream>>>(
        centers.as<uint32_t>(), rr->n_rem, gB.as<unsigned long long>(), G,
        rr_plab.as<unsigned long long>(), labels.as<unsigned long long>(),
        trim_bm.as<uint32_t>(), S);
    k_reach_sources_front<<<grid_for(rr->A), BLOCK, 0, stream>>>(S, C, rr->front,
                                                                 rr->nfront);
    kbytes[KK_MISC] += 12.0 * words + 24.0 * rr->n_rem;
  }
  prof_end();
  const uint32_t total_waves = (uint32_t)grid_step * WPB;
  const uint32_t tau = (uint32_t)std::min<uint64_t>(p.tau ? p.tau : 1, 1u << 30);
  // the kernels' VGC argument: 0 = off, else a wave's serial batch cap
  const int vgc_reach = p.local_reach ? vgc_batch_cap(p) : 0;
  uint32_t L = 0;
  // the previous run's level count (same graph: one poll, no idle levels;
  // the poll itself shows that the level after the batch has no frontier).
  // Single-pivot rounds share one hint across rounds of different depths:
  // one spare level saves them a second poll.
  uint32_t &hint = rr ? round_levels_hint : reach_levels_hint;
  const uint32_t hint0 = hint;
  int K = hint ? (int)std::min<uint32_t>(hint + (rr ? 1 : 0), 64) : 4;
  uint32_t lvl[2] = {0, 0};
  // timing >= 3: one level per poll and a per-level report on stderr
  const bool trace = p.timing >= 3 && !rr;
  if (trace) K = 1;
  ReachCtl hc_, prev_{};
  auto t_lvl = std::chrono::steady_clock::now();
  while (true) {
    auto levels = [&](uint32_t L0) {
      for (int k = 0; k < K; ++k) {
        const uint32_t Lk = L0 + (uint32_t)k;
        prof_begin(KK_STEP);
        k_reach_step<<<grid_step, BLOCK, 0, stream>>>(S, C, 2, Lk, tau, vgc_reach,
                                                        total_waves);
        prof_end();
        prof_begin(KK_PULL);
        k_reach_pull<<<grid_pull, BLOCK, 0, stream>>>(S, C, 2, Lk,
                                                      (uint32_t)grid_pull * WPB);
        prof_end();
        prof_begin(KK_HEAVY);
        k_reach_heavy<<<grid_heavy, BLOCK, 0, stream>>>(S, C, 2, Lk,
                                                         (uint32_t)grid_heavy * WPB);
        prof_end();
      }
    };
    if (p.capture && !trace && prof_level < 2) {
      const BatchKey key{0, L, (uint32_t)K, tau, total_waves, vgc_reach, grid_pull,
                         grid_heavy, S, C, nullptr};
      PSC_TRY(replay(key, [&] { levels(key.L0); }));
    } else {
      levels(L);
    }
    L += (uint32_t)K;
    PSC_HIP(hipGetLastError());
    const bool ran = after_levels && !trace && hint0 && L >= hint0;
    if (ran) {
      // ahead of the poll: the labelling checks on the device that the
      // reach is over at level L
      label_gate = true;
      label_gate_level = L;
      const int e = after_levels();
      label_gate = false;
      PSC_TRY(e);
    }
    PSC_TRY(poll(&hc_, C, sizeof(ReachCtl)));
    if (trace) {
      const auto now = std::chrono::steady_clock::now();
      const double us = std::chrono::duration<double, std::micro>(now - t_lvl).count();
      t_lvl = now;
      const uint32_t lv = L - 1;
      for (int s = 0; s < 2; ++s) {
        const ReachSlot &c = hc_.slot[s][lv % 3];
        const unsigned long long *a = hc_.cnt[s], *b = prev_.cnt[s];
        const double ds = c.from_pull ? (double)c.qsize * rs[s].avg_deg : (double)c.degsum;
        const double push = (double)c.qsize + ds;
        const double unv = nnz_total > c.cum_vis ? (double)(nnz_total - c.cum_vis) : 0.0;
        const bool pull = c.from_pull == PULL_UNSTORED || push >= (double)thresh ||
                          unv * PULL_VERTEX_COST + (double)(n / PULL_WALK_DIV) < push;
        const char *mode = c.qsize == 0 ? "none" : (pull ? "pull" : "push");
        std::fprintf(stderr,
                     "level %u %s %s nf=%u degsum=%llu heavy=%llu/%llu next=%u "
                     "push_exp=%llu push_edges=%llu pull_scans=%llu pull_edges=%llu "
                     "head_tests=%llu heavy_edges=%llu wall=%.1fus\n",
                     lv, s ? "bwd" : "fwd", mode, c.qsize, c.degsum,
                     c.heavy >> H_CNT_SHIFT, c.heavy & H_CHUNK_MASK,
                     hc_.slot[s][(lv + 1) % 3].qsize, a[R_PUSH_EXP] - b[R_PUSH_EXP],
                     a[R_PUSH_EDGES] - b[R_PUSH_EDGES], a[R_PULL_SCANS] - b[R_PULL_SCANS],
                     a[R_PULL_EDGES] - b[R_PULL_EDGES], a[R_HEAD_TESTS] - b[R_HEAD_TESTS],
                     a[R_HEAVY_EDGES] - b[R_HEAVY_EDGES], us);
      }
      prev_ = hc_;
    }
    // a search is over once a level starts without frontier: marked by that
    // level's step kernel, or seen here in the slot of level L (the next)
    bool fin = true;
    for (int s = 0; s < 2; ++s) {
      if (hc_.done[s]) lvl[s] = hc_.levels[s];
      else if (hc_.slot[s][L % 3].qsize == 0 && hc_.slot[s][L % 3].deferred == 0) lvl[s] = L;
      else fin = false;
    }
    if (fin) {
      if (after_ran) *after_ran = ran;
      break;
    }
    // past a hinted batch the search is usually a level or two from its
    // end: a short batch, then doubling
    if (!trace) K = (hint0 && L == hint0) ? 2 : std::min(K * 2, 64);
  }
  if (!rr) {
    st->levels_fwd = lvl[0];
    st->levels_bwd = lvl[1];
  } else {
    st->sub_rounds += std::max(lvl[0], lvl[1]);
    for (int s = 0; s < 2; ++s)
      rr->reached[s] = hc_.cnt[s][R_PUSH_VIS] + hc_.cnt[s][R_HEAVY_VIS] +
                       hc_.cnt[s][R_PULL_VIS] + hc_.cnt[s][R_HEAD_VIS];
  }
  // the first reach's depth varies by a level from run to run (which
  // vertices a level finds depends on timing): the hint is the deeper of
  // the last two runs, so a run rarely outlasts its hinted batch
  {
    const uint32_t d = std::max(lvl[0], lvl[1]);
    uint32_t &prev = rr ? round_levels_prev : reach_levels_prev;
    hint = rr ? d : std::max(d, prev);
    prev = d;
  }
  for (int s = 0; s < 2; ++s) {
    const unsigned long long *c = hc_.cnt[s];
    st->edges_scanned += c[R_PUSH_EDGES] + c[R_HEAVY_EDGES] + c[R_PULL_EDGES] +
                         c[R_HEAD_TESTS];
    // push: offsets of expanded vertices, edge ids, visited test per edge,
    // visited write per new vertex, frontier read/append, degree of appended
    const double push = 16.0 * c[R_PUSH_EXP] + 5.0 * c[R_PUSH_EDGES] +
                        c[R_PUSH_VIS] + 4.0 * c[R_PUSH_EXP] + 20.0 * c[R_PUSH_APP];
    const double heavy = 5.0 * c[R_HEAVY_EDGES] + c[R_HEAVY_VIS] +
                         20.0 * c[R_HEAVY_APP] + 32.0 * c[R_HEAVY_ITEMS];
    // pull: (visited | trimmed) words of every 64-word chunk, in-offsets of
    // scanned vertices, in-edge ids + visited test per scanned edge, new
    // vertices
    const double pull = 512.0 * c[R_PULL_UNITS] + 16.0 * c[R_PULL_SCANS] +
                        5.0 * c[R_PULL_EDGES] + 5.0 * c[R_PULL_VIS];
    // head-only pull levels (run by the step kernel): the same words per
    // chunk, 4 B head + visited test per tested vertex, new vertices
    const double head = 512.0 * c[R_HEAD_UNITS] + 5.0 * c[R_HEAD_TESTS] +
                        5.0 * c[R_HEAD_VIS];
    kbytes[KK_STEP] += push + head;
    kbytes[KK_PULL] += pull;
    kbytes[KK_HEAVY] += heavy;
    st->bytes_csr += 16 * (c[R_PUSH_EXP] + c[R_PULL_SCANS]) +
                     4 * (c[R_PUSH_EDGES] + c[R_HEAVY_EDGES] + c[R_PULL_EDGES] +
                          c[R_HEAD_TESTS]);
    st->bytes_label += c[R_PUSH_EDGES] + c[R_HEAVY_EDGES] + c[R_PULL_EDGES] +
                       c[R_HEAD_TESTS] + c[R_PUSH_VIS] + c[R_HEAVY_VIS] + c[R_PULL_VIS] +
                       c[R_HEAD_VIS] + 512 * (c[R_PULL_UNITS] + c[R_HEAD_UNITS]);
    st->bytes_front += 4 * (c[R_PUSH_APP] + c[R_HEAVY_APP] + c[R_PULL_VIS] +
                            c[R_HEAD_VIS] + c[R_PUSH_EXP]) +
                       16 * (c[R_PUSH_APP] + c[R_HEAVY_APP]);
  }
  return PSC_OK;
}

// A single-pivot round (every window one position): fused reaches from the
// round's pivots inside their subproblems, then the bitmap partition.
int Engine::round_reach(const psc_params &p, uint32_t A, const uint32_t *wgid_d,
                        uint64_t n_rem, psc_stats *st) {
  ReachRound rr{front.as<uint32_t>(), misc.as<unsigned long long>() + 3, A, wgid_d,
                n_rem, {0, 0}};
  PSC_TRY(first_reach(p, st, &rr));
  prof_begin(KK_PARTITION);
  k_rr_mark<<<grid_for(n_rem, BLOCK, 16384), BLOCK, 0, stream>>>(
      centers.as<uint32_t>(), n_rem, gB.as<unsigned long long>(), G,
      rr_plab.as<unsigned long long>(), trim_bm.as<uint32_t>(), vis[0].as<uint32_t>(),
      vis[1].as<uint32_t>(), gD.as<unsigned long long>(), labels.as<unsigned long long>());
  k_rr_advance<<<1, BLOCK, 0, stream>>>(gC.as<unsigned long long>(), A, wgid_d,
                                        gD.as<unsigned long long>());
  prof_end();
  PSC_HIP(hipGetLastError());
  kbytes[KK_PARTITION] += 30.0 * n_rem;
  st->bytes_label += 16 * n_rem;
  log_n[0] = rr.reached[0];
  log_n[1] = rr.reached[1];
  return PSC_OK;
}

// Allocate / size the multi-search buffers of search s for `slots` slots.
int Engine::ms_prepare(int s, uint64_t slots, MSSearch *d) {
  const uint64_t log_cap = std::max<uint64_t>(slots / 2, 64);
  PSC_TRY(table[s].ensure(slots * 8));
  PSC_TRY(log[s].ensure(log_cap * 8));
  PSC_TRY(mq[s][0].ensure(log_cap * 8));
  PSC_TRY(mq[s][1].ensure(log_cap * 8));
  PSC_TRY(mhp[s].ensure(log_cap * 8));
  PSC_TRY(mhc[s].ensure(log_cap * 8));
  d->off = s == 0 ? d_off : d_ioff;
  d->adj = s == 0 ? d_adj : d_iadj;
  d->table = table[s].as<unsigned long long>();
  d->mask = slots - 1;
  d->log = log[s].as<unsigned long long>();
  d->log_cap = log_cap;
  d->q[0] = mq[s][0].as<unsigned long long>();
  d->q[1] = mq[s][1].as<unsigned long long>();
  d->q_cap = log_cap;
  d->hp = mhp[s].as<unsigned long long>();
  d->hc = mhc[s].as<unsigned long long>();
  // at load <= 1/2 linear probes are short; a run this long means the
  // table overflowed (staged log appends may overshoot the load limit)
  d->probe_limit = (uint32_t)std::min<uint64_t>(slots, 4096);
  d->pad = 0;
  return PSC_OK;
}

static uint64_t pow2_at_least(uint64_t x) {
  uint64_t m = 1;
  while (m < x) m <<= 1;
  return m;
}

// multi_search_safe for both directions (src/scc.hpp:256-275,395-423):
// searches run fused; an overflowed search is rerun alone with a doubled
// table.
int Engine::multi_search(const psc_params &p, uint64_t nfront, uint32_t A,
                         const uint32_t *wgid_d, uint64_t slots[2], psc_stats *st) {
  MSSearch md[2];
  std::memcpy(md, h_ms, sizeof(md));
  MSSearch *S = ms_desc.as<MSSearch>();
  MSCtl *C = ms_ctl.as<MSCtl>();
  int want[2] = {1, 1};
  const uint32_t tau = (uint32_t)std::min<uint64_t>(p.tau ? p.tau : 1, 1u << 30);
  const int vgc_ms = p.local_scc ? vgc_batch_cap(p) : 0;
  const uint32_t total_waves = (uint32_t)grid_ms * WPB;
  int attempt = 0;
  while (want[0] || want[1]) {
    const auto t_attempt = std::chrono::steady_clock::now();
    for (int s = 0; s < 2; ++s) {
      if (!want[s]) continue;
      while (slots[s] / 2 < nfront) {   // seeds alone exceed the load limit
        slots[s] *= 2;
        ++st->resizes;
      }
      PSC_TRY(ms_prepare(s, slots[s], &md[s]));
    }
    std::memcpy(h_ms, md, sizeof(md));
    // descriptors go through pinned staging, only when they changed; the
    // stream is synchronised (poll) before the staging area is written again
    if (!ms_dev_valid || std::memcmp(h_ms_dev, md, sizeof(md)) != 0) {
      char *pin = (char *)h_pinned;
      std::memcpy(pin + PIN_MS, md, sizeof(md));
      PSC_HIP(hipMemcpyAsync(S, pin + PIN_MS, sizeof(md), hipMemcpyHostToDevice,
                             stream));
      std::memcpy(h_ms_dev, md, sizeof(md));
      ms_dev_valid = true;
    }
    const uint32_t wmask = (want[0] ? 1u : 0u) | (want[1] ? 2u : 0u);
    const uint64_t s0 = want[0] ? md[0].mask + 1 : 0, s1 = want[1] ? md[1].mask + 1 : 0;
    prof_begin(KK_SEED);
    k_ms_clear<<<grid_for(std::max(s0, s1), BLOCK, 8192), BLOCK, 0, stream>>>(
        C, wmask, md[0].table, s0, md[1].table, s1);
    k_ms_begin<<<grid_for(nfront), BLOCK, 0, stream>>>(
        S, C, wmask, front.as<uint32_t>(), misc.as<unsigned long long>() + 3,
        gC.as<unsigned long long>(), A, wgid_d, gD.as<unsigned long long>());
    prof_end();
    kbytes[KK_SEED] += 8.0 * (s0 + s1) + (4.0 * nfront + 32.0 * nfront) *
                                             ((want[0] ? 1 : 0) + (want[1] ? 1 : 0));
    PSC_HIP(hipGetLastError());
    uint32_t L = 0;
    int K = (int)std::min<uint32_t>(std::max<uint32_t>(ms_levels_hint, 2), 32);
    MSCtl hc_;
    while (true) {
      auto levels = [&](uint32_t L0) {
        for (int k = 0; k < K; ++k) {
          const uint32_t Lk = L0 + (uint32_t)k;
          prof_begin(KK_MS_EXPAND);
          k_ms_expand<<<grid_ms, BLOCK, 0, stream>>>(
              S, C, 2, Lk, labels.as<unsigned long long>(), tau, vgc_ms, total_waves);
          prof_end();
          prof_begin(KK_MS_HEAVY);
          k_ms_heavy<<<grid_heavy, BLOCK, 0, stream>>>(
              S, C, 2, Lk, labels.as<unsigned long long>(),
              (uint32_t)grid_heavy * WPB);
          prof_end();
        }
      };
      if (p.capture && prof_level < 2) {
        const BatchKey key{1, L, (uint32_t)K, tau, total_waves, vgc_ms, grid_heavy,
                           0, S, C, labels.ptr};
        PSC_TRY(replay(key, [&] { levels(key.L0); }));
      } else {
        levels(L);
      }
      L += (uint32_t)K;
      PSC_HIP(hipGetLastError());
      PSC_TRY(poll(&hc_, C, sizeof(MSCtl)));
      bool live = false;
      for (int s = 0; s < 2; ++s)
        if (want[s] && !hc_.full[s] && hc_.slot[s][L % 3].qsize) live = true;
      if (!live) break;
      K = std::min(K * 2, 32);
    }
    st->sub_rounds += L;
    ms_levels_hint = L;
    for (int s = 0; s < 2; ++s) {
      if (!want[s]) continue;
      const unsigned long long *c = hc_.cnt[s];
      st->edges_scanned += c[M_EDGES] + c[M_HEAVY_EDGES];
      st->table_inserts += c[M_INS] + c[M_HEAVY_INS];
      // expand: pair read + offsets + labels[v] per expanded pair; edge id
      // + labels[w] per edge; 8 B per probed slot; CAS write, log append
      // and frontier append per new pair
      const double ex = 32.0 * c[M_EXP] + 12.0 * c[M_EDGES] +
                        8.0 * c[M_PROBES] + 16.0 * c[M_INS] + 8.0 * c[M_APP];
      const double hv = 12.0 * c[M_HEAVY_EDGES] + 8.0 * c[M_HEAVY_PROBES] +
                        16.0 * c[M_HEAVY_INS] + 8.0 * c[M_HEAVY_APP] +
                        40.0 * c[M_HEAVY_ITEMS];
      kbytes[KK_MS_EXPAND] += ex;
      kbytes[KK_MS_HEAVY] += hv;
      st->bytes_csr += 16 * c[M_EXP] + 4 * (c[M_EDGES] + c[M_HEAVY_EDGES]);
      st->bytes_label += 8 * (c[M_EXP] + c[M_EDGES] + c[M_HEAVY_EDGES]);
      st->bytes_table += 8 * (c[M_PROBES] + c[M_HEAVY_PROBES] + c[M_INS] +
                              c[M_HEAVY_INS]);
      st->bytes_front += 8 * (c[M_EXP] + c[M_INS] + c[M_HEAVY_INS] + c[M_APP] +
                              c[M_HEAVY_APP]);
      if (hc_.full[s]) {
        slots[s] *= 2;   // double_size (src/resizable_table.h:196-204)
        ++st->resizes;
        want[s] = 1;
      } else {
        log_n[s] = hc_.log_n[s];
        want[s] = 0;
      }
    }
    // reruns after an overflow are the "hash table resize time" of
    // src/scc.hpp:112-113
    if (attempt++ > 0)
      st->ms_resize += std::chrono::duration<double, std::milli>(
                           std::chrono::steady_clock::now() - t_attempt).count();
  }
  return PSC_OK;
}

int Engine::finalize_labels(const unsigned long long *go) {
  if (direct_final) {
    k_finalize_rest<<<grid_for(final_rem + G, BLOCK, 16384), BLOCK, 0, stream>>>(
        labels.as<unsigned long long>(), centers.as<uint32_t>(), final_rem,
        first_trim.as<uint32_t>(), gA.as<unsigned long long>() + (G + 1), G, go);
  } else {
    k_finalize<<<grid_for(n, BLOCK, 16384), BLOCK, 0, stream>>>(
        labels.as<unsigned long long>(), n, go);
  }
  return PSC_OK;
}

int Engine::run(const psc_params &p, unsigned long long *d_out, psc_stats *st) {
  *st = psc_stats{};
  syncs = 0;
  finalized = false;
  direct_final = false;
  final_rem = 0;
  PSC_TRY(prof_reset(p.timing));
  unsigned long long *lab = labels.as<unsigned long long>();
  unsigned long long *voffd = voff_d.as<unsigned long long>();
  auto t_host0 = std::chrono::steady_clock::now();
  if (p.timing) PSC_HIP(hipEventRecord(ev[0], stream));
  // ---- init + trim, per graph (src/scc.hpp:284-306): one pass counts the
  // non-trimmed vertices per tile and writes the trimmed and dead-end
  // bitmaps; NZ (the non-trimmed vertices in order) is never listed: the
  // first pivot is found by rank, every label is written after the reach
  // the binding trimmed every graph (trim_static, dead, trim_off and the
  // per-graph starts); a run starts its searches' visited bitmaps as the
  // dead-end sets and clears its probe counter
  unsigned long long *nstart = gA.as<unsigned long long>();
  unsigned long long *zstart = nstart + (G + 1);
  if (n) {
    const uint64_t words = (n + 31) / 32;
    prof_begin(KK_MISC);
    k_run_start<<<grid_for(words, BLOCK, 4096), BLOCK, 0, stream>>>(
        dead.as<uint32_t>(), words, vis[0].as<uint32_t>(), vis[1].as<uint32_t>(),
        misc.as<unsigned long long>() + 2);
    prof_end();
    kbytes[KK_MISC] += 16.0 * words;
    st->bytes_front += 16 * words;
    PSC_HIP(hipGetLastError());
  }
  if (p.timing) PSC_HIP(hipEventRecord(ev[1], stream));
  bool counted = false;
  nnz_total = starts_h[G];
  const uint64_t *h_nstart = starts_h.data(), *h_zstart = starts_h.data() + G + 1;
  // per-graph label counters: trimmed get 1..nz_g, the first SCC nz_g + 1,
  // the rounds' labels start at nz_g + 2
  std::vector<uint64_t> label_off(G), label_next(G);
  auto set_label_offs = [&]() {
    for (uint32_t g = 0; g < G; ++g) {
      label_off[g] = h_zstart[g + 1] - h_zstart[g] + 1;
      label_next[g] = label_off[g] + 1;
    }
  };
  // The labelling after the first reach: every label (trimmed ranks,
  // first-SCC and subproblem labels) in one vertex-order pass, the
  // remaining vertices placed by pivot position, the position bitmap
  // compacted in order, and the per-graph starts of the remaining vertices
  // copied back.  It needs only the trim counts from the host, so it is
  // enqueued behind the batch of reach levels that is expected to end the
  // reach, before that batch's poll: the poll returns with it done.  If the
  // reach went on, it is enqueued again after the last batch (it rewrites
  // everything it writes; the trim's tile offsets live in their own buffer).
  auto label_pass = [&]() -> int {
    const uint64_t nnz = h_nstart[G];
    set_label_offs();
    // final values go straight into the labels unless a graph with
    // non-trimmed vertices has no trimmed one (its first SCC's final value
    // would be 0, which an unreached vertex's label also is)
    direct_final = true;
    for (uint32_t g = 0; g < G; ++g)
      if (h_nstart[g + 1] > h_nstart[g] && h_zstart[g + 1] == h_zstart[g])
        direct_final = false;
    // one graph: its two offsets go to k_first_label as arguments
    if (G > 1) {
      PSC_TRY(upload(gC, label_off.data(), G, 5));
      PSC_TRY(upload(gD, label_next.data(), G, 4));
    }
    const uint64_t pwords = (std::max<uint64_t>(nnz, 1) + 31) / 32;
    PSC_TRY(pbits.ensure(pwords * 4));
    uint32_t *posv = front.as<uint32_t>();   // front is free until the rounds
    const uint64_t nt = (n + SEL_TILE - 1) / SEL_TILE;
    // the remaining vertices are counted per tile of positions while they
    // are placed (the select's counts), into `tiles`, zeroed with the
    // position bitmap
    const uint64_t st_tiles = (std::max<uint64_t>(nnz, 1) + SEL_TILE - 1) / SEL_TILE;
    prof_begin(KK_MISC);
    k_zero2<<<grid_for(pwords + st_tiles, BLOCK, 4096), BLOCK, 0, stream>>>(
        pbits.as<uint32_t>(), pwords, tiles.as<uint32_t>(), st_tiles);
    k_first_label<<<(unsigned)nt, BLOCK, 0, stream>>>(
        trim_static.as<uint32_t>(), n, trim_off.as<unsigned long long>(), voffd, G, zstart,
        nstart, p.seed, vis[0].as<uint32_t>(), vis[1].as<uint32_t>(), lab,
        G > 1 ? gC.as<unsigned long long>() : nullptr, posv, pbits.as<uint32_t>(),
        direct_final ? 1 : 0, first_trim.as<uint32_t>(), tiles.as<uint32_t>(),
        label_gate ? reach_ctl.as<ReachCtl>() : nullptr, label_gate_level, label_off[0],
        label_next[0], gC.as<unsigned long long>(), gD.as<unsigned long long>());
    prof_end();
    PSC_HIP(hipGetLastError());
    if (nnz) {
      // the remaining count stays on the device: the per-graph starts
      // (cstart[G] = the count) come back with the poll
      unsigned long long *rem_d = misc.as<unsigned long long>();
      prof_begin(KK_SELECT);
      if (st_tiles <= TSCAN_MAX) {
        k_tile_scan<<<1, TSCAN_BLOCK, 0, stream>>>(tiles.as<uint32_t>(), st_tiles,
                                                   tile_off.as<unsigned long long>(), rem_d,
                                                   G == 1 ? gB.as<unsigned long long>() : nullptr);
      } else {
        PSC_TRY(scan_tiles(st_tiles, rem_d));
        if (G == 1) k_segment_starts<<<1, 64, 0, stream>>>(centers.as<uint32_t>(), 0, rem_d, voffd,
                                                         G, nullptr, G, gB.as<unsigned long long>());
      }
      k_sel_bits<<<(unsigned)((st_tiles + WPB - 1) / WPB), BLOCK, 0, stream>>>(
          pbits.as<uint32_t>(), posv, nnz, tile_off.as<unsigned long long>(),
          centers.as<uint32_t>());
      prof_end();
      if (G > 1) {
        prof_begin(KK_MISC);
        k_segment_starts<<<grid_for(G + 1), BLOCK, 0, stream>>>(
            centers.as<uint32_t>(), 0, rem_d, voffd, G, nullptr, G,
            gB.as<unsigned long long>());
        prof_end();
      }
      PSC_HIP(hipMemcpyAsync(region(3), gB.ptr, 8 * (G + 1), hipMemcpyDeviceToHost,
                             stream));
    }
    if (p.timing) PSC_HIP(hipEventRecord(ev[2], stream));
    if (!counted) {
      kbytes[KK_MISC] += 8.0 * n + 3 * n / 8.0 + (nnz + 31) / 8.0;
      kbytes[KK_SELECT] += 2 * nnz / 8.0;
      st->bytes_label += 8 * n;
      counted = true;
    }
    return PSC_OK;
  };
  if (n) {
    // ---- first SCC of every graph (src/scc.hpp:316-355)
    bool labelled = false;
    PSC_TRY(first_reach(p, st, nullptr, label_pass, &labelled));
    if (!labelled) {
      PSC_TRY(label_pass());
      PSC_HIP(hipStreamSynchronize(stream));
      ++syncs;
    }
  } else {
    set_label_offs();
    PSC_TRY(upload(gD, label_next.data(), G, 4));
    if (p.timing) PSC_HIP(hipEventRecord(ev[2], stream));
  }
  const uint64_t nnz = h_nstart[G], nz = n - nnz;
  nnz_total = nnz;
  st->n_trimmed = nz;
  uint64_t n_rem = 0;
  std::vector<uint64_t> h_cstart(G + 1, 0);
  if (nnz) {
    std::memcpy(h_cstart.data(), region(3), 8 * (G + 1));
    n_rem = h_cstart[G];
    final_rem = n_rem;
    kbytes[KK_SELECT] += 8.0 * n_rem;
  }
  // ---- multi-pivot rounds (src/scc.hpp:369-473), every graph on its own
  // pivot-batch schedule; the rounds of all graphs advance together
  std::vector<uint64_t> cur(G, 0), stepg(G, 1);
  std::vector<uint64_t> wpre, wsrc;
  std::vector<uint32_t> wgid;
  uint64_t ne_prev[2] = {1, 1};
  const double beta_def = p.beta > 0 ? p.beta : 1.5;
  double beta_max = beta_def;
  for (double b : beta_g) beta_max = std::max(beta_max, b);
  uint64_t live = 0;
  for (uint32_t g = 0; g < G; ++g) live += (h_cstart[g + 1] > h_cstart[g]);
  if (live && p.tail_max) {
    // small remainders: all their rounds in one launch (tail.hip)
    bool used = false;
    PSC_TRY(ms_tail(p, h_cstart.data(), n_rem, cur, stepg, st, &used));
    if (used) {
      live = 0;
      for (uint32_t g = 0; g < G; ++g) live += cur[g] < h_cstart[g + 1] - h_cstart[g];
    }
  }
  uint32_t tail_wait = 0, tail_backoff = 1;
  while (live) {
    wpre.assign(1, 0);
    wsrc.clear();
    wgid.clear();
    for (uint32_t g = 0; g < G; ++g) {
      const uint64_t rem = h_cstart[g + 1] - h_cstart[g];
      if (cur[g] == rem) continue;
      const uint64_t end = std::min(cur[g] + stepg[g], rem);
      const double b = beta_g.empty() ? beta_def : beta_g[g];
      stepg[g] = (uint64_t)std::ceil((double)stepg[g] * b);
      wsrc.push_back(h_cstart[g] + cur[g]);
      wpre.push_back(wpre.back() + (end - cur[g]));
      wgid.push_back(g);
      cur[g] = end;
      if (cur[g] == rem) --live;
    }
    const uint32_t A = (uint32_t)wgid.size();
    st->rounds++;
    // timing >= 3: one line per host-driven round on stderr
    const auto t_round = std::chrono::steady_clock::now();
    const uint64_t sub0 = st->sub_rounds, ins0 = st->table_inserts,
                   edg0 = st->edges_scanned;
    PSC_TRY(upload_round(wpre, wsrc, wgid));
    const unsigned long long *d_wpre = rnd.as<unsigned long long>();
    const unsigned long long *d_wsrc = d_wpre + (A + 1);
    const uint32_t *d_wgid = reinterpret_cast<const uint32_t *>(d_wsrc + A);
    // front = not-done pivots of every window, grouped by graph, and its
    // per-window starts (gC); its size stays on the device (misc[3]) -- one
    // host sync per round, in multi_search
    unsigned long long *nfront_d = misc.as<unsigned long long>() + 3;
    const uint64_t V = wpre[A];
    const SelWindow selw{centers.as<uint32_t>(), d_wpre, d_wsrc, A, lab};
    if (V <= SMALL_SELECT) {
      prof_begin(KK_SELECT);
      k_sel_small<SelWindow><<<1, SMALL_BLOCK, 0, stream>>>(
          selw, V, front.as<uint32_t>(), nfront_d, d_wpre, A,
          gC.as<unsigned long long>());
      prof_end();
    } else {
      PSC_TRY(select(selw, V, front.as<uint32_t>(), nullptr, nullptr, nfront_d));
      prof_begin(KK_MISC);
      k_segment_starts<<<grid_for(A + 1), BLOCK, 0, stream>>>(
          front.as<uint32_t>(), 0, nfront_d, voffd, G, d_wgid, A,
          gC.as<unsigned long long>());
      prof_end();
    }
    kbytes[KK_SELECT] += 2 * 12.0 * V;
    if (p.pivot_reach && V == A) {
      // one position per window: at most one pivot per graph
      PSC_TRY(round_reach(p, A, d_wgid, n_rem, st));
    } else {
      uint64_t slots[2];
      for (int s = 0; s < 2; ++s) {
        const uint64_t cap = 6000000ull * G;
        const uint64_t base = std::max<uint64_t>(
            std::min<uint64_t>((uint64_t)std::ceil(0.3 * (double)n_rem), cap),
            (uint64_t)beta_def * ne_prev[s]);
        slots[s] = pow2_at_least(p.table_init ? p.table_init : 2 * base);
      }
      PSC_TRY(multi_search(p, V, A, d_wgid, slots, st));
      const int small = (slots[0] <= slots[1]) ? 0 : 1;
      prof_begin(KK_PARTITION);
      k_partition<<<grid_for(log_n[0] + log_n[1], BLOCK, 2048), BLOCK, 0, stream>>>(
          ms_desc.as<MSSearch>(), ms_ctl.as<MSCtl>(), small, lab,
          misc.as<unsigned long long>() + 2, gC.as<unsigned long long>(), A, d_wgid,
          gD.as<unsigned long long>());
      prof_end();
      PSC_HIP(hipGetLastError());
      // log reads + label read-modify-write per pair; contains() probes of
      // the small log's pairs are added after the run from misc[2]
      kbytes[KK_PARTITION] += 24.0 * (log_n[0] + log_n[1]);
      st->bytes_label += 16 * (log_n[0] + log_n[1]);
      st->bytes_front += 8 * (log_n[0] + log_n[1]);
    }
    ne_prev[0] = log_n[0];
    ne_prev[1] = log_n[1];
    const uint64_t round_edges = st->edges_scanned - edg0;
    if (p.timing >= 3) {
      PSC_HIP(hipStreamSynchronize(stream));
      std::fprintf(stderr,
                   "round %llu: windows %u pivots %llu levels %llu pairs %llu/%llu "
                   "inserts %llu edges %llu %.1f us\n",
                   (unsigned long long)st->rounds, A, (unsigned long long)V,
                   (unsigned long long)(st->sub_rounds - sub0), (unsigned long long)log_n[0],
                   (unsigned long long)log_n[1],
                   (unsigned long long)(st->table_inserts - ins0),
                   (unsigned long long)(st->edges_scanned - edg0),
                   std::chrono::duration<double, std::micro>(
                       std::chrono::steady_clock::now() - t_round).count());
    }
    // Remainders whose rounds turned light (typically right after the round
    // that found their largest SCC) go back to the single-launch rounds.
    // A return that advances no graph waits twice as many rounds before
    // the next try.
    if (live && p.tail_max && (double)round_edges * beta_max <= TAIL_ROUND_EDGES &&
        ++tail_wait >= tail_backoff) {
      uint64_t before = 0, after = 0;
      for (uint32_t g = 0; g < G; ++g) before += cur[g];
      bool used = false;
      PSC_TRY(ms_tail(p, h_cstart.data(), n_rem, cur, stepg, st, &used));
      tail_wait = 0;
      for (uint32_t g = 0; g < G; ++g) after += cur[g];
      if (after == before) tail_backoff *= 2;
      if (used) {
        live = 0;
        for (uint32_t g = 0; g < G; ++g) live += cur[g] < h_cstart[g + 1] - h_cstart[g];
      }
    }
  }
  if (p.timing && !finalized) PSC_HIP(hipEventRecord(ev[3], stream));
  // ---- finalize (src/scc.hpp:475-476), unless the remainder rounds'
  // launch already enqueued it
  prof_begin(KK_MISC);
  if (!finalized) PSC_TRY(finalize_labels(nullptr));
  kbytes[KK_MISC] += 16.0 * (direct_final ? final_rem + G : n);
  if (p.canonical) {
    uint32_t *minv = scratch32.as<uint32_t>();
    k_fill_u32<<<grid_for(n + G, BLOCK, 16384), BLOCK, 0, stream>>>(
        minv, n + G, 0xFFFFFFFFu);
    k_canon_min<<<grid_for(n, BLOCK, 16384), BLOCK, 0, stream>>>(lab, n, voffd, G, minv);
    k_canon_apply<<<grid_for(n, BLOCK, 16384), BLOCK, 0, stream>>>(lab, n, voffd, G, minv);
    kbytes[KK_MISC] += 4.0 * n + 12.0 * n + 20.0 * n;
  }
  prof_end();
  st->bytes_label += 16 * n;
  if (d_out != lab && n)
    PSC_HIP(hipMemcpyAsync(d_out, lab, n * 8, hipMemcpyDeviceToDevice, stream));
  PSC_HIP(hipGetLastError());
  if (p.timing) PSC_HIP(hipEventRecord(ev[4], stream));
  // the contains() probe count rides on the run's final sync; the first
  // SCC holds every non-trimmed vertex that does not remain
  unsigned long long *probes = (unsigned long long *)((char *)h_pinned + PIN_POLL);
  PSC_HIP(hipMemcpyAsync(probes, misc.as<unsigned long long>() + 2, 8,
                         hipMemcpyDeviceToHost, stream));
  PSC_HIP(hipStreamSynchronize(stream));
  ++syncs;
  st->first_scc_size = nnz - n_rem;
  kbytes[KK_PARTITION] += 8.0 * probes[0];
  st->bytes_table += 8 * probes[0];
  if (p.timing) {
    float a = 0;
    PSC_HIP(hipEventElapsedTime(&a, ev[0], ev[4]));
    st->ms_total = a;
    PSC_HIP(hipEventElapsedTime(&a, ev[0], ev[1]));
    st->ms_init = a;
    PSC_HIP(hipEventElapsedTime(&a, ev[1], ev[2]));
    st->ms_first = a;
    PSC_HIP(hipEventElapsedTime(&a, ev[2], ev[3]));
    st->ms_multi = a;
    PSC_HIP(hipEventElapsedTime(&a, ev[3], ev[4]));
    st->ms_finalize = a;
  } else {
    st->ms_total = std::chrono::duration<double, std::milli>(
                       std::chrono::steady_clock::now() - t_host0).count();
  }
  PSC_TRY(prof_collect());
  st->host_syncs = syncs;
  return PSC_OK;
}

// ---------------------------------------------------------------- profiler
// timing >= 2: one event pair around every launch of the path, summed per
// kernel kind after the run (the per-launch average rocprofv3 reports).
int Engine::prof_reset(int level) {
  prof_level = level;
  pev_used = 0;
  pkind.clear();
  for (int k = 0; k < KK_N; ++k) {
    kprof[k] = KProf{};
    kbytes[k] = 0;
  }
  // the contains() probe counter (misc[2]) is cleared by the run's first
  // kernel (k_run_start); an empty graph has no kernel
  if (n == 0) (void)hipMemsetAsync(misc.as<unsigned long long>() + 2, 0, 8, stream);
  return PSC_OK;
}

void Engine::prof_begin(int kind) {
  if (prof_level < 2) return;
  if (pev_used + 2 > pev.size()) {
    GateGuard gate;
    for (int i = 0; i < 256; ++i) {
      hipEvent_t e;
      if (hipEventCreate(&e) != hipSuccess) return;
      pev.push_back(e);
    }
  }
  (void)hipEventRecord(pev[pev_used], stream);
  pkind.push_back(kind);
}

void Engine::prof_end() {
  if (prof_level < 2 || pkind.size() * 2 != pev_used + 2) return;
  (void)hipEventRecord(pev[pev_used + 1], stream);
  pev_used += 2;
}

int Engine::prof_collect() {
  for (int k = 0; k < KK_N; ++k) kprof[k].bytes = kbytes[k];
  if (prof_level < 2) return PSC_OK;
  for (size_t i = 0; i < pkind.size() && 2 * i + 1 < pev_used; ++i) {
    float a = 0;
    PSC_HIP(hipEventElapsedTime(&a, pev[2 * i], pev[2 * i + 1]));
    kprof[pkind[i]].ms += a;
    kprof[pkind[i]].launches += 1;
  }
  return PSC_OK;
}

const char *const KK_NAMES[KK_N] = {
    "k_reach_step", "k_reach_pull", "k_reach_heavy", "k_ms_expand", "k_ms_heavy",
    "k_partition", "select", "k_ms_clear+begin", "tail (compact + k_ms_tail)",
    "init/trim/finalize"};

int Engine::status(const unsigned long long *d_labels, uint64_t *n_scc,
                   uint64_t *largest) {
  uint32_t *cnt = scratch32.as<uint32_t>();
  unsigned long long *acc = misc.as<unsigned long long>() + 8;
  PSC_HIP(hipMemsetAsync(acc, 0, 32, stream));
  k_fill_u32<<<grid_for(n + 1, BLOCK, 16384), BLOCK, 0, stream>>>(cnt, n + 1, 0u);
  k_count_labels<<<grid_for(n, BLOCK, 16384), BLOCK, 0, stream>>>(
      d_labels, n, cnt, acc + 2);
  k_status_reduce<<<grid_for(n + 1, BLOCK, 4096), BLOCK, 0, stream>>>(
      cnt, n + 1, acc, (uint32_t *)(acc + 1));
  PSC_HIP(hipGetLastError());
  unsigned long long h[3];
  PSC_TRY(poll(h, acc, sizeof(h)));
  if (h[2]) {
    set_error("psc_status_device: labels outside [0, n]");
    return PSC_ERR_INVALID;
  }
  *n_scc = h[0];
  *largest = (uint32_t)h[1];
  return PSC_OK;
}

}  // namespace psc

// ======================================================================
//                                C ABI
// ======================================================================
using psc::Engine;

extern "C" {

void psc_params_default(psc_params *p) {
  if (!p) return;
  std::memset(p, 0, sizeof(*p));
  p->beta = 1.5;           // src/scc.cpp:40
  p->local_reach = 1;      // the experiments pass -local_reach -local_scc
  p->local_scc = 1;        // (scripts/Experiment1.py:29)
  p->tau = 512;            // src/scc.cpp:46
  p->min_bag = 1024;       // src/scc.cpp:45
  p->dense_div = 20;       // src/reach.hpp:49
  p->seed = 1;
  p->table_init = 0;
  p->canonical = 0;
  p->timing = 1;
  p->tail_max = 65536;
  p->pivot_reach = 1;
  p->capture = 1;
  p->vgc_batches = 0;
}

const char *psc_last_error(void) { return psc::g_last_error.c_str(); }

static int validate_host_graph(const psc_graph *g) {
  if (!g) return PSC_ERR_INVALID;
  if (g->n >= 0xFFFFFFFFull) {
    psc::set_error("n must be < 2^32-1 (UINT_N_MAX is the empty key)");
    return PSC_ERR_INVALID;
  }
  if (g->n && (!g->offset || !g->in_offset)) {
    psc::set_error("missing offset arrays");
    return PSC_ERR_INVALID;
  }
  if (g->m && (!g->edges || !g->in_edges)) {
    psc::set_error("missing edge arrays");
    return PSC_ERR_INVALID;
  }
  return PSC_OK;
}

int psc_create(const psc_graph *g, psc_handle **out) {
  psc::GateGuard gate;
  if (!out) return PSC_ERR_INVALID;
  *out = nullptr;
  int r = validate_host_graph(g);
  if (r) return r;
  if (g->n) {
    if (g->offset[0] != 0 || g->offset[g->n] != g->m || g->in_offset[0] != 0 ||
        g->in_offset[g->n] != g->m) {
      psc::set_error("offsets must start at 0 and end at m");
      return PSC_ERR_INVALID;
    }
    for (uint64_t i = 0; i < g->m; ++i)
      if (g->edges[i] >= g->n || g->in_edges[i] >= g->n) {
        psc::set_error("edge endpoint out of range");
        return PSC_ERR_INVALID;
      }
  }
  auto *h = new (std::nothrow) psc_handle;
  if (!h) return PSC_ERR_NOMEM;
  Engine &e = h->eng;
  const uint64_t n1 = g->n + 1;
  r = e.setup();
  if (!r) r = e.g_off.ensure(n1 * 8);
  if (!r) r = e.g_ioff.ensure(n1 * 8);
  if (!r) r = e.g_adj.ensure(std::max<uint64_t>(g->m, 1) * 4);
  if (!r) r = e.g_iadj.ensure(std::max<uint64_t>(g->m, 1) * 4);
  if (r) {
    for (psc::DevBuf *b : {&e.g_off, &e.g_adj, &e.g_ioff, &e.g_iadj}) b->release();
    delete h;
    return r;
  }
  hipError_t he = hipSuccess;
  const uint64_t z = 0;
  const hipMemcpyKind h2d = hipMemcpyHostToDevice;
  he = hipMemcpyAsync(e.g_off.ptr, g->n ? g->offset : &z, (g->n ? n1 : 1) * 8, h2d,
                      e.stream);
  if (he == hipSuccess)
    he = hipMemcpyAsync(e.g_ioff.ptr, g->n ? g->in_offset : &z, (g->n ? n1 : 1) * 8,
                        h2d, e.stream);
  if (he == hipSuccess && g->m)
    he = hipMemcpyAsync(e.g_adj.ptr, g->edges, g->m * 4, h2d, e.stream);
  if (he == hipSuccess && g->m)
    he = hipMemcpyAsync(e.g_iadj.ptr, g->in_edges, g->m * 4, h2d, e.stream);
  if (he == hipSuccess) he = hipStreamSynchronize(e.stream);
  if (he != hipSuccess) {
    psc::set_error(std::string("hipMemcpy graph: ") + hipGetErrorString(he));
    delete h;
    return PSC_ERR_HIP;
  }
  psc_graph dg{g->n, g->m, e.g_off.as<uint64_t>(), e.g_adj.as<uint32_t>(),
               e.g_ioff.as<uint64_t>(), e.g_iadj.as<uint32_t>()};
  r = e.init(&dg, true);
  if (r) {
    delete h;
    return r;
  }
  *out = h;
  return PSC_OK;
}

int psc_create_device(const psc_graph *dg, psc_handle **out) {
  psc::GateGuard gate;
  if (!out) return PSC_ERR_INVALID;
  *out = nullptr;
  int r = validate_host_graph(dg);
  if (r) return r;
  auto *h = new (std::nothrow) psc_handle;
  if (!h) return PSC_ERR_NOMEM;
  r = h->eng.init(dg, false);
  if (r) {
    delete h;
    return r;
  }
  *out = h;
  return PSC_OK;
}

static int run_common(psc_handle *h, const psc_params *p,
                      unsigned long long *d_out, psc_stats *stats) {
  psc_params def;
  psc_params_default(&def);
  if (!p) p = &def;
  if (p->beta < 1.0) {
    psc::set_error("beta must be >= 1");
    return PSC_ERR_INVALID;
  }
  psc_stats tmp;
  return h->eng.run(*p, d_out, stats ? stats : &tmp);
}

int psc_run(psc_handle *h, const psc_params *p, uint64_t *labels_out,
            psc_stats *stats) {
  if (!h) return PSC_ERR_INVALID;
  Engine &e = h->eng;
  int r = run_common(h, p, e.labels.as<unsigned long long>(), stats);
  if (r) return r;
  if (labels_out && e.n) return e.download_labels(labels_out);
  return PSC_OK;
}

int psc_run_device(psc_handle *h, const psc_params *p, uint64_t *d_labels_out,
                   psc_stats *stats) {
  if (!h || (!d_labels_out && h->eng.n)) return PSC_ERR_INVALID;
  return run_common(h, p, (unsigned long long *)d_labels_out, stats);
}

int psc_status_device(psc_handle *h, const uint64_t *d_labels, uint64_t *n_scc,
                      uint64_t *largest) {
  if (!h || !n_scc || !largest) return PSC_ERR_INVALID;
  if (h->eng.n == 0) {
    *n_scc = *largest = 0;
    return PSC_OK;
  }
  return h->eng.status((const unsigned long long *)d_labels, n_scc, largest);
}

int psc_kernel_profile(psc_handle *h, psc_kernel_prof *out, int cap) {
  if (!h || (!out && cap > 0)) return PSC_ERR_INVALID;
  int k = 0;
  for (; k < psc::KK_N && k < cap; ++k) {
    std::memset(&out[k], 0, sizeof(out[k]));
    std::strncpy(out[k].name, psc::KK_NAMES[k], sizeof(out[k].name) - 1);
    out[k].launches = h->eng.kprof[k].launches;
    out[k].total_ms = h->eng.kprof[k].ms;
    out[k].bytes = h->eng.kprof[k].bytes;
  }
  return k;
}

void psc_destroy(psc_handle *h) {
  psc::GateGuard gate;
  delete h;
}

int psc_device_alloc(size_t bytes, void **ptr) {
  psc::GateGuard gate;
  if (!ptr) return PSC_ERR_INVALID;
  *ptr = nullptr;
  hipError_t e = hipMalloc(ptr, bytes ? bytes : 1);
  if (e != hipSuccess) {
    psc::set_error(std::string("hipMalloc: ") + hipGetErrorString(e));
    return PSC_ERR_NOMEM;
  }
  return PSC_OK;
}

int psc_device_free(void *ptr) {
  psc::GateGuard gate;
  hipError_t e = hipFree(ptr);
  if (e != hipSuccess) {
    psc::set_error(std::string("hipFree: ") + hipGetErrorString(e));
    return PSC_ERR_HIP;
  }
  return PSC_OK;
}

uint64_t psc_pivot_perm(uint64_t n, uint64_t seed, uint64_t i, int inverse) {
  if (n == 0 || i >= n) return 0;
  const psc::PivotPerm perm(n, seed);
  return inverse ? perm.inverse(i) : perm(i);
}

int psc_host_alloc(size_t bytes, void **ptr) {
  psc::GateGuard gate;
  if (!ptr) return PSC_ERR_INVALID;
  *ptr = nullptr;
  hipError_t e = hipHostMalloc(ptr, bytes ? bytes : 1);
  if (e != hipSuccess) {
    *ptr = nullptr;
    psc::set_error(std::string("hipHostMalloc: ") + hipGetErrorString(e));
    return PSC_ERR_NOMEM;
  }
  return PSC_OK;
}

int psc_host_free(void *ptr) {
  psc::GateGuard gate;
  hipError_t e = hipHostFree(ptr);
  if (e != hipSuccess) {
    psc::set_error(std::string("hipHostFree: ") + hipGetErrorString(e));
    return PSC_ERR_HIP;
  }
  return PSC_OK;
}

int psc_device_copy(void *dst, const void *src, size_t bytes) {
  psc::GateGuard gate;
  hipError_t e = hipMemcpy(dst, src, bytes, hipMemcpyDefault);
  if (e != hipSuccess) {
    psc::set_error(std::string("hipMemcpy: ") + hipGetErrorString(e));
    return PSC_ERR_HIP;
  }
  return PSC_OK;
}

void *psc_handle_stream(psc_handle *h) { return h ? (void *)h->eng.stream : nullptr; }

}  // extern "C"
