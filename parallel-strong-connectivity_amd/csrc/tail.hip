// Multi-pivot rounds of small remainders in one launch (src/scc.hpp:369-473).
//
// After the first SCC, what is left of a graph is often small (a few
// thousand vertices on RMAT).  Host-driven rounds then cost a host sync and
// a dozen launches per round while the device does microseconds of work.
// Here every graph whose remainder is small gets one workgroup that runs
// all of its rounds: windows of the pivot order, the forward and backward
// multi-searches over (vertex, label) pairs, the partition -- the same
// steps and the same pair sets as the host-driven rounds, so the labels are
// identical.  The searches run on a compact copy of the remainder: its
// vertices renumbered by pivot position, and only the edges whose ends are
// both unfinished and in the same subproblem.  Labels only split
// subproblems, never merge them, so a dropped edge could never be followed
// by a later round either.
//
// A graph whose pair tables, logs or frontiers overflow stops at the start
// of that round (its labels and schedule are written back as of the last
// completed round) and is relaunched with larger tables.  A graph whose
// rounds grow heavy (TAIL_ROUND_EDGES) stops after a completed round and
// the host-driven rounds, which use the whole device, finish it.
#include <hip/hip_runtime.h>

#include <cstring>
#include <rocprim/device/device_scan.hpp>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <string>
#include <vector>

#include "device_common.hpp"
#include "engine.hpp"
#include "scc_kernels.hpp"

namespace psc {

namespace {

// Workgroup of a remainder: half of its waves expand the forward search,
// half the backward one.  Windows are scanned in parts of 32 positions a
// thread (one bit mask).  A remainder small enough to live in LDS (mode 2
// below) runs with four waves (one workgroup per CU, LDS-bound): its rounds
// are short chains of LDS steps (measured on RMAT 2^25's remainder: 111 us
// with two waves, 100 with four, 103 with eight, 163 with one).
constexpr int TAIL_BLOCK = 1024;
constexpr int TAIL_SMALL_BLOCK = 256;
// Remainder positions (finished vertices included) a graph may have here:
// every round scans its window's positions.
constexpr uint64_t TAIL_MAX_POS = 1ull << 20;
constexpr uint64_t TAIL_MAX_SLOTS = 1ull << 23;   // per table of one graph
constexpr uint32_t TAIL_RING = 512;   // local-search ring per wave (LDS)
// A graph leaves the single-workgroup rounds (the host-driven rounds, which
// use the whole device, take over) once a round scanned so many edges that
// the next one, beta times larger, would exceed TAIL_ROUND_EDGES.
constexpr uint32_t TAIL_PROBE_LIMIT = 4096;
// Compact edges (both directions) above which a remainder is not taken.
constexpr uint64_t TAIL_MAX_EDGES = 1ull << 20;

// Global scratch of the tail, one region per graph (offsets in slots).
struct TailGraph {
  uint64_t base;        // first center index of the graph
  uint64_t nr;          // remaining vertices
  uint64_t tbase;       // first table slot (two tables of cap slots)
  uint64_t lbase;       // first log / frontier entry (6 lists of cap / 2)
  uint64_t cap;         // table slots, power of two
  uint64_t cap_max;     // largest size a relaunch may use
  uint64_t cur, step;   // pivot window schedule
  double beta;
  uint32_t g;
  uint32_t nu;          // unfinished vertices (host-side table sizing)
};

struct TailState {    // written back per graph
  unsigned long long cur, step, rounds, full, ins, edges, ticks;
};

// *go = 1 if every graph of the launch finished all its rounds (no table
// overflow), else 0.  One wave.
__global__ void k_tail_go(const TailState *ts, const TailGraph *tg, uint32_t nl,
                          unsigned long long *go) {
  bool ok = true;
  for (uint32_t k = lane(); k < nl; k += WAVE) ok = ok && ts[k].cur == tg[k].nr && !ts[k].full;
  const bool all = ballot(!ok) == 0;
  if (threadIdx.x == 0) *go = all ? 1ull : 0ull;
}

// The writeback of the compact labels and, when every graph of the launch
// finished its rounds (k_tail_go's test, made by every workgroup), the
// finalize of the remaining vertices and of each graph's first trimmed
// vertex (k_finalize_rest's values) in one pass.
__global__ void k_tail_final(const uint32_t *centers, uint64_t nrem,
                             const unsigned long long *tlab, unsigned long long *labels,
                             const TailState *ts, const TailGraph *tg, uint32_t nl,
                             const uint32_t *first_trim, const unsigned long long *zstart,
                             uint32_t G, unsigned long long *go) {
  bool ok = true;
  for (uint32_t k = lane(); k < nl; k += WAVE) ok = ok && ts[k].cur == tg[k].nr && !ts[k].full;
  const bool fin = ballot(!ok) == 0;
  if (blockIdx.x == 0 && threadIdx.x == 0) *go = fin ? 1ull : 0ull;
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < nrem + G;
       i += (uint64_t)gridDim.x * blockDim.x) {
    if (i < nrem) {
      const unsigned long long l = tlab[i];
      labels[centers[i]] = fin ? (l & VAL_MASK) - 1 : l;
    } else if (fin) {
      const uint64_t g = i - nrem;
      if (zstart[g + 1] == zstart[g]) continue;   // no trimmed vertex
      const uint32_t v = first_trim[g];
      labels[v] = (labels[v] & VAL_MASK) - 1;
    }
  }
}

// Remainder vertices are scanned by TAIL_GRP-lane groups, WAVE / TAIL_GRP
// vertices a wave (their lists are short), the windows of both directions
// at once, so a vertex's chain is four dependent loads.
constexpr int TAIL_GRP = 16;
static_assert(WAVE % TAIL_GRP == 0, "whole groups per wave");

// lid[centers[j]] = j, compact labels, and the compact degrees: edges to
// vertices that are unfinished and carry the same label.
__global__ void k_tail_degrees(const uint32_t *centers, uint64_t nrem,
                               const unsigned long long *labels,
                               const uint64_t *off, const uint32_t *adj,
                               const uint64_t *ioff, const uint32_t *iadj,
                               uint32_t *cnt_out, uint32_t *cnt_in, uint32_t *lid,
                               unsigned long long *tlab) {
  const int gl = lane() % TAIL_GRP, gs = lane() / TAIL_GRP * TAIL_GRP;
  if (blockIdx.x == 0 && threadIdx.x == 0) {   // scan sentinels
    cnt_out[nrem] = 0;
    cnt_in[nrem] = 0;
  }
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x / TAIL_GRP;
  for (uint64_t j0 = (blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) / WAVE * (WAVE / TAIL_GRP);
       j0 < nrem; j0 += stride) {
    const uint64_t j = j0 + lane() / TAIL_GRP;
    const bool own = j < nrem;
    const uint32_t v = own ? centers[j] : 0u;
    const unsigned long long lv = own ? labels[v] : 0ull;
    const uint64_t o0 = own ? off[v] : 0, o1 = own ? off[v + 1] : 0;
    const uint64_t i0 = own ? ioff[v] : 0, i1 = own ? ioff[v + 1] : 0;
    if (own && gl == 0) {
      lid[v] = (uint32_t)j;
      tlab[j] = lv;
    }
    uint32_t co = 0, ci = 0;
    const uint64_t dmax = max(o1 - o0, i1 - i0);
    for (uint64_t k = 0; ballot(k < dmax); k += TAIL_GRP) {
      const uint64_t eo = o0 + k + gl, ei = i0 + k + gl;
      const uint32_t xo = eo < o1 ? adj[eo] : 0u, xi = ei < i1 ? iadj[ei] : 0u;
      const unsigned long long lo = eo < o1 ? labels[xo] : TOP_BIT;
      const unsigned long long li = ei < i1 ? labels[xi] : TOP_BIT;
      const uint32_t mo = (uint32_t)(ballot(!(lo & TOP_BIT) && lo == lv) >> gs) & 0xFFFFu;
      const uint32_t mi = (uint32_t)(ballot(!(li & TOP_BIT) && li == lv) >> gs) & 0xFFFFu;
      co += (uint32_t)__popc(mo);
      ci += (uint32_t)__popc(mi);
    }
    if (own && gl == 0) {
      cnt_out[j] = co;
      cnt_in[j] = ci;
    }
  }
}

// Compact adjacency in original order (ballot ranks within each group's
// window), both directions' windows at once.
__global__ void k_tail_fill(const uint32_t *centers, uint64_t nrem,
                            const unsigned long long *labels, const uint32_t *lid,
                            const uint64_t *off, const uint32_t *adj,
                            const uint64_t *ioff, const uint32_t *iadj,
                            const uint32_t *toff, const uint32_t *tioff,
                            uint32_t *tadj, uint32_t *tiadj) {
  const int gl = lane() % TAIL_GRP, gs = lane() / TAIL_GRP * TAIL_GRP;
  const uint32_t below = (1u << gl) - 1u;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x / TAIL_GRP;
  for (uint64_t j0 = (blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) / WAVE * (WAVE / TAIL_GRP);
       j0 < nrem; j0 += stride) {
    const uint64_t j = j0 + lane() / TAIL_GRP;
    const bool own = j < nrem;
    const uint32_t v = own ? centers[j] : 0u;
    const unsigned long long lv = own ? labels[v] : 0ull;
    const uint64_t o0 = own ? off[v] : 0, o1 = own ? off[v + 1] : 0;
    const uint64_t i0 = own ? ioff[v] : 0, i1 = own ? ioff[v + 1] : 0;
    uint32_t ato = own ? toff[j] : 0u, ati = own ? tioff[j] : 0u;
    const uint64_t dmax = max(o1 - o0, i1 - i0);
    for (uint64_t k = 0; ballot(k < dmax); k += TAIL_GRP) {
      const uint64_t eo = o0 + k + gl, ei = i0 + k + gl;
      const uint32_t xo = eo < o1 ? adj[eo] : 0u, xi = ei < i1 ? iadj[ei] : 0u;
      const unsigned long long lo = eo < o1 ? labels[xo] : TOP_BIT;
      const unsigned long long li = ei < i1 ? labels[xi] : TOP_BIT;
      const bool ko = !(lo & TOP_BIT) && lo == lv, ki = !(li & TOP_BIT) && li == lv;
      const uint32_t lo_id = ko ? lid[xo] : 0u, li_id = ki ? lid[xi] : 0u;
      const uint32_t mo = (uint32_t)(ballot(ko) >> gs) & 0xFFFFu;
      const uint32_t mi = (uint32_t)(ballot(ki) >> gs) & 0xFFFFu;
      if (ko) tadj[ato + (uint32_t)__popc(mo & below)] = lo_id;
      if (ki) tiadj[ati + (uint32_t)__popc(mi & below)] = li_id;
      ato += (uint32_t)__popc(mo);
      ati += (uint32_t)__popc(mi);
    }
  }
}

// Insert (k, v) with linear probing; returns 1 and the slot if new, 0 if
// present, -1 past the probe limit.
__device__ __forceinline__ int tail_insert(unsigned long long *T, uint64_t mask,
                                           uint32_t k, uint32_t v, uint64_t *slot) {
  const unsigned long long kv = make_pair(k, v);
  uint64_t h = pair_slot(kv) & mask;
  for (uint32_t i = 0; i < TAIL_PROBE_LIMIT && i <= mask; ++i) {
    const unsigned long long s = __atomic_load_n(&T[h], __ATOMIC_RELAXED);
    if (s == kv) return 0;
    if (s == EMPTY_SLOT) {
      const unsigned long long old = atomicCAS(&T[h], EMPTY_SLOT, kv);
      if (old == EMPTY_SLOT) {
        *slot = h;
        return 1;
      }
      if (old == kv) return 0;
    }
    h = (h + 1) & mask;
  }
  return -1;
}

// Special positions of a remainder (bit j of `spec`, span positions):
// those with compact edges or already finished when the graph entered the
// tail.  Every other position is isolated: nothing reaches it and it
// reaches nothing, so it stays unfinished until its own round makes it a
// singleton SCC.  A round therefore reads the done flags of its window's
// special positions only, and labels the isolated ones by arithmetic.
//
// One part [c0, c1) of a round's window (at most 32 positions a thread): thread t
// owns a contiguous segment of at most 32 positions starting at span
// position p0 (its bits come from one or two bitmap words).  done marks
// its special positions that are finished now; u0 / ns0 are the
// block-wide exclusive prefixes of unfinished positions / unfinished
// special positions (seeds), tot / nstot their totals.  Every thread of
// the block calls it (it synchronises the block).
// Block-wide exclusive scan of one value per thread (NB threads); *total =
// the sum.  ws holds NB / WAVE entries; every thread of the block calls it.
template <int NB>
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t x, uint32_t *ws, uint32_t *total) {
  const uint32_t w = threadIdx.x / WAVE;
  const uint32_t inc = wave_incl_scan(x);
  __syncthreads();   // earlier readers of ws are done
  if ((threadIdx.x & (WAVE - 1)) == WAVE - 1) ws[w] = inc;
  __syncthreads();
  uint32_t before = 0, t = 0;
  for (uint32_t k = 0; k < NB / WAVE; ++k) {
    before += k < w ? ws[k] : 0u;
    t += ws[k];
  }
  *total = t;
  return before + inc - x;
}

struct SpecWord {
  uint64_t p0;                 // span position of bit 0
  uint32_t inwin, spec, done;  // bit masks
  uint32_t u0, tot, ns0, nstot;
};

// The bits of positions [p, p + len) (len <= 32) of a bitmap.
__device__ __forceinline__ uint32_t bits_at(const uint32_t *bm, uint64_t p, uint32_t len) {
  if (!len) return 0u;
  const uint64_t wd = p >> 5;
  const uint32_t sh = (uint32_t)(p & 31);
  uint64_t bits = bm[wd];
  if (sh + len > 32) bits |= (uint64_t)bm[wd + 1] << 32;
  return (uint32_t)(bits >> sh) & (len >= 32 ? ~0u : ((1u << len) - 1u));
}

// specw(p, len): special bits of span positions [p, p + len); labp(p): the
// label of special span position p.
template <int NB, class SpecF, class LabF>
__device__ __forceinline__ SpecWord tail_spec_part(SpecF specw, LabF labp, uint64_t base,
                                                   uint64_t c0, uint64_t c1,
                                                   uint32_t *wsum, uint32_t *wsum2) {
  const uint32_t tid = threadIdx.x, w = tid / WAVE;
  const uint64_t W = c1 - c0, per = (W + NB - 1) / NB;   // <= 32
  const uint64_t s0 = min((uint64_t)tid * per, W), s1 = min(s0 + per, W);
  SpecWord r;
  r.p0 = base + c0 + s0;
  const uint32_t len = (uint32_t)(s1 - s0);
  r.inwin = len >= 32 ? ~0u : ((1u << len) - 1u);
  r.spec = specw(r.p0, len);
  r.done = 0;
  for (uint32_t m = r.spec; m; m &= m - 1) {
    const int b = __ffs(m) - 1;
    if (labp(r.p0 + b) & TOP_BIT) r.done |= 1u << b;
  }
  const uint32_t unf = (uint32_t)__popc(r.inwin & ~r.done);
  const uint32_t sd = (uint32_t)__popc(r.spec & ~r.done);
  const uint32_t inc = wave_incl_scan(unf), incs = wave_incl_scan(sd);
  if ((tid & (WAVE - 1)) == WAVE - 1) {
    wsum[w] = inc;
    wsum2[w] = incs;
  }
  __syncthreads();
  uint32_t before = 0, befores = 0;
  r.tot = 0;
  r.nstot = 0;
  for (uint32_t x = 0; x < NB / WAVE; ++x) {
    before += x < w ? wsum[x] : 0u;
    r.tot += wsum[x];
    befores += x < w ? wsum2[x] : 0u;
    r.nstot += wsum2[x];
  }
  r.u0 = before + inc - unf;
  r.ns0 = befores + incs - sd;
  __syncthreads();   // wsum is reused by the next part
  return r;
}

// spec bitmap of the span positions [0, nrem) (bits past nrem clear).
__global__ void k_tail_special(const unsigned long long *tlab, const uint32_t *toff,
                               const uint32_t *tioff, uint64_t nrem, uint32_t *spec) {
  const int l = lane();
  for (uint64_t i0 = (blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) & ~(uint64_t)(WAVE - 1);
       i0 < nrem; i0 += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t i = i0 + l;
    bool sp = false;
    if (i < nrem)
      sp = (tlab[i] & TOP_BIT) || toff[i + 1] != toff[i] || tioff[i + 1] != tioff[i];
    const uint64_t m = ballot(sp);
    if (l == 0) spec[i0 >> 5] = (uint32_t)m;
    if (l == 32 && i0 + 32 < nrem) spec[(i0 >> 5) + 1] = (uint32_t)(m >> 32);
  }
}

struct TailArgs {
  const TailGraph *graphs;
  TailState *state;
  const uint32_t *toff, *tadj, *tioff, *tiadj;
  unsigned long long *tlab;
  const uint32_t *spec;        // special positions (k_tail_special)
  unsigned long long *table;   // 2 cap slots per graph, all EMPTY on entry
  uint32_t *lists;             // 6 lists of cap / 2 entries per graph
  unsigned long long *lab_off; // per graph label counter
  unsigned long long *stats;   // [0] inserts, [1] edges
  uint32_t tau;                // local-search budget per frontier pair
  int vgc;                     // local search on (-local_scc)
  // timing >= 3: per round of the first graph, (start tick, window,
  // seeds, search levels), TAIL_TRACE rounds at most; else null
  unsigned long long *trace;
  uint32_t first;              // descriptor / state index of the launch's first graph
  uint32_t count;              // graphs of the launch
  unsigned int *queue;         // work counter of a persistent launch (zero on entry)
};
constexpr uint32_t TAIL_TRACE = 64;
// Table slots of a graph whose pair tables, logs and frontiers live in LDS
// (two tables of TAIL_LDS_CAP u64 and six lists of TAIL_LDS_CAP / 2 u32).
// Every table, log and frontier access of its rounds is then an LDS access;
// only the compact graph and its labels are read from memory.  A graph whose
// rounds overflow them is relaunched with tables in memory.
constexpr uint64_t TAIL_LDS_CAP = 2048;
// A graph of at most TAIL_LG_POS remainder positions, TAIL_LG_VERTS special
// positions and TAIL_LG_EDGES compact edge entries (both directions) also
// keeps its special positions in LDS: their labels and compact lists under
// dense ids (rank among the special positions), so its searches and its
// partition touch no memory at all.  The labels go back at the end.
constexpr uint32_t TAIL_LG_POS = 32768;
constexpr uint32_t TAIL_LG_VERTS = 1024;
constexpr uint32_t TAIL_LG_EDGES = 2048;
// lg_pos entries: the local position, and bit 31 once the position is known
// to have been finished before its own round
constexpr uint32_t LG_PREDONE = 1u << 31, LG_POS_MASK = LG_PREDONE - 1;

// One workgroup per graph: all rounds of its remainder.  The host orders the
// graphs of a launch by decreasing work, and workgroups are dispatched in
// order as CUs free up, so the longest remainders start first and the short
// ones fill in behind them.  MODE 0: tables, logs and frontiers in the
// per-graph regions of A.table / A.lists; 1: in LDS (cap <= TAIL_LDS_CAP);
// 2: in LDS, and the special positions too (dense ids).  A graph that does
// not fit mode 2 stops before its first round with full = 2.
template <int MODE, int NB>
__global__ __launch_bounds__(NB) void k_ms_tail(TailArgs A) {
  constexpr bool LT = MODE >= 1, LG = MODE == 2;
  constexpr uint32_t HALF_WAVES = NB / WAVE / 2;   // waves per search direction
  constexpr uint64_t TAIL_PART = 32ull * NB;        // window part (<= 32 positions a thread)
  __shared__ unsigned long long lt_tab[LT ? 2 * TAIL_LDS_CAP : 1];
  __shared__ uint32_t lt_list[LT ? 3 * TAIL_LDS_CAP : 1];
  // mode 2: special bits of the graph's positions, specials before each
  // word, and per dense id: position, label, out- then in-offsets, lists
  __shared__ uint32_t lg_sw[LG ? TAIL_LG_POS / 32 : 1];
  __shared__ uint16_t lg_rp[LG ? TAIL_LG_POS / 32 : 1];
  __shared__ uint32_t lg_pos[LG ? TAIL_LG_VERTS : 1];
  __shared__ unsigned long long lg_lab[LG ? TAIL_LG_VERTS : 1];
  __shared__ uint32_t lg_off[LG ? 2 * (TAIL_LG_VERTS + 1) : 1];
  __shared__ uint32_t lg_adj[LG ? TAIL_LG_EDGES : 1];
  __shared__ uint32_t wsum[NB / WAVE], wsum2[NB / WAVE];
  __shared__ unsigned int cnt_q[2], cnt_log[2], flag_full;
  __shared__ unsigned int front_n, seed_n;
  __shared__ unsigned long long round_edges;
  // per-wave local-search ring and window scratch
  __shared__ uint32_t ring_s[NB / WAVE][TAIL_RING];
  __shared__ uint32_t sc_s[NB / WAVE][WAVE];
  __shared__ uint32_t st_s[NB / WAVE][WAVE];
  __shared__ uint32_t el_s[NB / WAVE][WAVE];
  __shared__ unsigned long long lw_s[NB / WAVE][WAVE];
  // one graph of the launch, by its descriptor / state index
  auto run_graph = [&](const uint32_t gi) {
  const TailGraph tg = A.graphs[gi];
  const uint32_t tid = threadIdx.x, w = tid / WAVE;
  const uint64_t base = tg.base, nr = tg.nr, half = tg.cap / 2, mask = tg.cap - 1;
  unsigned long long *const T0 = LT ? lt_tab : A.table + tg.tbase, *const T1 = T0 + tg.cap;
  uint32_t *const L0 = LT ? lt_list : A.lists + tg.lbase;   // slot indices of inserted pairs
  uint32_t *const L1 = L0 + half;
  // frontier ping-pong of each direction
  uint32_t *const qa0 = L0 + 2 * half, *const qb0 = L0 + 3 * half;
  uint32_t *const qa1 = L0 + 4 * half, *const qb1 = L0 + 5 * half;
  // labels of the searched vertices: dense ids (mode 2) or span positions
  unsigned long long *const lab = LG ? lg_lab : A.tlab;
  const uint32_t tau = A.tau;
  const int vgc = A.vgc;
  uint64_t cur = tg.cur, step = tg.step;
  unsigned long long loff = A.lab_off[tg.g];
  const uint64_t c0 = cur;                   // the launch's first position
  const unsigned long long loff0 = loff;     // and its label offset
  const unsigned long long t0 = wall_clock64();
  unsigned long long rounds = 0, ins = 0, edges = 0;
  uint32_t nlog0 = 0, nlog1 = 0;
  bool full = false, handover = false, lg_fail = false;
  if (LT) {   // LDS tables start empty (memory ones are cleared by the host)
    for (uint32_t i = tid; i < 2 * TAIL_LDS_CAP; i += NB) T0[i] = EMPTY_SLOT;
    __syncthreads();
  }
  // dense id of the special local position q (mode 2)
  auto rank = [&](uint64_t q) -> uint32_t {
    const uint32_t wd = (uint32_t)(q >> 5);
    return lg_rp[wd] + (uint32_t)__popc(lg_sw[wd] & ((1u << (q & 31)) - 1u));
  };
  uint32_t lg_n = 0;
  // dense id of the first special position at or after q (q <= nr)
  auto rank_end = [&](uint64_t q) -> uint32_t { return q >= nr ? lg_n : rank(q); };
  if (LG) {
    // special bits by local position (WPT words per thread, in order), the
    // specials before each word, then per dense id its position, label and
    // degrees; then the offsets (out-lists first, then in-lists, one array)
    // and the lists with neighbours by dense id
    constexpr uint32_t WPT = TAIL_LG_POS / 32 / NB, VPT = TAIL_LG_VERTS / NB;
    const uint32_t nw = (uint32_t)((nr + 31) / 32);
    const bool fits = nr <= TAIL_LG_POS;
    uint32_t c = 0;
    if (fits)
      for (uint32_t k = 0; k < WPT; ++k) {
        const uint32_t wd = tid * WPT + k;
        uint32_t bits = 0;
        if (wd < nw)
          bits = bits_at(A.spec, base + 32ull * wd, (uint32_t)min<uint64_t>(32, nr - 32ull * wd));
        lg_sw[wd] = bits;
        c += (uint32_t)__popc(bits);
      }
    uint32_t total = 0;
    uint32_t d = block_excl_scan<NB>(c, wsum, &total);
    lg_n = total;
    lg_fail = !fits || total > TAIL_LG_VERTS;
    if (!lg_fail)
      for (uint32_t k = 0; k < WPT; ++k) {
        const uint32_t wd = tid * WPT + k;
        if (wd >= nw) break;
        lg_rp[wd] = (uint16_t)d;
        for (uint32_t m = lg_sw[wd]; m; m &= m - 1, ++d)
          lg_pos[d] = 32 * wd + (uint32_t)(__ffs(m) - 1);
      }
    __syncthreads();
    // then a dense id per thread, the loads of all of a thread's ids in
    // flight together; the offsets stay in registers for the lists
    uint32_t o0[VPT], o1[VPT], i0[VPT], i1[VPT];
    if (!lg_fail) {
      unsigned long long lb[VPT];
#pragma unroll
      for (uint32_t k = 0; k < VPT; ++k) {
        const uint32_t dd = tid + k * NB;
        lb[k] = 0;
        o0[k] = o1[k] = i0[k] = i1[k] = 0;
        if (dd < lg_n) {
          const uint64_t p = base + lg_pos[dd];
          lb[k] = A.tlab[p];
          o0[k] = A.toff[p];
          o1[k] = A.toff[p + 1];
          i0[k] = A.tioff[p];
          i1[k] = A.tioff[p + 1];
        }
      }
#pragma unroll
      for (uint32_t k = 0; k < VPT; ++k) {
        const uint32_t dd = tid + k * NB;
        if (dd < lg_n) {
          lg_lab[dd] = lb[k];
          lg_off[dd] = o1[k] - o0[k];
          lg_off[TAIL_LG_VERTS + 1 + dd] = i1[k] - i0[k];
        }
      }
    }
    __syncthreads();
    if (!lg_fail) {
      uint32_t so = 0, si = 0;
      for (uint32_t k = 0; k < VPT; ++k) {
        const uint32_t dd = tid * VPT + k;
        if (dd < lg_n) {
          so += lg_off[dd];
          si += lg_off[TAIL_LG_VERTS + 1 + dd];
        }
      }
      uint32_t to = 0, ti = 0;
      uint32_t ro = block_excl_scan<NB>(so, wsum, &to);
      uint32_t ri = block_excl_scan<NB>(si, wsum2, &ti);
      lg_fail = to + ti > TAIL_LG_EDGES;
      if (!lg_fail) {
        ri += to;
        // each thread rewrites the degrees it read into offsets
        for (uint32_t k = 0; k < VPT; ++k) {
          const uint32_t dd = tid * VPT + k;
          if (dd >= lg_n) break;
          const uint32_t a = lg_off[dd], b = lg_off[TAIL_LG_VERTS + 1 + dd];
          lg_off[dd] = ro;
          lg_off[TAIL_LG_VERTS + 1 + dd] = ri;
          ro += a;
          ri += b;
        }
        if (tid == 0) {   // end sentinels
          lg_off[lg_n] = to;
          lg_off[TAIL_LG_VERTS + 1 + lg_n] = to + ti;
        }
      }
      __syncthreads();
      // every compact neighbour is special
      if (!lg_fail)
#pragma unroll
        for (uint32_t k = 0; k < VPT; ++k) {
          const uint32_t dd = tid + k * NB;
          if (dd >= lg_n) break;
          uint32_t at = lg_off[dd];
          for (uint32_t e = o0[k]; e < o1[k]; ++e) lg_adj[at++] = rank(A.tadj[e] - base);
          at = lg_off[TAIL_LG_VERTS + 1 + dd];
          for (uint32_t e = i0[k]; e < i1[k]; ++e) lg_adj[at++] = rank(A.tiadj[e] - base);
        }
      __syncthreads();
    }
    full = lg_fail;
  }
  // one part of a round's window (tail_spec_part), from LDS in mode 2
  auto spec_part = [&](uint64_t c0, uint64_t c1) -> SpecWord {
    if (LG)
      return tail_spec_part<NB>(
          [&](uint64_t p, uint32_t len) { return bits_at(lg_sw, p - base, len); },
          [&](uint64_t p) { return lg_lab[rank(p - base)]; }, base, c0, c1, wsum, wsum2);
    return tail_spec_part<NB>(
        [&](uint64_t p, uint32_t len) { return bits_at(A.spec, p, len); },
        [&](uint64_t p) { return A.tlab[p]; }, base, c0, c1, wsum, wsum2);
  };
  while (cur < nr && !full && !handover) {
    const uint64_t end = min(cur + step, nr);
    const bool tr = A.trace && gi == 0 && tid == 0 && rounds < TAIL_TRACE;
    if (tr) {
      A.trace[4 * rounds] = wall_clock64() - t0;
      A.trace[4 * rounds + 1] = end - cur;
    }
    SpecWord sw_keep{};
    const unsigned long long edges0 = edges;
    // ---- front: unfinished vertices of the window, in order; seeds
    // (pivot_i, loff + i) in both tables (src/scc.hpp:147-150)
    if (tid == 0) {
      front_n = 0;
      flag_full = 0;
      round_edges = 0;
    }
    __syncthreads();
    if (LG) {
      // Mode 2: only the window's special positions are read (dense ids
      // [d0, d1), in position order).  Every position of the window that was
      // unfinished at the round's start is a pivot, in order: pivot q gets
      // loff + (q - cur) - (specials of the window before q that were
      // finished before this round).  The seeds are the unfinished special
      // positions; the isolated ones (no compact edges) take their labels
      // from the same count after the last round (they are never finished
      // by another pivot, so nothing before then reads them).
      const uint32_t d0 = rank_end(cur), d1 = rank_end(end);
      uint32_t kb = 0, npre = 0;
      for (uint32_t x0 = d0; x0 < d1; x0 += NB) {
        const uint32_t d = x0 + tid;
        uint32_t pre = 0, sd = 0;
        if (d < d1) {
          pre = (lg_lab[d] & TOP_BIT) ? 1u : 0u;
          sd = 1u - pre;
        }
        uint32_t tot = 0;   // pre-done count << 16 | seed count (<= NB each)
        const uint32_t ex = block_excl_scan<NB>((pre << 16) | sd, wsum, &tot);
        if (sd) {
          const uint32_t q = lg_pos[d] & LG_POS_MASK;
          const uint32_t lb = (uint32_t)(loff + (q - cur) - (npre + (ex >> 16)));
          const uint32_t k = kb + (ex & 0xFFFFu);
          uint64_t q0 = 0, q1 = 0;
          const int r0 = tail_insert(T0, mask, d, lb, &q0);
          const int r1 = tail_insert(T1, mask, d, lb, &q1);
          if (r0 < 0 || r1 < 0 || k >= half) {
            flag_full = 1;
          } else {
            L0[k] = (uint32_t)q0;
            L1[k] = (uint32_t)q1;
          }
        }
        if (pre) lg_pos[d] |= LG_PREDONE;
        npre += tot >> 16;
        kb += tot & 0xFFFFu;
      }
      if (tid == 0) {
        front_n = (uint32_t)(end - cur) - npre;
        seed_n = kb;
      }
      __syncthreads();
    } else {
      // Unfinished vertices of the window, part by part: pivot i of the
      // round gets label loff + i.  A pivot without compact edges in either
      // direction is its own SCC this round (nothing reaches it, it reaches
      // nothing): it is labelled (loff + i) | TOP_BIT at the round's commit,
      // exactly what its pair would give in the partition; only the others
      // are seeded into the tables.
      uint32_t fb = 0, kb = 0;
      for (uint64_t c0 = cur; c0 < end; c0 += TAIL_PART) {
        const SpecWord sw = spec_part(c0, min(c0 + TAIL_PART, end));
        sw_keep = sw;   // the commit reuses a one-part window's scan
        const uint32_t seeds = sw.spec & ~sw.done, unf = sw.inwin & ~sw.done;
        uint32_t k = kb + sw.ns0;
        for (uint32_t mm = seeds; mm; mm &= mm - 1) {
          const int b = __ffs(mm) - 1;
          const uint32_t v = LG ? rank(sw.p0 + b - base) : (uint32_t)(sw.p0 + b);
          const uint32_t i = fb + sw.u0 + (uint32_t)__popc(unf & ((1u << b) - 1u));
          const uint32_t lb = (uint32_t)(loff + i);
          uint64_t q0 = 0, q1 = 0;
          const int r0 = tail_insert(T0, mask, v, lb, &q0);
          const int r1 = tail_insert(T1, mask, v, lb, &q1);
          if (r0 < 0 || r1 < 0 || k >= half) {
            flag_full = 1;
          } else {
            L0[k] = (uint32_t)q0;
            L1[k] = (uint32_t)q1;
          }
          ++k;
        }
        fb += sw.tot;
        kb += sw.nstot;
      }
      if (tid == 0) {
        front_n = fb;
        seed_n = kb;
      }
      __syncthreads();
    }
    const uint32_t nfront = front_n, nseed = seed_n;
    uint32_t slevels = 0;
    // ---- forward and backward multi-search over pairs (src/scc.hpp:160-252),
    // both at once: waves [0, HALF_WAVES) expand the forward search, the
    // others the backward one, so a round waits for the longer of the two
    // searches instead of their sum
    nlog0 = nlog1 = 0;
    if (nseed && !flag_full) {
      for (uint32_t i = tid; i < nseed; i += NB) {
        qa0[i] = L0[i];
        qa1[i] = L1[i];
      }
      uint32_t nq0 = nseed, nq1 = nseed, nl0 = nseed, nl1 = nseed;
      uint32_t *qc0 = qa0, *qn0 = qb0, *qc1 = qa1, *qn1 = qb1;
      const int dw = w >= HALF_WAVES ? 1 : 0;
      const uint32_t wd = w - (uint32_t)dw * HALF_WAVES;
      unsigned long long *const Td = dw ? T1 : T0;
      uint32_t *const Ld = dw ? L1 : L0;
      const uint32_t *const eo = LG ? lg_off + dw * (TAIL_LG_VERTS + 1) : (dw ? A.tioff : A.toff);
      const uint32_t *const ea = LG ? lg_adj : (dw ? A.tiadj : A.tadj);
      __syncthreads();
      while ((nq0 || nq1) && !flag_full) {
        ++slevels;
        if (tid < 2) {
          cnt_q[tid] = 0;
          cnt_log[tid] = tid ? nl1 : nl0;
        }
        __syncthreads();
        const uint32_t nq = dw ? nq1 : nq0;
        const uint32_t *const qc = dw ? qc1 : qc0;
        uint32_t *const qn = dw ? qn1 : qn0;
        // each wave takes 64 frontier pairs at a time and expands them with
        // load-balanced 64-edge windows; new pairs of compact degree <=
        // VGC_MAX_DEG stay in the wave's LDS ring while the budget lasts
        // (the VGC of src/scc.hpp:177-219), the rest go to the next frontier
        uint32_t *ring = ring_s[w];
        uint32_t *sc = sc_s[w];
        uint32_t *st = st_s[w];
        uint32_t *el = el_s[w];
        unsigned long long *lw = lw_s[w];
        const int l = (int)(tid & (WAVE - 1));
        for (uint32_t c0 = wd * WAVE; c0 < nq && !flag_full; c0 += HALF_WAVES * WAVE) {
          const uint32_t cnt = min((uint32_t)WAVE, nq - c0);
          if ((uint32_t)l < cnt) ring[l] = qc[c0 + l];
          uint32_t head = 0, tail = cnt;
          uint32_t budget = vgc ? tau : 0u;
          uint64_t ebudget = (uint64_t)budget * VGC_EDGE_MULT;
          wave_sync();
          while (head < tail) {
            const uint32_t k = min((uint32_t)WAVE, tail - head);
            const bool valid = (uint32_t)l < k;
            uint32_t deg = 0, b = 0, lb = 0;
            unsigned long long lv = 0;
            if (valid) {
              const unsigned long long pr = Td[ring[(head + l) & (TAIL_RING - 1)]];
              const uint32_t v = pair_vertex(pr);
              lb = pair_label(pr);
              b = eo[v];
              deg = eo[v + 1] - b;
              lv = lab[v];
            }
            head += k;
            const uint32_t inc = wave_incl_scan(deg);
            const uint32_t total = __shfl(inc, WAVE - 1);
            wave_sync();
            sc[l] = inc;
            st[l] = b;
            el[l] = lb;
            lw[l] = lv;
            ebudget = ebudget > total ? ebudget - total : 0;
            if (!ebudget) budget = 0;
            wave_sync();
            for (uint32_t base = 0; base < total; base += WAVE) {
              const uint32_t e = base + l;
              bool fresh = false;
              uint32_t x = 0, slot32 = 0, lbx = 0;
              if (e < total) {
                int lo = 0, hi = (int)k - 1;   // lanes >= k hold no list
                while (lo < hi) {
                  const int mid = (lo + hi) >> 1;
                  if (sc[mid] > e) hi = mid; else lo = mid + 1;
                }
                const uint32_t ex = lo ? sc[lo - 1] : 0u;
                x = ea[st[lo] + (e - ex)];
                ++edges;
                const unsigned long long lx = lab[x];
                if (!(lx & TOP_BIT) && lx == lw[lo]) {
                  lbx = el[lo];
                  uint64_t slot = 0;
                  const int r = tail_insert(Td, mask, x, lbx, &slot);
                  if (r < 0) flag_full = 1;
                  fresh = r > 0;
                  slot32 = (uint32_t)slot;
                }
              }
              const uint64_t fm = ballot(fresh);
              if (!fm) continue;
              ins += fresh;
              // log: one LDS atomic per window
              uint32_t lbase = 0;
              if (l == 0) lbase = atomicAdd(&cnt_log[dw], (uint32_t)__popcll(fm));
              lbase = __shfl(lbase, 0);
              if (fresh) {
                const uint32_t at = lbase + lanes_below(fm);
                if (at < half) Ld[at] = slot32; else flag_full = 1;
              }
              bool want = false;
              if (fresh && budget) want = eo[x + 1] - eo[x] <= VGC_MAX_DEG;
              const uint64_t wm = ballot(want);
              const uint32_t room = TAIL_RING - (tail - head);
              const uint32_t take = min(min((uint32_t)__popcll(wm), room), budget);
              const bool loc = want && lanes_below(wm) < take;
              if (loc) ring[(tail + lanes_below(wm)) & (TAIL_RING - 1)] = slot32;
              tail += take;
              budget -= take;
              const bool glob = fresh && !loc;
              const uint64_t gm = ballot(glob);
              if (gm) {
                uint32_t qb0 = 0;
                if (l == 0) qb0 = atomicAdd(&cnt_q[dw], (uint32_t)__popcll(gm));
                qb0 = __shfl(qb0, 0);
                if (glob) {
                  const uint32_t aq = qb0 + lanes_below(gm);
                  if (aq < half) qn[aq] = slot32; else flag_full = 1;
                }
              }
            }
            wave_sync();
          }
        }
        __syncthreads();
        if (nq0) {
          nq0 = min(cnt_q[0], (unsigned int)half);
          nl0 = min(cnt_log[0], (unsigned int)half);
          uint32_t *t = qc0;
          qc0 = qn0;
          qn0 = t;
        }
        if (nq1) {
          nq1 = min(cnt_q[1], (unsigned int)half);
          nl1 = min(cnt_log[1], (unsigned int)half);
          uint32_t *t = qc1;
          qc1 = qn1;
          qn1 = t;
        }
        __syncthreads();
      }
      nlog0 = nl0;
      nlog1 = nl1;
    }
    if (tr) {
      A.trace[4 * rounds + 2] = nseed;
      A.trace[4 * rounds + 3] = slevels;
    }
    if (edges != edges0) atomicAdd(&round_edges, edges - edges0);
    __syncthreads();
    full = flag_full;
    if (!full) {
      // isolated pivots: their own SCC, (loff + i) | TOP_BIT (written only
      // now: a stopped round leaves every label as it was).  The window's
      // parts are scanned again; only isolated positions are written, and
      // only special positions are read.  (Mode 2 writes them after the
      // last round.)
      if (!LG) {
        uint32_t fb = 0;
        for (uint64_t c0 = cur; c0 < end; c0 += TAIL_PART) {
          const SpecWord sw = end - cur <= TAIL_PART ? sw_keep
                                                     : spec_part(c0, min(c0 + TAIL_PART, end));
          const uint32_t unf = sw.inwin & ~sw.done;
          for (uint32_t mm = sw.inwin & ~sw.spec; mm; mm &= mm - 1) {
            const int b = __ffs(mm) - 1;
            const uint32_t i = fb + sw.u0 + (uint32_t)__popc(unf & ((1u << b) - 1u));
            A.tlab[sw.p0 + b] = (unsigned long long)(loff + i) | TOP_BIT;
          }
          fb += sw.tot;
        }
      }
      // ---- partition (src/scc.hpp:437-464): a pair in both tables is in
      // the pivot's SCC; labels take the maximum
      for (uint32_t i = tid; i < nlog0; i += NB) {
        const unsigned long long pr = T0[L0[i]];
        const uint32_t v = pair_vertex(pr), lb = pair_label(pr);
        uint32_t probes = 0;
        const unsigned long long val =
            table_contains(T1, mask, v, lb, &probes) ? ((unsigned long long)lb | TOP_BIT)
                                                     : (unsigned long long)lb;
        if (lab[v] < val) atomicMax(&lab[v], val);
      }
      for (uint32_t i = tid; i < nlog1; i += NB) {
        const unsigned long long pr = T1[L1[i]];
        const unsigned long long val = pair_label(pr);
        const uint32_t v = pair_vertex(pr);
        if (lab[v] < val) atomicMax(&lab[v], val);
      }
      loff += nfront;
      cur = end;
      step = (uint64_t)ceil((double)step * tg.beta);
      ++rounds;
      handover = (double)round_edges * tg.beta > TAIL_ROUND_EDGES;
    }
    __syncthreads();
    // ---- empty the slots this round used
    if (!full) {
      for (uint32_t i = tid; i < nlog0; i += NB) T0[L0[i]] = EMPTY_SLOT;
      for (uint32_t i = tid; i < nlog1; i += NB) T1[L1[i]] = EMPTY_SLOT;
    }
    __syncthreads();
  }
  if (full && !LT) {
    // slots of a stopped round are not all logged: clear the whole region
    for (uint64_t i = tid; i < 2 * tg.cap; i += NB) T0[i] = EMPTY_SLOT;
  }
  if (LG && !lg_fail) {
    // the isolated positions of the completed rounds, [c0, cur): pivot q
    // has loff0 + (q - c0) - (specials in [c0, q) finished before their own
    // round), the running pivot index over all rounds of this launch; the
    // count comes from an exclusive prefix of those flags over dense ids
    // [rank(c0), rank(cur)), kept in lg_adj (the lists are no longer read)
    const uint32_t dc0 = rank_end(c0), dcur = rank_end(cur);
    constexpr uint32_t VPT = TAIL_LG_VERTS / NB;
    uint32_t cnt = 0;
    for (uint32_t k = 0; k < VPT; ++k) {
      const uint32_t d = dc0 + tid * VPT + k;
      if (d < dcur) cnt += lg_pos[d] >> 31;
    }
    uint32_t tot = 0;
    uint32_t run = block_excl_scan<NB>(cnt, wsum, &tot);
    uint32_t *const pdx = lg_adj;
    for (uint32_t k = 0; k < VPT; ++k) {
      const uint32_t d = dc0 + tid * VPT + k;
      if (d < dcur) {
        pdx[d - dc0] = run;
        run += lg_pos[d] >> 31;
      }
    }
    if (tid == 0) pdx[dcur - dc0] = tot;
    __syncthreads();
    // per bitmap word of positions (a thread each): its specials' flags as
    // a mask (flags are set only for specials in [c0, cur)) and the flags
    // before the word, in the first table's LDS (no longer read)
    uint32_t *const pdm_w = reinterpret_cast<uint32_t *>(lt_tab);
    uint32_t *const pd0_w = pdm_w + TAIL_LG_POS / 32;
    static_assert(2 * (TAIL_LG_POS / 32) * 4 <= 2 * TAIL_LDS_CAP * 8, "word flags fit the tables");
    const uint64_t w_end = cur > c0 ? (cur + 31) / 32 : 0;
    for (uint64_t w = c0 / 32 + tid; w < w_end; w += NB) {
      const uint32_t bits = lg_sw[w], r0 = lg_rp[w];
      uint32_t pdm = 0, d = r0;
      for (uint32_t m = bits; m; m &= m - 1, ++d)
        if (lg_pos[d] >> 31) pdm |= m & (~m + 1u);
      pdm_w[w] = pdm;
      pd0_w[w] = r0 >= dc0 ? pdx[r0 - dc0] : 0u;
    }
    __syncthreads();
    // then a position per thread, so the label stores are coalesced, and
    // EPI_STEPS positions' LDS reads in flight together
    constexpr uint32_t EPI_STEPS = 8;
    for (uint64_t q0 = c0 + tid; q0 < cur; q0 += NB * EPI_STEPS) {
      uint32_t sw[EPI_STEPS], pd[EPI_STEPS], pm[EPI_STEPS];
#pragma unroll
      for (uint32_t u = 0; u < EPI_STEPS; ++u) {
        const uint64_t q = q0 + u * NB;
        const uint32_t w = q < cur ? (uint32_t)(q >> 5) : 0u;
        sw[u] = lg_sw[w];
        pd[u] = pd0_w[w];
        pm[u] = pdm_w[w];
      }
#pragma unroll
      for (uint32_t u = 0; u < EPI_STEPS; ++u) {
        const uint64_t q = q0 + u * NB;
        const uint32_t b = (uint32_t)(q & 31);
        if (q >= cur || ((sw[u] >> b) & 1u)) continue;
        const uint32_t p = pd[u] + (uint32_t)__popc(pm[u] & ((1u << b) - 1u));
        A.tlab[base + q] = (unsigned long long)(loff0 + (q - c0) - p) | TOP_BIT;
      }
    }
    // the special positions' labels as of the last round
    for (uint32_t d = tid; d < lg_n; d += NB) A.tlab[base + (lg_pos[d] & LG_POS_MASK)] = lg_lab[d];
  }
  ins = wave_sum(ins);
  edges = wave_sum(edges);
  if ((tid & (WAVE - 1)) == 0) {
    if (ins) atomicAdd(&A.stats[0], ins);
    if (edges) atomicAdd(&A.stats[1], edges);
  }
  __syncthreads();
  if (tid == 0) {
    // full: 1 = a table, log or frontier overflowed, 2 = the graph does not
    // fit mode 2 (no round ran)
    A.state[gi] = TailState{cur, step, rounds, lg_fail ? 2ull : full ? 1ull : 0ull, 0, 0,
                            wall_clock64() - t0};
    A.lab_off[tg.g] = loff;
  }
  };
  if constexpr (NB == TAIL_SMALL_BLOCK) {
    // a persistent grid: each workgroup takes the launch's graphs from its
    // work counter, in the host's order (largest remainders first)
    __shared__ uint32_t item_s;
    for (;;) {
      __syncthreads();   // the previous graph's LDS use and item reads are done
      if (threadIdx.x == 0) item_s = atomicAdd(A.queue, 1u);
      __syncthreads();
      const uint32_t item = item_s;
      if (item >= A.count) break;
      run_graph(A.first + item);
    }
  } else {
    run_graph(A.first + blockIdx.x);   // one workgroup per graph
  }
}

// Unfinished vertices of each described graph at positions [cur, nr) of
// its remainder (every earlier position was a pivot and is finished).
__global__ void k_tail_unfinished(const TailGraph *tg, const uint32_t *centers,
                                  const unsigned long long *labels,
                                  unsigned long long *out) {
  __shared__ unsigned long long red[TAIL_BLOCK / WAVE];
  const TailGraph t = tg[blockIdx.x];
  unsigned long long c = 0;
  for (uint64_t j = t.cur + threadIdx.x; j < t.nr; j += blockDim.x)
    c += !(labels[centers[t.base + j]] & TOP_BIT);
  c = wave_sum(c);
  if (lane() == 0) red[threadIdx.x / WAVE] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long tot = 0;
    for (uint32_t x = 0; x < blockDim.x / WAVE; ++x) tot += red[x];
    out[blockIdx.x] = tot;
  }
}

// Compact edge count (both directions) of every graph, from the per-graph
// starts of the remaining list; out[G] and out[G + 1] get the total out-
// and in-edge counts.
__global__ void k_tail_sizes(const unsigned long long *cstart, uint32_t G, uint64_t lo,
                             const uint32_t *toff, const uint32_t *tioff, uint64_t nrem,
                             unsigned long long *out) {
  for (uint32_t g = blockIdx.x * blockDim.x + threadIdx.x; g <= G;
       g += gridDim.x * blockDim.x) {
    if (g == G) {
      out[G] = toff[nrem];
      out[G + 1] = tioff[nrem];
      continue;
    }
    // graphs outside the compacted range [lo, lo + nrem) count 0
    const uint64_t b = min(max(cstart[g], lo), lo + nrem) - lo;
    const uint64_t e = min(max(cstart[g + 1], lo), lo + nrem) - lo;
    out[g] = (unsigned long long)(toff[e] - toff[b]) + (tioff[e] - tioff[b]);
  }
}

__global__ void k_tail_writeback(const uint32_t *centers, uint64_t nrem,
                                 const unsigned long long *tlab,
                                 unsigned long long *labels) {
  for (uint64_t j = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; j < nrem;
       j += (uint64_t)gridDim.x * blockDim.x)
    labels[centers[j]] = tlab[j];
}

__global__ void k_tail_fill_u64(unsigned long long *p, uint64_t n,
                                unsigned long long *zero2, unsigned int *queue) {
  if (blockIdx.x == 0 && threadIdx.x < 2) zero2[threadIdx.x] = 0;
  if (blockIdx.x == 0 && threadIdx.x == 0) *queue = 0;
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = EMPTY_SLOT;
}

int blocks(uint64_t n, int per = 256, int cap = 16384) {
  uint64_t g = (n + per - 1) / per;
  return (int)std::max<uint64_t>(1, std::min<uint64_t>(g, cap));
}

}  // namespace

int Engine::tail_pin_ensure(size_t bytes) {
  if (bytes <= tail_pin_cap) return PSC_OK;
  if (host_static) {
    set_error("allocation-free run: remainder descriptors exceed the staging area");
    return PSC_ERR_INVALID;
  }
  GateGuard gate;
  PSC_HIP(hipStreamSynchronize(stream));
  if (tail_pin) (void)hipHostFree(tail_pin);
  tail_pin = nullptr;
  tail_pin_cap = 0;
  PSC_HIP(hipHostMalloc(&tail_pin, bytes));
  tail_pin_cap = bytes;
  return PSC_OK;
}

// Runs the remaining rounds of every unfinished graph with at most
// p.tail_max unfinished vertices (and at most TAIL_MAX_POS remainder
// positions, TAIL_MAX_EDGES compact edges) on the device; cur / step of
// those graphs are advanced (fully, unless a graph's tables overflowed or
// its rounds grew heavy).  Returns PSC_OK with *used = false when no graph
// qualifies.
int Engine::ms_tail(const psc_params &p, const uint64_t *h_cstart, uint64_t n_rem,
                    std::vector<uint64_t> &cur, std::vector<uint64_t> &stepg,
                    psc_stats *st, bool *used) {
  *used = false;
  const double beta_def = p.beta > 0 ? p.beta : 1.5;
  // Candidates: unfinished graphs of at most TAIL_MAX_POS remainder
  // positions.  A graph qualifies with at most tail_max unfinished
  // vertices; there are at most nr - cur of them (every position before
  // cur was a pivot and is finished), and only graphs above that bound
  // are counted (one small kernel and one poll).
  std::vector<TailGraph> tgs;
  bool count = false;
  for (uint32_t g = 0; g < G; ++g) {
    const uint64_t r = h_cstart[g + 1] - h_cstart[g];
    if (r == 0 || cur[g] == r || r > TAIL_MAX_POS) continue;
    TailGraph t{};
    t.base = h_cstart[g];
    t.nr = r;
    t.cur = cur[g];
    t.step = stepg[g];
    t.beta = beta_g.empty() ? beta_def : beta_g[g];
    t.g = g;
    t.nu = (uint32_t)(r - cur[g]);
    count |= t.nu > p.tail_max;
    tgs.push_back(t);
  }
  if (tgs.empty() || n_rem == 0) return PSC_OK;
  const size_t ns = tgs.size();   // descriptor / state slots
  PSC_TRY(tail_meta.ensure(ns * (sizeof(TailGraph) + sizeof(TailState)) + (2 * G + 4) * 8 +
                           64));
  TailGraph *d_tg = tail_meta.as<TailGraph>();
  TailState *d_ts = reinterpret_cast<TailState *>(d_tg + ns);
  unsigned long long *d_stats = reinterpret_cast<unsigned long long *>(d_ts + ns);
  unsigned long long *d_e = d_stats + 2;
  unsigned long long *d_go = d_stats + (2 * G + 4);   // in the 64 spare bytes
  if (count) {
    unsigned long long *d_u = d_e + (G + 2);
    PSC_TRY(tail_pin_ensure(ns * sizeof(TailGraph)));
    std::memcpy(tail_pin, tgs.data(), ns * sizeof(TailGraph));
    PSC_HIP(hipMemcpyAsync(d_tg, tail_pin, ns * sizeof(TailGraph), hipMemcpyHostToDevice,
                           stream));
    prof_begin(KK_TAIL);
    k_tail_unfinished<<<(unsigned)ns, TAIL_BLOCK, 0, stream>>>(
        d_tg, centers.as<uint32_t>(), labels.as<unsigned long long>(), d_u);
    prof_end();
    std::vector<unsigned long long> h_u(ns);
    PSC_TRY(poll(h_u.data(), d_u, ns * 8));
    std::vector<TailGraph> keep;
    for (size_t i = 0; i < ns; ++i) {
      if (h_u[i] > p.tail_max) continue;
      tgs[i].nu = (uint32_t)h_u[i];
      keep.push_back(tgs[i]);
    }
    tgs.swap(keep);
    if (tgs.empty()) return PSC_OK;
  }
  // compact only the span of the selected graphs' remainders
  const uint64_t lo = tgs.front().base, span_end = h_cstart[tgs.back().g + 1];
  n_rem = span_end - lo;
  for (TailGraph &t : tgs) t.base -= lo;
  const unsigned long long *lab = labels.as<unsigned long long>();
  const uint32_t *cen = centers.as<uint32_t>() + lo;
  // ---- compact remainder: ids by pivot position, same-subproblem edges
  PSC_TRY(tail_cnt.ensure(2 * (n_rem + 1) * 4));
  PSC_TRY(tail_lab.ensure(n_rem * 8));
  uint32_t *cnt_out = tail_cnt.as<uint32_t>(), *cnt_in = cnt_out + (n_rem + 1);
  uint32_t *lid = scratch32.as<uint32_t>();
  prof_begin(KK_TAIL);
  k_tail_degrees<<<blocks(n_rem * TAIL_GRP), 256, 0, stream>>>(
      cen, n_rem, lab, d_off, d_adj, d_ioff, d_iadj, cnt_out, cnt_in, lid,
      tail_lab.as<unsigned long long>());
  // one scan over both degree arrays (contiguous, each with a 0 sentinel):
  // out-edges at [toff[j], toff[j+1]), in-edges at [tioff[j], tioff[j+1])
  // of the same compact edge array
  PSC_TRY(tail_off.ensure(2 * (n_rem + 1) * 4));
  uint32_t *toff = tail_off.as<uint32_t>(), *tioff = toff + (n_rem + 1);
  size_t tb = 0;
  PSC_HIP(rocprim::exclusive_scan(nullptr, tb, cnt_out, toff, 0u, 2 * (n_rem + 1),
                                  rocprim::plus<uint32_t>(), stream));
  PSC_TRY(sort_tmp.ensure(std::max<size_t>(tb, 8)));
  tb = sort_tmp.cap;
  PSC_HIP(rocprim::exclusive_scan(sort_tmp.ptr, tb, cnt_out, toff, 0u, 2 * (n_rem + 1),
                                  rocprim::plus<uint32_t>(), stream));
  prof_end();
  // per-graph compact edge counts size the pair tables (one poll)
  k_tail_sizes<<<blocks(G + 1), 256, 0, stream>>>(gB.as<unsigned long long>(), G, lo,
                                                  toff, tioff, n_rem, d_e);
  std::vector<unsigned long long> h_e(G + 2);
  PSC_TRY(poll(h_e.data(), d_e, (G + 2) * 8));
  // A remainder that keeps many edges runs its next rounds on the whole
  // device: one workgroup scans them far slower, and the host-driven rounds
  // return the graph here once its rounds are light.
  {
    std::vector<TailGraph> keep;
    for (const TailGraph &t : tgs)
      if (h_e[t.g] <= TAIL_MAX_EDGES) keep.push_back(t);
    tgs.swap(keep);
  }
  if (tgs.empty()) return PSC_OK;
  const size_t nt = tgs.size();
  // out- and in-edges of the compact copy: the same edges, so twice the
  // out-edges unless the 32-bit offsets wrapped (then the host-driven
  // rounds take the remainders)
  const uint64_t e_all = h_e[G + 1];
  if (h_e[G] >= (1ull << 31) || e_all != 2 * h_e[G]) return PSC_OK;
  PSC_TRY(tail_adj.ensure((e_all + 2) * 4));
  uint32_t *tadj = tail_adj.as<uint32_t>(), *tiadj = tadj;
  PSC_TRY(tail_spec.ensure(((n_rem + 31) / 32 + 1) * 4));
  prof_begin(KK_TAIL);
  k_tail_fill<<<blocks(n_rem * TAIL_GRP), 256, 0, stream>>>(
      cen, n_rem, lab, lid, d_off, d_adj, d_ioff, d_iadj, toff, tioff, tadj, tiadj);
  k_tail_special<<<blocks(n_rem), 256, 0, stream>>>(tail_lab.as<unsigned long long>(), toff,
                                                    tioff, n_rem, tail_spec.as<uint32_t>());
  prof_end();
  for (size_t i = 0; i < nt; ++i) {
    TailGraph &t = tgs[i];
    uint64_t cap = 1024;
    while (cap < 4 * ((uint64_t)t.nu + h_e[t.g])) cap *= 2;
    if (p.table_init) {   // forced small tables: exercises the hand-over
      cap = 16;
      while (cap < p.table_init) cap *= 2;
    }
    t.cap = cap;
    t.cap_max = std::min<uint64_t>(TAIL_MAX_SLOTS, cap * 64);
  }
  // A graph with few compact edges (its pairs per round are bounded by its
  // special positions) runs with its tables in LDS first (lds 1), and with
  // its special positions in LDS too when it may fit them (lds 2; the
  // kernel checks the count of special positions).
  // (Forced small tables, p.table_init, apply to the LDS tables too: the
  // overflow of every kind is then exercised.)
  std::vector<char> lds(nt, 0);
  for (size_t i = 0; i < nt; ++i) {
    const uint64_t e = h_e[tgs[i].g];
    lds[i] = e > TAIL_LDS_CAP / 4 ? 0
             : tgs[i].nr <= TAIL_LG_POS && 2 * e <= TAIL_LG_EDGES ? 2 : 1;
  }
  // Launch the graphs still pending; a graph whose tables overflowed stops
  // at the start of that round and is relaunched: from LDS to memory tables
  // of its computed size, from memory tables with 4x the slots, up to 64x
  // its first size and TAIL_MAX_SLOTS (then the host-driven rounds finish
  // it).
  std::vector<size_t> pending(nt);
  for (size_t i = 0; i < nt; ++i) pending[i] = i;
  std::vector<uint64_t> rounds(nt, 0);
  unsigned long long hst[2] = {0, 0};
  std::vector<TailGraph> launch;
  std::vector<TailState> hs;
  while (!pending.empty()) {
    // LDS graphs, then the others; within each, largest remainders first
    // (positions left plus compact edges estimate a graph's rounds)
    std::stable_sort(pending.begin(), pending.end(), [&](size_t x, size_t y) {
      if (lds[x] != lds[y]) return lds[x] > lds[y];
      return tgs[x].nr - tgs[x].cur + h_e[tgs[x].g] > tgs[y].nr - tgs[y].cur + h_e[tgs[y].g];
    });
    launch.clear();
    uint64_t slots = 0, lslots = 0;
    size_t nmode[3] = {0, 0, 0};
    std::vector<char> lmode;   // the launch's mode per graph
    for (size_t i : pending) {
      TailGraph t = tgs[i];
      ++nmode[(int)lds[i]];
      lmode.push_back(lds[i]);
      if (lds[i]) {
        t.cap = p.table_init ? std::min<uint64_t>(t.cap, TAIL_LDS_CAP) : TAIL_LDS_CAP;
        t.tbase = t.lbase = 0;
      } else {
        t.tbase = slots;   // two tables of cap slots
        t.lbase = lslots;   // six lists of cap / 2 u32
        slots += 2 * t.cap;
        lslots += 3 * t.cap;
      }
      launch.push_back(t);
    }
    const size_t nl = launch.size();
    PSC_TRY(tail_table.ensure(slots * 8));
    PSC_TRY(tail_lists.ensure(lslots * 4));
    // descriptors go through pinned staging (the stream is synchronised
    // by the poll below before the staging is written again)
    PSC_TRY(tail_pin_ensure(nl * sizeof(TailGraph)));
    std::memcpy(tail_pin, launch.data(), nl * sizeof(TailGraph));
    PSC_HIP(hipMemcpyAsync(d_tg, tail_pin, nl * sizeof(TailGraph), hipMemcpyHostToDevice,
                           stream));
    prof_begin(KK_TAIL);
    unsigned int *d_queue = reinterpret_cast<unsigned int *>(d_go + 1);
    k_tail_fill_u64<<<blocks(slots), 256, 0, stream>>>(tail_table.as<unsigned long long>(),
                                                       slots, d_stats, d_queue);
    TailArgs a;
    a.graphs = d_tg;
    a.state = d_ts;
    a.toff = toff;
    a.tadj = tadj;
    a.tioff = tioff;
    a.tiadj = tiadj;
    a.tlab = tail_lab.as<unsigned long long>();
    a.spec = tail_spec.as<uint32_t>();
    a.table = tail_table.as<unsigned long long>();
    a.lists = tail_lists.as<uint32_t>();
    a.lab_off = gD.as<unsigned long long>();
    a.stats = d_stats;
    a.tau = (uint32_t)std::min<uint64_t>(p.tau ? p.tau : 1, 1u << 30);
    a.vgc = p.local_scc;
    a.trace = nullptr;
    if (p.timing >= 3) {
      PSC_TRY(tail_trace.ensure((4 * TAIL_TRACE + 1) * 8));
      a.trace = tail_trace.as<unsigned long long>();
    }
    // the launch lists mode 2 graphs, then mode 1, then mode 0; mode 2 runs
    // as a persistent grid (one workgroup per CU, LDS-bound) over the work
    // counter, the others one workgroup per graph
    a.queue = d_queue;
    a.first = 0;
    a.count = (uint32_t)nmode[2];
    if (nmode[2])
      k_ms_tail<2, TAIL_SMALL_BLOCK>
          <<<(unsigned)std::min<size_t>(nmode[2], (size_t)num_cu), TAIL_SMALL_BLOCK, 0,
             stream>>>(a);
    a.first = (uint32_t)nmode[2];
    a.count = (uint32_t)nmode[1];
    if (nmode[1]) k_ms_tail<1, TAIL_BLOCK><<<(unsigned)nmode[1], TAIL_BLOCK, 0, stream>>>(a);
    a.first = (uint32_t)(nmode[2] + nmode[1]);
    a.count = (uint32_t)nmode[0];
    if (nmode[0]) k_ms_tail<0, TAIL_BLOCK><<<(unsigned)nmode[0], TAIL_BLOCK, 0, stream>>>(a);
    // When every graph with rounds left is in this launch, the finalize
    // goes behind it before the poll, skipped on the device unless all of
    // them finished: the poll then returns with the labels final.
    bool spec_fin = false, all_in = true;
    {
      std::vector<char> in_launch(G, 0);
      for (const TailGraph &t : launch) in_launch[t.g] = 1;
      for (uint32_t g = 0; g < G && all_in; ++g)
        if (!in_launch[g] && cur[g] < h_cstart[g + 1] - h_cstart[g]) all_in = false;
    }
    // labels as of every completed round, enqueued before the poll (a
    // relaunch continues from the compact labels; the host rounds and the
    // finalize read these); with the finalize when it can go with them
    if (all_in && direct_final && final_rem == n_rem) {
      k_tail_final<<<blocks(n_rem + G), 256, 0, stream>>>(
          cen, n_rem, tail_lab.as<unsigned long long>(), labels.as<unsigned long long>(), d_ts,
          d_tg, (uint32_t)nl, first_trim.as<uint32_t>(), gA.as<unsigned long long>() + (G + 1),
          G, d_go);
      if (p.timing) PSC_HIP(hipEventRecord(ev[3], stream));   // the finalize is in it
      spec_fin = true;
      all_in = false;   // nothing more to enqueue
    } else {
      k_tail_writeback<<<blocks(n_rem), 256, 0, stream>>>(
          cen, n_rem, tail_lab.as<unsigned long long>(), labels.as<unsigned long long>());
    }
    prof_end();
    PSC_HIP(hipGetLastError());
    {
      if (all_in) {
        if (p.timing) PSC_HIP(hipEventRecord(ev[3], stream));
        prof_begin(KK_MISC);
        k_tail_go<<<1, WAVE, 0, stream>>>(d_ts, d_tg, (uint32_t)nl, d_go);
        PSC_TRY(finalize_labels(d_go));
        prof_end();
        PSC_HIP(hipGetLastError());
        spec_fin = true;
      }
    }
    // the ns state slots and the two statistics counters are contiguous:
    // one poll (slots past this launch's nl are not read)
    std::vector<unsigned char> back(ns * sizeof(TailState) + 16);
    PSC_TRY(poll(back.data(), d_ts, back.size()));
    hs.resize(nl);
    std::memcpy(hs.data(), back.data(), nl * sizeof(TailState));
    unsigned long long part[2] = {0, 0};
    std::memcpy(part, back.data() + ns * sizeof(TailState), 16);
    hst[0] += part[0];
    hst[1] += part[1];
    std::vector<size_t> again;
    for (size_t k = 0; k < nl; ++k) {
      const size_t i = pending[k];
      TailGraph &t = tgs[i];
      t.cur = hs[k].cur;
      t.step = hs[k].step;
      cur[t.g] = hs[k].cur;
      stepg[t.g] = hs[k].step;
      rounds[i] += hs[k].rounds;
      if (hs[k].full == 2) {   // more special positions than mode 2 holds
        lds[i] = 1;
        again.push_back(i);
      } else if (hs[k].full) {
        ++st->resizes;
        if (lds[i]) {   // memory tables of the computed size
          lds[i] = 0;
          again.push_back(i);
        } else if (t.cap * 4 <= t.cap_max) {
          t.cap *= 4;
          again.push_back(i);
        }
      }
    }
    if (spec_fin) {   // the same test as k_tail_go's, on the polled states
      bool all_done = true;
      for (uint32_t g = 0; g < G && all_done; ++g)
        if (cur[g] < h_cstart[g + 1] - h_cstart[g]) all_done = false;
      finalized = all_done;
    }
    if (p.timing >= 3 && nl) {   // rounds of the launch's first graph
      std::vector<unsigned long long> tr(4 * TAIL_TRACE);
      PSC_HIP(hipMemcpy(tr.data(), tail_trace.ptr, tr.size() * 8, hipMemcpyDeviceToHost));
      const uint64_t nrd = std::min<uint64_t>(hs[0].rounds, TAIL_TRACE);
      for (uint64_t r = 0; r < nrd; ++r) {
        const unsigned long long t1 = r + 1 < nrd ? tr[4 * (r + 1)] : hs[0].ticks;
        std::fprintf(stderr, "tail round %llu: window %llu seeds %llu search levels %llu %.1f us\n",
                     (unsigned long long)r, tr[4 * r + 1], tr[4 * r + 2], tr[4 * r + 3],
                     (t1 - tr[4 * r]) / 100.0);
      }
    }
    if (p.timing >= 3) {   // per-graph report, slowest first (wall clock 100 MHz)
      std::vector<size_t> ord(nl);
      for (size_t k = 0; k < nl; ++k) ord[k] = k;
      std::sort(ord.begin(), ord.end(),
                [&](size_t x, size_t y) { return hs[x].ticks > hs[y].ticks; });
      for (size_t r = 0; r < std::min<size_t>(nl, 8); ++r) {
        const size_t k = ord[r];
        const TailGraph &t = launch[k];
        std::fprintf(stderr,
                     "tail graph %u: remainder %llu compact edges %llu cap %llu mode %d "
                     "rounds %llu %s %.1f us\n",
                     t.g, (unsigned long long)t.nr, (unsigned long long)h_e[t.g],
                     (unsigned long long)t.cap, (int)lmode[k], hs[k].rounds,
                     hs[k].full ? "overflow" : (hs[k].cur < t.nr ? "handed over" : "done"),
                     hs[k].ticks / 100.0);
      }
    }
    pending.swap(again);
  }
  st->rounds += *std::max_element(rounds.begin(), rounds.end());
  st->table_inserts += hst[0];
  st->edges_scanned += hst[1];
  st->bytes_csr += 8 * e_all + 8 * hst[1];
  st->bytes_label += 16 * hst[1];
  st->bytes_table += 16 * hst[0];
  *used = true;
  return PSC_OK;
}

}  // namespace psc
