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
// completed round) and the host-driven rounds finish it.
#include <hip/hip_runtime.h>

#include <cstring>
#include <rocprim/device/device_scan.hpp>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "device_common.hpp"
#include "engine.hpp"
#include "scc_kernels.hpp"

namespace psc {

namespace {

constexpr int TAIL_BLOCK = 1024;
constexpr uint32_t TAIL_PROBE_LIMIT = 4096;

// Global scratch of the tail, one region per graph (offsets in slots).
struct TailGraph {
  uint64_t base;        // first center index of the graph
  uint64_t nr;          // remaining vertices
  uint64_t tbase;       // first table slot (two tables of cap slots)
  uint64_t lbase;       // first log / frontier entry (cap / 2 per list)
  uint64_t cap;         // table slots, power of two
  uint64_t cur, step;   // pivot window schedule
  double beta;
  uint32_t g;
  uint32_t pad;
};

struct TailState {    // written back per graph
  unsigned long long cur, step, rounds, full;
};

// lid[centers[j]] = j, compact labels, and the compact degrees: edges to
// vertices that are unfinished and carry the same label (one wave per
// remaining vertex, 64-edge windows).
__global__ void k_tail_degrees(const uint32_t *centers, uint64_t nrem,
                               const unsigned long long *labels,
                               const uint64_t *off, const uint32_t *adj,
                               const uint64_t *ioff, const uint32_t *iadj,
                               uint32_t *cnt_out, uint32_t *cnt_in) {
  const int l = lane();
  for (uint64_t j = (blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) / WAVE; j < nrem;
       j += (uint64_t)gridDim.x * blockDim.x / WAVE) {
    const uint32_t v = centers[j];
    const unsigned long long lv = labels[v];
    for (int dir = 0; dir < 2; ++dir) {
      const uint64_t *o = dir ? ioff : off;
      const uint32_t *a = dir ? iadj : adj;
      const uint64_t e0 = o[v], e1 = o[v + 1];
      uint32_t c = 0;
      for (uint64_t e = e0; e < e1; e += WAVE) {
        bool keep = false;
        if (e + l < e1) {
          const unsigned long long lx = labels[a[e + l]];
          keep = !(lx & TOP_BIT) && lx == lv;
        }
        c += (uint32_t)__popcll(ballot(keep));
      }
      if (l == 0) (dir ? cnt_in : cnt_out)[j] = c;
    }
  }
}

__global__ void k_tail_lid(const uint32_t *centers, uint64_t nrem,
                           const unsigned long long *labels, uint32_t *lid,
                           unsigned long long *tlab) {
  for (uint64_t j = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; j < nrem;
       j += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t v = centers[j];
    lid[v] = (uint32_t)j;
    tlab[j] = labels[v];
  }
}

// Compact adjacency in original order (ballot ranks within each window).
__global__ void k_tail_fill(const uint32_t *centers, uint64_t nrem,
                            const unsigned long long *labels, const uint32_t *lid,
                            const uint64_t *off, const uint32_t *adj,
                            const uint64_t *ioff, const uint32_t *iadj,
                            const uint32_t *toff, const uint32_t *tioff,
                            uint32_t *tadj, uint32_t *tiadj) {
  const int l = lane();
  for (uint64_t j = (blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) / WAVE; j < nrem;
       j += (uint64_t)gridDim.x * blockDim.x / WAVE) {
    const uint32_t v = centers[j];
    const unsigned long long lv = labels[v];
    for (int dir = 0; dir < 2; ++dir) {
      const uint64_t *o = dir ? ioff : off;
      const uint32_t *a = dir ? iadj : adj;
      uint32_t *out = dir ? tiadj : tadj;
      uint32_t at = (dir ? tioff : toff)[j];
      const uint64_t e0 = o[v], e1 = o[v + 1];
      for (uint64_t e = e0; e < e1; e += WAVE) {
        bool keep = false;
        uint32_t x = 0;
        if (e + l < e1) {
          x = a[e + l];
          const unsigned long long lx = labels[x];
          keep = !(lx & TOP_BIT) && lx == lv;
        }
        const uint64_t m = ballot(keep);
        if (keep) out[at + lanes_below(m)] = lid[x];
        at += (uint32_t)__popcll(m);
      }
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

struct TailArgs {
  const TailGraph *graphs;
  TailState *state;
  const uint32_t *toff, *tadj, *tioff, *tiadj;
  unsigned long long *tlab;
  unsigned long long *table;   // 2 cap slots per graph, all EMPTY on entry
  uint32_t *lists;             // 4 lists of cap / 2 entries per graph
  unsigned long long *lab_off; // per graph label counter
  unsigned long long *stats;   // [0] inserts, [1] edges
};

// One workgroup per graph: all rounds of its remainder.
__global__ __launch_bounds__(TAIL_BLOCK) void k_ms_tail(TailArgs A) {
  __shared__ uint32_t wsum[TAIL_BLOCK / WAVE];
  __shared__ unsigned int cnt_q, cnt_log, flag_full;
  __shared__ unsigned int front_n;
  const TailGraph tg = A.graphs[blockIdx.x];
  const uint32_t tid = threadIdx.x, w = tid / WAVE;
  const uint64_t base = tg.base, nr = tg.nr, half = tg.cap / 2, mask = tg.cap - 1;
  unsigned long long *const T0 = A.table + tg.tbase, *const T1 = T0 + tg.cap;
  uint32_t *const L0 = A.lists + tg.lbase;   // slot indices of inserted pairs
  uint32_t *const L1 = L0 + half;
  uint32_t *const qa = L0 + 2 * half, *const qb = L0 + 3 * half;   // frontiers
  unsigned long long *lab = A.tlab;
  uint64_t cur = tg.cur, step = tg.step;
  unsigned long long loff = A.lab_off[tg.g];
  unsigned long long rounds = 0, ins = 0, edges = 0;
  uint32_t nlog0 = 0, nlog1 = 0;
  bool full = false;
  while (cur < nr && !full) {
    const uint64_t end = min(cur + step, nr);
    // ---- front: unfinished vertices of the window, in order; seeds
    // (pivot_i, loff + i) in both tables (src/scc.hpp:147-150)
    if (tid == 0) {
      front_n = 0;
      cnt_log = 0;
      flag_full = 0;
    }
    __syncthreads();
    for (uint64_t b0 = cur; b0 < end; b0 += TAIL_BLOCK) {
      const uint64_t j = b0 + tid;
      const bool take = j < end && !(lab[base + j] & TOP_BIT);
      const uint64_t m = ballot(take);
      if ((tid & (WAVE - 1)) == 0) wsum[w] = (uint32_t)__popcll(m);
      __syncthreads();
      uint32_t before = 0, tot = 0;
      for (uint32_t x = 0; x < TAIL_BLOCK / WAVE; ++x) {
        before += x < w ? wsum[x] : 0u;
        tot += wsum[x];
      }
      const uint32_t i = front_n + before + lanes_below(m);
      if (take) {
        const uint32_t v = (uint32_t)(base + j), lb = (uint32_t)(loff + i);
        uint64_t s0 = 0, s1 = 0;
        const int r0 = tail_insert(T0, mask, v, lb, &s0);
        const int r1 = tail_insert(T1, mask, v, lb, &s1);
        if (r0 < 0 || r1 < 0 || i >= half) {
          flag_full = 1;
        } else {
          L0[i] = (uint32_t)s0;
          L1[i] = (uint32_t)s1;
        }
      }
      __syncthreads();
      if (tid == 0) front_n += tot;
      __syncthreads();
    }
    const uint32_t nfront = front_n;
    // ---- forward and backward multi-search over pairs (src/scc.hpp:160-252)
    for (int d = 0; d < 2 && !flag_full; ++d) {
      unsigned long long *const Td = d ? T1 : T0;
      uint32_t *const Ld = d ? L1 : L0;
      const uint32_t *const eo = d ? A.tioff : A.toff;
      const uint32_t *const ea = d ? A.tiadj : A.tadj;
      uint32_t nl = nfront;
      for (uint32_t i = tid; i < nfront; i += TAIL_BLOCK) qa[i] = Ld[i];
      uint32_t *qc = qa, *qn = qb;
      uint32_t nq = nfront;
      __syncthreads();
      while (nq && !flag_full) {
        if (tid == 0) {
          cnt_q = 0;
          cnt_log = nl;
        }
        __syncthreads();
        for (uint32_t i = tid; i < nq; i += TAIL_BLOCK) {
          const unsigned long long pr = Td[qc[i]];
          const uint32_t v = pair_vertex(pr), lb = pair_label(pr);
          const unsigned long long lv = lab[v];
          const uint32_t e1 = eo[v + 1];
          for (uint32_t e = eo[v]; e < e1; ++e) {
            const uint32_t x = ea[e];
            ++edges;
            const unsigned long long lx = lab[x];
            if ((lx & TOP_BIT) || lx != lv) continue;
            uint64_t slot = 0;
            const int r = tail_insert(Td, mask, x, lb, &slot);
            if (r < 0) {
              flag_full = 1;
              break;
            }
            if (r > 0) {
              ++ins;
              const uint32_t at = atomicAdd(&cnt_log, 1u);
              const uint32_t aq = atomicAdd(&cnt_q, 1u);
              if (at >= half || aq >= half) {
                flag_full = 1;
                break;
              }
              Ld[at] = (uint32_t)slot;
              qn[aq] = (uint32_t)slot;
            }
          }
        }
        __syncthreads();
        nq = min(cnt_q, (unsigned int)half);
        nl = min(cnt_log, (unsigned int)half);
        uint32_t *t = qc;
        qc = qn;
        qn = t;
        __syncthreads();
      }
      if (d) nlog1 = nl; else nlog0 = nl;
    }
    __syncthreads();
    full = flag_full;
    if (!full) {
      // ---- partition (src/scc.hpp:437-464): a pair in both tables is in
      // the pivot's SCC; labels take the maximum
      for (uint32_t i = tid; i < nlog0; i += TAIL_BLOCK) {
        const unsigned long long pr = T0[L0[i]];
        const uint32_t v = pair_vertex(pr), lb = pair_label(pr);
        uint32_t probes = 0;
        const unsigned long long val =
            table_contains(T1, mask, v, lb, &probes) ? ((unsigned long long)lb | TOP_BIT)
                                                     : (unsigned long long)lb;
        if (lab[v] < val) atomicMax(&lab[v], val);
      }
      for (uint32_t i = tid; i < nlog1; i += TAIL_BLOCK) {
        const unsigned long long pr = T1[L1[i]];
        const unsigned long long val = pair_label(pr);
        const uint32_t v = pair_vertex(pr);
        if (lab[v] < val) atomicMax(&lab[v], val);
      }
      loff += nfront;
      cur = end;
      step = (uint64_t)ceil((double)step * tg.beta);
      ++rounds;
    }
    __syncthreads();
    // ---- empty the slots this round used
    if (!full) {
      for (uint32_t i = tid; i < nlog0; i += TAIL_BLOCK) T0[L0[i]] = EMPTY_SLOT;
      for (uint32_t i = tid; i < nlog1; i += TAIL_BLOCK) T1[L1[i]] = EMPTY_SLOT;
    }
    __syncthreads();
  }
  if (full) {
    // slots of a stopped round are not all logged: clear the whole region
    for (uint64_t i = tid; i < 2 * tg.cap; i += TAIL_BLOCK) T0[i] = EMPTY_SLOT;
  }
  if (tid == 0) {
    A.state[blockIdx.x] = TailState{cur, step, rounds, full ? 1ull : 0ull};
    A.lab_off[tg.g] = loff;
  }
  ins = wave_sum(ins);
  edges = wave_sum(edges);
  if ((tid & (WAVE - 1)) == 0) {
    if (ins) atomicAdd(&A.stats[0], ins);
    if (edges) atomicAdd(&A.stats[1], edges);
  }
}

// Compact edge count (both directions) of each tail graph; out[count] and
// out[count + 1] get the total out- and in-edge counts.
__global__ void k_tail_sizes(const TailGraph *tg, uint32_t count, const uint32_t *toff,
                             const uint32_t *tioff, uint64_t nrem,
                             unsigned long long *out) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i <= count;
       i += gridDim.x * blockDim.x) {
    if (i == count) {
      out[count] = toff[nrem];
      out[count + 1] = tioff[nrem];
      continue;
    }
    const uint64_t b = tg[i].base, e = b + tg[i].nr;
    out[i] = (unsigned long long)(toff[e] - toff[b]) + (tioff[e] - tioff[b]);
  }
}

__global__ void k_tail_writeback(const uint32_t *centers, uint64_t nrem,
                                 const unsigned long long *tlab,
                                 unsigned long long *labels) {
  for (uint64_t j = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; j < nrem;
       j += (uint64_t)gridDim.x * blockDim.x)
    labels[centers[j]] = tlab[j];
}

__global__ void k_tail_fill_u64(unsigned long long *p, uint64_t n) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = EMPTY_SLOT;
}

int blocks(uint64_t n, int per = 256, int cap = 16384) {
  uint64_t g = (n + per - 1) / per;
  return (int)std::max<uint64_t>(1, std::min<uint64_t>(g, cap));
}

}  // namespace

// Runs the rounds of every graph with 0 < remainder <= p.tail_max on the
// device; cur / step of those graphs are advanced (fully, unless a graph's
// tables overflowed).  Returns PSC_OK with *used = false when no graph
// qualifies.
int Engine::ms_tail(const psc_params &p, const uint64_t *h_cstart, uint64_t n_rem,
                    std::vector<uint64_t> &cur, std::vector<uint64_t> &stepg,
                    psc_stats *st, bool *used) {
  *used = false;
  const double beta_def = p.beta > 0 ? p.beta : 1.5;
  std::vector<TailGraph> tgs;
  for (uint32_t g = 0; g < G; ++g) {
    const uint64_t r = h_cstart[g + 1] - h_cstart[g];
    if (r == 0 || cur[g] == r || r > p.tail_max) continue;
    TailGraph t{};
    t.base = h_cstart[g];
    t.nr = r;
    t.cur = cur[g];
    t.step = stepg[g];
    t.beta = beta_g.empty() ? beta_def : beta_g[g];
    t.g = g;
    tgs.push_back(t);
  }
  if (tgs.empty() || n_rem == 0) return PSC_OK;
  const unsigned long long *lab = labels.as<unsigned long long>();
  const uint32_t *cen = centers.as<uint32_t>();
  // ---- compact remainder: ids by pivot position, same-subproblem edges
  PSC_TRY(tail_cnt.ensure(2 * (n_rem + 1) * 4));
  PSC_TRY(tail_lab.ensure(n_rem * 8));
  uint32_t *cnt_out = tail_cnt.as<uint32_t>(), *cnt_in = cnt_out + (n_rem + 1);
  uint32_t *lid = scratch32.as<uint32_t>();
  prof_begin(KK_TAIL);
  k_tail_lid<<<blocks(n_rem), 256, 0, stream>>>(cen, n_rem, lab, lid,
                                                tail_lab.as<unsigned long long>());
  k_tail_degrees<<<blocks(n_rem * WAVE), 256, 0, stream>>>(
      cen, n_rem, lab, d_off, d_adj, d_ioff, d_iadj, cnt_out, cnt_in);
  PSC_HIP(hipMemsetAsync(cnt_out + n_rem, 0, 4, stream));
  PSC_HIP(hipMemsetAsync(cnt_in + n_rem, 0, 4, stream));
  PSC_TRY(tail_off.ensure(2 * (n_rem + 1) * 4));
  uint32_t *toff = tail_off.as<uint32_t>(), *tioff = toff + (n_rem + 1);
  size_t tb = 0, tb2 = 0;
  PSC_HIP(rocprim::exclusive_scan(nullptr, tb, cnt_out, toff, 0u, n_rem + 1,
                                  rocprim::plus<uint32_t>(), stream));
  PSC_HIP(rocprim::exclusive_scan(nullptr, tb2, cnt_in, tioff, 0u, n_rem + 1,
                                  rocprim::plus<uint32_t>(), stream));
  PSC_TRY(sort_tmp.ensure(std::max<size_t>(std::max(tb, tb2), 8)));
  tb = sort_tmp.cap;
  PSC_HIP(rocprim::exclusive_scan(sort_tmp.ptr, tb, cnt_out, toff, 0u, n_rem + 1,
                                  rocprim::plus<uint32_t>(), stream));
  tb = sort_tmp.cap;
  PSC_HIP(rocprim::exclusive_scan(sort_tmp.ptr, tb, cnt_in, tioff, 0u, n_rem + 1,
                                  rocprim::plus<uint32_t>(), stream));
  prof_end();
  // per-graph compact edge counts size the pair tables
  const size_t nt = tgs.size();
  PSC_TRY(tail_meta.ensure(nt * (sizeof(TailGraph) + sizeof(TailState) + 8) + 64));
  TailGraph *d_tg = tail_meta.as<TailGraph>();
  TailState *d_ts = reinterpret_cast<TailState *>(d_tg + nt);
  unsigned long long *d_e = reinterpret_cast<unsigned long long *>(d_ts + nt);
  unsigned long long *d_stats = d_e + nt + 2;
  PSC_HIP(hipMemcpyAsync(d_tg, tgs.data(), nt * sizeof(TailGraph), hipMemcpyHostToDevice,
                         stream));
  k_tail_sizes<<<blocks(nt + 1), 256, 0, stream>>>(d_tg, (uint32_t)nt, toff, tioff, n_rem,
                                                   d_e);
  std::vector<unsigned long long> h_e(nt + 2);
  PSC_TRY(poll(h_e.data(), d_e, (nt + 2) * 8));
  const uint64_t e_out = h_e[nt], e_in = h_e[nt + 1];
  PSC_TRY(tail_adj.ensure((e_out + e_in + 2) * 4));
  uint32_t *tadj = tail_adj.as<uint32_t>(), *tiadj = tadj + e_out + 1;
  prof_begin(KK_TAIL);
  k_tail_fill<<<blocks(n_rem * WAVE), 256, 0, stream>>>(
      cen, n_rem, lab, lid, d_off, d_adj, d_ioff, d_iadj, toff, tioff, tadj, tiadj);
  prof_end();
  uint64_t slots = 0;
  for (size_t i = 0; i < tgs.size(); ++i) {
    TailGraph &t = tgs[i];
    uint64_t cap = 1024;
    while (cap < 4 * (t.nr + h_e[i])) cap *= 2;
    if (p.table_init) {   // forced small tables: exercises the hand-over
      cap = 16;
      while (cap < p.table_init) cap *= 2;
    }
    t.cap = cap;
    t.tbase = slots;   // two tables of cap slots
    t.lbase = slots;   // four lists of cap / 2 u32
    slots += 2 * cap;
  }
  PSC_TRY(tail_table.ensure(slots * 8));
  PSC_TRY(tail_lists.ensure(slots * 4));
  PSC_HIP(hipMemcpyAsync(d_tg, tgs.data(), nt * sizeof(TailGraph), hipMemcpyHostToDevice,
                         stream));
  PSC_HIP(hipMemsetAsync(d_stats, 0, 16, stream));
  prof_begin(KK_TAIL);
  k_tail_fill_u64<<<blocks(slots), 256, 0, stream>>>(tail_table.as<unsigned long long>(),
                                                     slots);
  TailArgs a;
  a.graphs = d_tg;
  a.state = d_ts;
  a.toff = toff;
  a.tadj = tadj;
  a.tioff = tioff;
  a.tiadj = tiadj;
  a.tlab = tail_lab.as<unsigned long long>();
  a.table = tail_table.as<unsigned long long>();
  a.lists = tail_lists.as<uint32_t>();
  a.lab_off = gD.as<unsigned long long>();
  a.stats = d_stats;
  k_ms_tail<<<(unsigned)tgs.size(), TAIL_BLOCK, 0, stream>>>(a);
  k_tail_writeback<<<blocks(n_rem), 256, 0, stream>>>(
      cen, n_rem, tail_lab.as<unsigned long long>(), labels.as<unsigned long long>());
  prof_end();
  PSC_HIP(hipGetLastError());
  std::vector<TailState> hs(tgs.size());
  PSC_TRY(poll(hs.data(), d_ts, tgs.size() * sizeof(TailState)));
  unsigned long long hst[2] = {0, 0};
  PSC_TRY(poll(hst, d_stats, 16));
  uint64_t max_rounds = 0;
  for (size_t i = 0; i < tgs.size(); ++i) {
    cur[tgs[i].g] = hs[i].cur;
    stepg[tgs[i].g] = hs[i].step;
    max_rounds = std::max<uint64_t>(max_rounds, hs[i].rounds);
    if (hs[i].full) ++st->resizes;
  }
  st->rounds += max_rounds;
  st->table_inserts += hst[0];
  st->edges_scanned += hst[1];
  st->bytes_csr += 8 * (e_out + e_in) + 8 * hst[1];
  st->bytes_label += 16 * hst[1];
  st->bytes_table += 16 * hst[0];
  *used = true;
  return PSC_OK;
}

}  // namespace psc
