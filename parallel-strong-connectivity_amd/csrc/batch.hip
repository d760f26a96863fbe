// Batched SCC: many graphs in one handle.
//
// Input is the packed form (per-graph CSRs back to back, uint32 local ids,
// per-graph vertex / edge prefixes).  On the device the batch becomes one
// disjoint union with global ids -- offsets rebased by the edge prefix,
// neighbour ids by the vertex prefix -- so every kernel of the single-graph
// path runs over all graphs at once, while the engine keeps each graph's
// trim numbering, pivot permutation, label namespace and pivot-batch
// schedule separate.  Each graph's labels therefore equal its stand-alone
// run bit for bit.
//
// The seeded batch generator builds the mixed workload (RMAT with average
// out-degree 16, directed grids, directed 3D k-NN) straight into the union
// layout with the same counter-based draws as the host generators.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_select.hpp>

#include "../../include/psc.h"
#include "device_common.hpp"
#include "engine.hpp"
#include "gen_common.hpp"

namespace {

constexpr unsigned long long DROP = ~0ull;

#define B_HIP(call)                                                      \
  do {                                                                   \
    hipError_t e_ = (call);                                              \
    if (e_ != hipSuccess) {                                              \
      psc::set_error(std::string(#call) + ": " + hipGetErrorString(e_)); \
      rc = PSC_ERR_HIP;                                                  \
      goto done;                                                         \
    }                                                                    \
  } while (0)

int blocks_for(uint64_t n) {
  uint64_t g = (n + 255) / 256;
  return (int)std::min<uint64_t>(std::max<uint64_t>(g, 1), 65536);
}

// global offsets: goff[vbase[g] + i] = ebase[g] + loff[vbase[g] + g + i]
__global__ void k_rebase_offsets(const uint64_t *loff, const uint64_t *vbase,
                                 const uint64_t *ebase, uint32_t G,
                                 uint64_t ntot, uint64_t *goff) {
  for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v <= ntot;
       v += (uint64_t)gridDim.x * blockDim.x) {
    if (v == ntot) {
      goff[v] = ebase[G];
      continue;
    }
    const uint32_t g = psc::seg_of((const unsigned long long *)vbase, G, v);
    goff[v] = ebase[g] + loff[v + g];
  }
}

// global neighbour ids: gadj[j] = vbase[g] + ladj[j], g owning edge j
__global__ void k_rebase_edges(const uint32_t *ladj, const uint64_t *vbase,
                               const uint64_t *ebase, uint32_t G, uint64_t mtot,
                               uint32_t *gadj) {
  for (uint64_t j = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; j < mtot;
       j += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t g = psc::seg_of((const unsigned long long *)ebase, G, j);
    gadj[j] = (uint32_t)(vbase[g] + ladj[j]);
  }
}

// local copy of graph g out of the union (for the one-graph-per-launch path)
__global__ void k_extract(const uint64_t *goff, const uint32_t *gadj,
                          uint64_t v0, uint64_t n, uint64_t e0, uint64_t m,
                          uint64_t *loff, uint32_t *ladj) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i <= n;
       i += (uint64_t)gridDim.x * blockDim.x)
    loff[i] = goff[v0 + i] - e0;
  for (uint64_t j = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; j < m;
       j += (uint64_t)gridDim.x * blockDim.x)
    ladj[j] = (uint32_t)(gadj[e0 + j] - v0);
}

__global__ void k_fill_one(uint64_t *p, uint64_t v) { *p = v; }

// one device graph into its slot of a union: source offsets start at
// `se0` and source ids at `sv0` (0 for a stand-alone graph, the graph's
// prefixes for a view into another union)
__global__ void k_rebase_one(const uint64_t *loff, const uint32_t *ladj,
                             uint64_t n, uint64_t m, uint64_t v0, uint64_t e0,
                             uint64_t sv0, uint64_t se0, uint64_t *goff,
                             uint32_t *gadj) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i <= n;
       i += (uint64_t)gridDim.x * blockDim.x)
    goff[i] = e0 + (loff[i] - se0);
  for (uint64_t j = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; j < m;
       j += (uint64_t)gridDim.x * blockDim.x)
    gadj[j] = (uint32_t)(v0 + (ladj[j] - sv0));
}

template <class T>
int dev_copy_in(const T *src, size_t count, int on_device, T **dst) {
  psc::GateGuard gate;
  *dst = nullptr;
  if (hipMalloc(dst, std::max<size_t>(count, 1) * sizeof(T)) != hipSuccess) {
    psc::set_error("hipMalloc (batch input)");
    return PSC_ERR_NOMEM;
  }
  if (count &&
      hipMemcpy(*dst, src, count * sizeof(T),
                on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice) !=
          hipSuccess) {
    psc::set_error("hipMemcpy (batch input)");
    return PSC_ERR_HIP;
  }
  return PSC_OK;
}

// Device batches live in the device's stream-ordered memory pool (the
// engine's buffers do too): a plain hipFree waits for all work on the
// device, which would stall every other handle's stream each time a batch
// is released.  Frees are ordered on a library stream.
template <class T>
static hipError_t pool_alloc(T **p, size_t bytes, hipStream_t st) {
  psc::GateGuard gate;
  return hipMallocAsync(reinterpret_cast<void **>(p), std::max<size_t>(bytes, 8), st);
}

static void pool_free(void *p) {
  psc::GateGuard gate;
  if (p) (void)hipFreeAsync(p, psc::pool_free_stream());
}

__global__ void k_gather_at(const uint64_t *__restrict__ src,
                            const uint64_t *__restrict__ idx, uint32_t count,
                            uint64_t *__restrict__ out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) out[i] = src[idx[i]];
}

// ------------------------------------------------------ batch generator
enum : uint32_t { KIND_RMAT = 0, KIND_GRID = 1, KIND_KNN = 2 };

struct GenGraph {   // one graph of the generated batch (device copy)
  uint64_t vbase, nitems, ibase;   // items: RMAT samples / grid pairs / points
  uint64_t kbase;                  // first key slot of the graph
  uint64_t seed;
  uint32_t kind, scale, rows, cols;
  uint64_t cbase;   // k-NN: first global cell
  uint32_t cells1;  // k-NN: cells per axis
  uint32_t pad;
  double cw;        // k-NN: cell width
};

__device__ __forceinline__ uint32_t find_graph(const GenGraph *gg, uint32_t G,
                                               uint64_t item) {
  uint32_t lo = 0, hi = G - 1;
  while (lo < hi) {
    const uint32_t mid = (lo + hi + 1) >> 1;
    if (gg[mid].ibase <= item) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// RMAT samples and grid pairs of all non-k-NN graphs -> (u << 32 | v) keys
// with global ids (one key per RMAT sample, two per grid pair; DROP marks
// self-loops and absent directions).
__global__ void k_batch_keys(const GenGraph *gg, uint32_t G, uint64_t items,
                             uint32_t t0, uint32_t t1, uint32_t t2,
                             uint32_t thr_both, unsigned long long *keys) {
  for (uint64_t it = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; it < items;
       it += (uint64_t)gridDim.x * blockDim.x) {
    const GenGraph g = gg[find_graph(gg, G, it)];
    const uint64_t i = it - g.ibase;
    unsigned long long a = DROP, b = DROP;
    if (g.kind == KIND_RMAT) {
      uint32_t u, v;
      psc::rmat_sample(g.seed, i, (int)g.scale, t0, t1, t2, &u, &v);
      const psc::IdPerm perm((int)g.scale, g.seed);
      u = (uint32_t)perm(u);
      v = (uint32_t)perm(v);
      if (u != v) a = ((g.vbase + u) << 32) | (g.vbase + v);
      keys[g.kbase + i] = a;
    } else {
      const uint64_t cols = g.cols, rows = g.rows;
      const uint64_t nh = rows * (cols - 1);
      uint64_t x, y;
      if (i < nh) {
        const uint64_t r = i / (cols - 1), c = i % (cols - 1);
        x = r * cols + c;
        y = x + 1;
      } else {
        const uint64_t q = i - nh, r = q / cols, c = q % cols;
        x = r * cols + c;
        y = x + cols;
      }
      const int kind = psc::grid_pair_kind(g.seed, i, thr_both);
      x += g.vbase;
      y += g.vbase;
      if (kind != 2) a = (x << 32) | y;
      if (kind != 1) b = (y << 32) | x;
      keys[g.kbase + 2 * i] = a;
      keys[g.kbase + 2 * i + 1] = b;
    }
  }
}

__device__ __forceinline__ uint32_t knn_cell(uint32_t x, double cw, uint32_t G1) {
  uint32_t c = (uint32_t)((double)x / cw);
  return c >= G1 ? G1 - 1 : c;
}

// k-NN points -> (global cell << 32 | global point) for the cell sort
__global__ void k_knn_cellkeys(const GenGraph *gg, uint32_t G, uint64_t items,
                               unsigned long long *keys) {
  for (uint64_t it = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; it < items;
       it += (uint64_t)gridDim.x * blockDim.x) {
    const GenGraph g = gg[find_graph(gg, G, it)];
    const uint64_t i = it - g.ibase;
    const uint32_t c0 = knn_cell(psc::knn_coord(g.seed, i, 0), g.cw, g.cells1);
    const uint32_t c1 = knn_cell(psc::knn_coord(g.seed, i, 1), g.cw, g.cells1);
    const uint32_t c2 = knn_cell(psc::knn_coord(g.seed, i, 2), g.cw, g.cells1);
    const uint64_t cell = g.cbase + ((uint64_t)c0 * g.cells1 + c1) * g.cells1 + c2;
    keys[it] = (cell << 32) | (g.vbase + i);
  }
}

// cstart[c] = first sorted index with cell >= c
__global__ void k_cell_starts(const unsigned long long *sorted, uint64_t items,
                              uint64_t ncells, uint64_t *cstart) {
  for (uint64_t c = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; c <= ncells;
       c += (uint64_t)gridDim.x * blockDim.x) {
    const unsigned long long t = (unsigned long long)c << 32;
    uint64_t lo = 0, hi = items;
    while (lo < hi) {
      const uint64_t mid = (lo + hi) >> 1;
      if (sorted[mid] < t) lo = mid + 1; else hi = mid;
    }
    cstart[c] = lo;
  }
}

// Exact k nearest neighbours (k <= 8) by shell search over the cell grid,
// the same search and stopping rule as the host generator; the result is
// the k smallest (distance, id) pairs, so it does not depend on visiting
// order.
__global__ void k_knn_search(const GenGraph *gg, uint32_t G, uint64_t items,
                             const unsigned long long *sorted,
                             const uint64_t *cstart, uint32_t k,
                             unsigned long long *keys) {
  for (uint64_t it = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; it < items;
       it += (uint64_t)gridDim.x * blockDim.x) {
    const GenGraph g = gg[find_graph(gg, G, it)];
    const uint64_t i = it - g.ibase;
    const int64_t xi = psc::knn_coord(g.seed, i, 0), yi = psc::knn_coord(g.seed, i, 1),
                  zi = psc::knn_coord(g.seed, i, 2);
    const int G1 = (int)g.cells1;
    const int cx = (int)knn_cell((uint32_t)xi, g.cw, g.cells1);
    const int cy = (int)knn_cell((uint32_t)yi, g.cw, g.cells1);
    const int cz = (int)knn_cell((uint32_t)zi, g.cw, g.cells1);
    uint64_t bd[8];
    uint32_t bj[8];
    uint32_t nb = 0;
    for (int r = 0;; ++r) {
      for (int dx = -r; dx <= r; ++dx)
        for (int dy = -r; dy <= r; ++dy)
          for (int dz = -r; dz <= r; ++dz) {
            const int ax = dx < 0 ? -dx : dx, ay = dy < 0 ? -dy : dy,
                      az = dz < 0 ? -dz : dz;
            if (max(ax, max(ay, az)) != r) continue;
            const int x = cx + dx, y = cy + dy, z = cz + dz;
            if (x < 0 || y < 0 || z < 0 || x >= G1 || y >= G1 || z >= G1) continue;
            const uint64_t c = g.cbase + ((uint64_t)x * G1 + y) * G1 + z;
            for (uint64_t t = cstart[c]; t < cstart[c + 1]; ++t) {
              const uint32_t jg = (uint32_t)sorted[t];
              const uint64_t j = jg - g.vbase;
              if (j == i) continue;
              const int64_t ddx = xi - (int64_t)psc::knn_coord(g.seed, j, 0);
              const int64_t ddy = yi - (int64_t)psc::knn_coord(g.seed, j, 1);
              const int64_t ddz = zi - (int64_t)psc::knn_coord(g.seed, j, 2);
              const uint64_t d = (uint64_t)(ddx * ddx + ddy * ddy + ddz * ddz);
              // insert (d, j) into the sorted top-k
              if (nb == k && (d > bd[k - 1] || (d == bd[k - 1] && j > bj[k - 1])))
                continue;
              int pos = (int)(nb < k ? nb : k - 1);
              while (pos > 0 && (bd[pos - 1] > d || (bd[pos - 1] == d && bj[pos - 1] > j))) {
                bd[pos] = bd[pos - 1];
                bj[pos] = bj[pos - 1];
                --pos;
              }
              bd[pos] = d;
              bj[pos] = (uint32_t)j;
              if (nb < k) ++nb;
            }
          }
      if (nb == k) {
        const double reach = (double)r * g.cw;
        if (reach * reach > (double)bd[k - 1]) break;
      }
      if (r > G1) break;
    }
    for (uint32_t t = 0; t < k; ++t)
      keys[g.kbase + i * k + t] =
          t < nb ? (((g.vbase + i) << 32) | (g.vbase + bj[t])) : DROP;
  }
}

__global__ void k_low32(const unsigned long long *keys, uint32_t *E, uint64_t m) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < m;
       i += (uint64_t)gridDim.x * blockDim.x)
    E[i] = (uint32_t)keys[i];
}
__global__ void k_swap32(const unsigned long long *in, unsigned long long *out,
                         uint64_t m) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < m;
       i += (uint64_t)gridDim.x * blockDim.x)
    out[i] = (in[i] << 32) | (in[i] >> 32);
}
// off[i] = e0 + first index with key >= (v0 + i) << 32, i in [0, n]
__global__ void k_offsets_from_keys(const unsigned long long *keys, uint64_t m,
                                    uint64_t n, uint64_t *off, uint64_t v0,
                                    uint64_t e0) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i <= n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const unsigned long long t = (unsigned long long)(v0 + i) << 32;
    uint64_t lo = 0, hi = m;
    while (lo < hi) {
      const uint64_t mid = (lo + hi) >> 1;
      if (keys[mid] < t) lo = mid + 1; else hi = mid;
    }
    off[i] = e0 + lo;
  }
}

}  // namespace

struct psc_dev_batch {
  psc_graph g;                   // the union, device pointers
  uint64_t num_graphs;
  std::vector<uint64_t> vbase, ebase;
  std::vector<uint32_t> kind;
};

extern "C" {

int psc_batch_create(const psc_batch_desc *b, int on_device, psc_handle **out) {
  psc::GateGuard gate;
  if (!b || !out || b->num_graphs == 0 || b->num_graphs > 0xFFFFFFF0ull)
    return PSC_ERR_INVALID;
  *out = nullptr;
  const uint32_t G = (uint32_t)b->num_graphs;
  std::vector<uint64_t> vb(G + 1), eb(G + 1);
  if (on_device) {
    if (hipMemcpy(vb.data(), b->vbase, (G + 1) * 8, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(eb.data(), b->ebase, (G + 1) * 8, hipMemcpyDeviceToHost) != hipSuccess) {
      psc::set_error("hipMemcpy (batch prefixes)");
      return PSC_ERR_HIP;
    }
  } else {
    std::memcpy(vb.data(), b->vbase, (G + 1) * 8);
    std::memcpy(eb.data(), b->ebase, (G + 1) * 8);
  }
  for (uint32_t g = 0; g < G; ++g)
    if (vb[g + 1] < vb[g] || eb[g + 1] < eb[g] || vb[0] || eb[0]) {
      psc::set_error("batch prefixes must start at 0 and be non-decreasing");
      return PSC_ERR_INVALID;
    }
  const uint64_t ntot = vb[G], mtot = eb[G];
  if (ntot >= 0xFFFFFFFFull) {
    psc::set_error("batch too large: total vertices must be < 2^32-1");
    return PSC_ERR_INVALID;
  }
  if (!on_device) {   // validate the packed host input
    for (uint32_t g = 0; g < G; ++g) {
      const uint64_t n = vb[g + 1] - vb[g], m = eb[g + 1] - eb[g];
      const uint64_t *o = b->offset + vb[g] + g, *io = b->in_offset + vb[g] + g;
      if (o[0] != 0 || o[n] != m || io[0] != 0 || io[n] != m) {
        psc::set_error("graph " + std::to_string(g) + ": offsets must run 0..m_g");
        return PSC_ERR_INVALID;
      }
      for (uint64_t j = eb[g]; j < eb[g + 1]; ++j)
        if (b->edges[j] >= n || b->in_edges[j] >= n) {
          psc::set_error("graph " + std::to_string(g) + ": edge endpoint out of range");
          return PSC_ERR_INVALID;
        }
    }
  }
  int rc = PSC_OK;
  uint64_t *d_vb = nullptr, *d_eb = nullptr, *d_lo = nullptr, *d_lio = nullptr;
  uint32_t *d_le = nullptr, *d_lie = nullptr;
  auto *h = new (std::nothrow) psc_handle;
  psc::Engine *e = h ? &h->eng : nullptr;
  if (!h) return PSC_ERR_NOMEM;
  rc = e->setup();
  if (!rc) rc = dev_copy_in(vb.data(), G + 1, 0, &d_vb);
  if (!rc) rc = dev_copy_in(eb.data(), G + 1, 0, &d_eb);
  if (!rc) rc = dev_copy_in(b->offset, ntot + G, on_device, &d_lo);
  if (!rc) rc = dev_copy_in(b->in_offset, ntot + G, on_device, &d_lio);
  if (!rc) rc = dev_copy_in(b->edges, mtot, on_device, &d_le);
  if (!rc) rc = dev_copy_in(b->in_edges, mtot, on_device, &d_lie);
  if (!rc) rc = e->g_off.ensure((ntot + 1) * 8);
  if (!rc) rc = e->g_ioff.ensure((ntot + 1) * 8);
  if (!rc) rc = e->g_adj.ensure(std::max<uint64_t>(mtot, 1) * 4);
  if (!rc) rc = e->g_iadj.ensure(std::max<uint64_t>(mtot, 1) * 4);
  if (rc) goto done;
  k_rebase_offsets<<<blocks_for(ntot + 1), 256, 0, e->stream>>>(
      d_lo, d_vb, d_eb, G, ntot, e->g_off.as<uint64_t>());
  k_rebase_offsets<<<blocks_for(ntot + 1), 256, 0, e->stream>>>(
      d_lio, d_vb, d_eb, G, ntot, e->g_ioff.as<uint64_t>());
  if (mtot) {
    k_rebase_edges<<<blocks_for(mtot), 256, 0, e->stream>>>(d_le, d_vb, d_eb, G, mtot,
                                                            e->g_adj.as<uint32_t>());
    k_rebase_edges<<<blocks_for(mtot), 256, 0, e->stream>>>(d_lie, d_vb, d_eb, G, mtot,
                                                            e->g_iadj.as<uint32_t>());
  }
  B_HIP(hipGetLastError());
  B_HIP(hipStreamSynchronize(e->stream));
  {
    psc_graph ug{ntot, mtot, e->g_off.as<uint64_t>(), e->g_adj.as<uint32_t>(),
                 e->g_ioff.as<uint64_t>(), e->g_iadj.as<uint32_t>()};
    rc = e->init(&ug, true, G, vb.data());
  }
done:
  for (void *p : {(void *)d_vb, (void *)d_eb, (void *)d_lo, (void *)d_lio,
                  (void *)d_le, (void *)d_lie})
    if (p) (void)hipFree(p);
  if (rc) {
    delete h;
    return rc;
  }
  *out = h;
  return PSC_OK;
}

// Gather device-resident graphs (local ids, as made by the device
// generators) into the union buffers of handle h: offsets rebased by the
// edge prefix, ids by the vertex prefix.  The handle's buffers are reused
// and only grow, so a serving lane re-binds the same handle batch after
// batch.
static int gather_graphs(psc_handle *h, const psc_graph *graphs, uint64_t count,
                         const uint64_t *id_base, const uint64_t *edge_base) {
  const uint32_t G = (uint32_t)count;
  std::vector<uint64_t> vb(G + 1, 0), eb(G + 1, 0);
  for (uint32_t g = 0; g < G; ++g) {
    vb[g + 1] = vb[g] + graphs[g].n;
    eb[g + 1] = eb[g] + graphs[g].m;
  }
  if (vb[G] >= 0xFFFFFFFFull) {
    psc::set_error("batch too large: total vertices must be < 2^32-1");
    return PSC_ERR_INVALID;
  }
  psc::Engine &e = h->eng;
  const uint64_t ntot = vb[G], mtot = eb[G];
  int rc = e.setup();
  if (!rc && e.arena) e.arena_reset();   // a lane carves the new binding afresh
  if (!rc) rc = e.g_off.ensure((ntot + 1) * 8);
  if (!rc) rc = e.g_ioff.ensure((ntot + 1) * 8);
  if (!rc) rc = e.g_adj.ensure(std::max<uint64_t>(mtot, 1) * 4);
  if (!rc) rc = e.g_iadj.ensure(std::max<uint64_t>(mtot, 1) * 4);
  if (rc) return rc;
  hipStream_t st = e.stream;
  hipError_t he = hipSuccess;
  for (uint32_t g = 0; g < G && he == hipSuccess; ++g) {
    const psc_graph &x = graphs[g];
    if (x.n == 0) continue;
    const uint64_t sv0 = id_base ? id_base[g] : 0, se0 = edge_base ? edge_base[g] : 0;
    k_rebase_one<<<blocks_for(std::max(x.n + 1, x.m)), 256, 0, st>>>(
        x.offset, x.edges, x.n, x.m, vb[g], eb[g], sv0, se0,
        e.g_off.as<uint64_t>() + vb[g], e.g_adj.as<uint32_t>() + eb[g]);
    k_rebase_one<<<blocks_for(std::max(x.n + 1, x.m)), 256, 0, st>>>(
        x.in_offset, x.in_edges, x.n, x.m, vb[g], eb[g], sv0, se0,
        e.g_ioff.as<uint64_t>() + vb[g], e.g_iadj.as<uint32_t>() + eb[g]);
    he = hipGetLastError();
  }
  // an empty graph owns no offset slot of its own: its neighbours' rebased
  // offsets already hold the right value; only an all-empty batch needs
  // its single offset written
  if (he == hipSuccess && ntot == 0) {
    k_fill_one<<<1, 1, 0, st>>>(e.g_off.as<uint64_t>(), 0);
    k_fill_one<<<1, 1, 0, st>>>(e.g_ioff.as<uint64_t>(), 0);
    he = hipGetLastError();
  }
  if (he == hipSuccess) he = hipStreamSynchronize(st);
  if (he != hipSuccess) {
    psc::set_error(std::string("batch gather: ") + hipGetErrorString(he));
    return PSC_ERR_HIP;
  }
  psc_graph ug{ntot, mtot, e.g_off.as<uint64_t>(), e.g_adj.as<uint32_t>(),
               e.g_ioff.as<uint64_t>(), e.g_iadj.as<uint32_t>()};
  return e.init(&ug, true, G, vb.data());
}

int psc_batch_create_from_graphs(const psc_graph *graphs, uint64_t count,
                                 const uint64_t *id_base, const uint64_t *edge_base,
                                 psc_handle **out) {
  psc::GateGuard gate;
  if (!graphs || !out || count == 0 || count > 0xFFFFFFF0ull) return PSC_ERR_INVALID;
  *out = nullptr;
  auto *h = new (std::nothrow) psc_handle;
  if (!h) return PSC_ERR_NOMEM;
  int rc = gather_graphs(h, graphs, count, id_base, edge_base);
  if (rc) {
    delete h;
    return rc;
  }
  *out = h;
  return PSC_OK;
}

int psc_batch_rebind(psc_handle *h, const psc_graph *graphs, uint64_t count,
                     const uint64_t *id_base, const uint64_t *edge_base) {
  psc::GateGuard gate;
  if (!h || !graphs || count == 0 || count > 0xFFFFFFF0ull) return PSC_ERR_INVALID;
  if (!h->eng.owned_graph && h->eng.ready && !h->eng.lane) {
    psc::set_error("psc_batch_rebind: handle borrows its graph");
    return PSC_ERR_INVALID;
  }
  return gather_graphs(h, graphs, count, id_base, edge_base);
}

// A serving lane: a re-bindable handle whose every device buffer -- the
// graphs' union copy, the working set, the tables -- is carved from one
// region of arena_bytes allocated here, so the lane holds exactly that much
// device memory whatever it runs.  A binding or run that does not fit
// fails with PSC_ERR_NOMEM and leaves the lane usable.
int psc_lane_create(uint64_t arena_bytes, psc_handle **out) {
  psc::GateGuard gate;
  if (!out || arena_bytes < (1u << 20)) return PSC_ERR_INVALID;
  *out = nullptr;
  auto *h = new (std::nothrow) psc_handle;
  if (!h) return PSC_ERR_NOMEM;
  psc::Engine &e = h->eng;
  int rc = e.setup();
  void *p = nullptr;
  if (!rc && hipMalloc(&p, arena_bytes) != hipSuccess) {
    psc::set_error("psc_lane_create: hipMalloc(" + std::to_string(arena_bytes) + " B) failed");
    rc = PSC_ERR_NOMEM;
  }
  if (rc) {
    delete h;
    return rc;
  }
  e.lane_arena.base = static_cast<char *>(p);
  e.lane_arena.cap = arena_bytes;
  e.arena = &e.lane_arena;
  e.lane = true;
  e.owned_graph = true;
  e.voff_h.assign(2, 0);   // no graph bound yet
  *out = h;
  return PSC_OK;
}

// Device bytes a handle holds: its arena, or its pool buffers, set-up
// blocks and owned graph copy; *peak gets the most of an arena ever in use
// (the held bytes for a pool handle).
int psc_handle_device_bytes(psc_handle *h, uint64_t *held, uint64_t *peak) {
  if (!h) return PSC_ERR_INVALID;
  psc::Engine &e = h->eng;
  uint64_t b = 0, pk = 0;
  if (e.arena) {
    b = e.arena->cap;
    pk = e.arena->peak;
  } else {
    for (psc::DevBuf *d : e.all_bufs())
      if (!d->carved) b += d->cap;
    b += e.setup_bytes;
    if (e.owned_graph)
      for (psc::DevBuf *d : {&e.g_off, &e.g_adj, &e.g_ioff, &e.g_iadj}) b += d->cap;
    pk = b;
  }
  if (held) *held = b;
  if (peak) *peak = pk;
  return PSC_OK;
}

// Arena bytes a lane binding of G graphs with n vertices and m edges in all
// takes before its first run: the graphs' union copy and the working set
// of init_requests, each piece 256-byte aligned as the arena carves it.
uint64_t psc_batch_layout_bytes(uint64_t n, uint64_t m, uint64_t num_graphs) {
  auto a = [](uint64_t b) { return (b + 255) & ~(uint64_t)255; };
  psc::Engine e;
  uint64_t t = 2 * a((n + 1) * 8) + 2 * a(std::max<uint64_t>(m, 1) * 4);
  for (const auto &r : e.init_requests(n, m, std::max<uint64_t>(num_graphs, 1))) t += a(r.second);
  return t;
}

// The device's stream-ordered pool, which the library's handles and device
// batches allocate from: bytes it holds from the driver (reserved) and
// bytes of live allocations (used).  Freed memory stays reserved for reuse
// until psc_pool_trim gives it back.
int psc_pool_usage(uint64_t *reserved, uint64_t *used) {
  psc::GateGuard gate;
  int dev = 0;
  hipMemPool_t pool;
  if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetDefaultMemPool(&pool, dev) != hipSuccess) {
    psc::set_error("psc_pool_usage: no default memory pool");
    return PSC_ERR_HIP;
  }
  uint64_t r = 0, u = 0;
  if (hipMemPoolGetAttribute(pool, hipMemPoolAttrReservedMemCurrent, &r) != hipSuccess ||
      hipMemPoolGetAttribute(pool, hipMemPoolAttrUsedMemCurrent, &u) != hipSuccess) {
    psc::set_error("psc_pool_usage: hipMemPoolGetAttribute failed");
    return PSC_ERR_HIP;
  }
  if (reserved) *reserved = r;
  if (used) *used = u;
  return PSC_OK;
}

// Give the pool's unused memory back to the driver, keeping keep_bytes
// reserved (frees already ordered on the library's free stream complete
// first).
int psc_pool_trim(uint64_t keep_bytes) {
  psc::GateGuard gate;
  int dev = 0;
  hipMemPool_t pool;
  if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetDefaultMemPool(&pool, dev) != hipSuccess) {
    psc::set_error("psc_pool_trim: no default memory pool");
    return PSC_ERR_HIP;
  }
  (void)hipStreamSynchronize(psc::pool_free_stream());
  if (hipMemPoolTrimTo(pool, keep_bytes) != hipSuccess) {
    psc::set_error("psc_pool_trim: hipMemPoolTrimTo failed");
    return PSC_ERR_HIP;
  }
  return PSC_OK;
}

// Free and total device memory as the driver reports them (hipMemGetInfo).
int psc_device_mem_info(uint64_t *free_bytes, uint64_t *total_bytes) {
  size_t f = 0, t = 0;
  if (hipMemGetInfo(&f, &t) != hipSuccess) {
    psc::set_error("hipMemGetInfo failed");
    return PSC_ERR_HIP;
  }
  if (free_bytes) *free_bytes = f;
  if (total_bytes) *total_bytes = t;
  return PSC_OK;
}

// Graph g of a device batch as a view into its union (global ids from
// vbase[g], offsets from ebase[g]); with id_base = vbase[g] and edge_base
// = ebase[g] it can be gathered by psc_batch_rebind.
int psc_dev_batch_view(const psc_dev_batch *db, uint64_t g, psc_graph *out,
                       uint64_t *id_base, uint64_t *edge_base) {
  if (!db || !out || g >= db->num_graphs) return PSC_ERR_INVALID;
  const uint64_t v0 = db->vbase[g], e0 = db->ebase[g];
  *out = psc_graph{db->vbase[g + 1] - v0, db->ebase[g + 1] - e0,
                   db->g.offset + v0, db->g.edges + e0, db->g.in_offset + v0,
                   db->g.in_edges + e0};
  if (id_base) *id_base = v0;
  if (edge_base) *edge_base = e0;
  return PSC_OK;
}

int psc_dev_batch_union(const psc_dev_batch *db, psc_graph *out) {
  if (!db || !out) return PSC_ERR_INVALID;
  *out = db->g;
  return PSC_OK;
}

uint64_t psc_handle_graphs(psc_handle *h) { return h ? h->eng.G : 0; }

int psc_handle_vbase(psc_handle *h, uint64_t *vbase_out) {
  if (!h || !vbase_out) return PSC_ERR_INVALID;
  const psc::Engine &e = h->eng;
  // a lane that was never bound holds no graph: one empty graph
  if (e.voff_h.size() < (size_t)e.G + 1) {
    std::memset(vbase_out, 0, ((size_t)e.G + 1) * 8);
    return PSC_OK;
  }
  std::memcpy(vbase_out, e.voff_h.data(), ((size_t)e.G + 1) * 8);
  return PSC_OK;
}

int psc_batch_run(psc_handle *h, const psc_params *p, const double *beta_per_graph,
                  uint64_t *labels_out, psc_stats *stats) {
  if (!h) return PSC_ERR_INVALID;
  psc_params def;
  psc_params_default(&def);
  if (!p) p = &def;
  psc::Engine &e = h->eng;
  e.beta_g.clear();
  if (beta_per_graph) {
    e.beta_g.assign(beta_per_graph, beta_per_graph + e.G);
    for (double b : e.beta_g)
      if (!(b >= 1.0)) {
        psc::set_error("beta must be >= 1");
        return PSC_ERR_INVALID;
      }
  }
  psc_stats tmp;
  int rc = e.run(*p, e.labels.as<unsigned long long>(), stats ? stats : &tmp);
  e.beta_g.clear();
  if (rc) return rc;
  if (labels_out && e.n) return e.download_labels(labels_out);
  return PSC_OK;
}

// Seeded description of a batch: graph g is RMAT (kind 0), a grid (1) or a
// 3D k-NN graph (2), with a vertex count log-uniform in [2^lo, 2^hi].
// RMAT counts are rounded to a power of two and grids to a square.
int psc_batch_spec(uint64_t num_graphs, double lo_log2, double hi_log2,
                   uint64_t seed, uint32_t kinds_mask, uint32_t *kind_out,
                   uint64_t *n_out) {
  if (!kind_out || !n_out || hi_log2 < lo_log2 || lo_log2 < 1 || hi_log2 > 30 ||
      (kinds_mask & 7) == 0)
    return PSC_ERR_INVALID;
  uint32_t allowed[3];
  uint32_t na = 0;
  for (uint32_t k = 0; k < 3; ++k)
    if (kinds_mask & (1u << k)) allowed[na++] = k;
  for (uint64_t g = 0; g < num_graphs; ++g) {
    const uint64_t r = psc::draw64(seed, 0x62617463680000ull, g);
    const uint32_t kind = allowed[(r >> 40) % na];
    const double u = (double)(r & 0xFFFFFFFFull) / 4294967296.0;
    double n = std::floor(std::exp2(lo_log2 + (hi_log2 - lo_log2) * u));
    if (kind == KIND_RMAT) {
      n = std::exp2(std::round(std::log2(n)));
    } else if (kind == KIND_GRID) {
      const double side = std::max(2.0, std::round(std::sqrt(n)));
      n = side * side;
    }
    kind_out[g] = kind;
    n_out[g] = (uint64_t)n;
  }
  return PSC_OK;
}

// Build a batch on the device from its description: RMAT graphs get
// 16 n samples with (a,b,c,d) = (.57,.19,.19,.05) and scrambled ids, grids
// keep each 4-neighbour pair in both directions with probability 0.6, k-NN
// graphs use k = 5.  Graph g draws from seeds[g], or mix(seed, g) when
// seeds is NULL.  Graphs are built in groups of at most 2^30 edge keys
// (generate, radix sort, de-duplicate, both CSR directions); the groups'
// CSRs are concatenated into the union.
constexpr uint32_t GEN_KNN_K = 5;
constexpr uint64_t GEN_GROUP_KEYS = 1ull << 30;

// What a batch generation builds and allocates: per graph its description
// and key count, the vertex prefix, the key count of the largest group of
// graphs (groups of at most GEN_GROUP_KEYS keys are generated one after the
// other), the total key count and the largest k-NN cell count of a group.
struct GenPlan {
  std::vector<GenGraph> gg;
  std::vector<uint64_t> vb, nkeys;
  uint64_t max_group = 0, total_keys = 0, max_cells = 0;
};

static int gen_plan(uint64_t num_graphs, const uint32_t *kind, const uint64_t *nreq,
                    uint64_t seed, const uint64_t *seeds, GenPlan *P) {
  const uint32_t G = (uint32_t)num_graphs;
  const uint32_t KNN_K = GEN_KNN_K;
  P->gg.assign(G, GenGraph{});
  P->vb.assign(G + 1, 0);
  P->nkeys.assign(G, 0);
  std::vector<GenGraph> &gg = P->gg;
  std::vector<uint64_t> &vb = P->vb, &nkeys_g = P->nkeys;
  for (uint32_t g = 0; g < G; ++g) {
    GenGraph x{};
    x.kind = kind[g];
    x.seed = seeds ? seeds[g] : psc::mix64(seed * 0x9E3779B97F4A7C15ull + g + 1);
    x.vbase = vb[g];
    uint64_t n = nreq[g];
    if (x.kind == KIND_RMAT) {
      x.scale = 1;
      while ((1ull << x.scale) < n) ++x.scale;
      n = 1ull << x.scale;
      x.nitems = 16 * n;
      nkeys_g[g] = x.nitems;
    } else if (x.kind == KIND_GRID) {
      const uint64_t side = std::max<uint64_t>(2, (uint64_t)std::llround(std::sqrt((double)n)));
      x.rows = x.cols = (uint32_t)side;
      n = side * side;
      x.nitems = x.rows * (uint64_t)(x.cols - 1) + (uint64_t)(x.rows - 1) * x.cols;
      nkeys_g[g] = 2 * x.nitems;
    } else if (x.kind == KIND_KNN) {
      if (n < KNN_K + 1) n = KNN_K + 1;
      x.nitems = n;
      x.cells1 = (uint32_t)std::max(1.0, std::floor(std::cbrt((double)n / 2.0)));
      x.cw = 2097152.0 / x.cells1;
      nkeys_g[g] = KNN_K * n;
    } else {
      psc::set_error("unknown graph kind");
      return PSC_ERR_INVALID;
    }
    vb[g + 1] = vb[g] + n;
    gg[g] = x;
  }
  if (vb[G] >= 0xFFFFFFFFull) {
    psc::set_error("batch too large: total vertices must be < 2^32-1");
    return PSC_ERR_INVALID;
  }
  uint64_t acc = 0, cells = 0;
  for (uint32_t g = 0; g < G; ++g) {
    if (acc && acc + nkeys_g[g] > GEN_GROUP_KEYS) {   // the groups of psc_dev_gen_batch
      P->max_group = std::max(P->max_group, acc);
      P->max_cells = std::max(P->max_cells, cells);
      acc = cells = 0;
    }
    acc += nkeys_g[g];
    if (gg[g].kind == KIND_KNN) cells += (uint64_t)gg[g].cells1 * gg[g].cells1 * gg[g].cells1;
  }
  P->max_group = std::max(P->max_group, acc);
  P->max_cells = std::max(P->max_cells, cells);
  for (uint32_t g = 0; g < G; ++g) P->total_keys += nkeys_g[g];
  return PSC_OK;
}

// Device bytes of psc_dev_gen_batch for the same description: *keep = what
// the returned batch holds, *peak = everything allocated at once while it
// is built (the keep part included), each allocation rounded up to the
// pool's 2 MiB granularity.
int psc_dev_gen_batch_bytes(uint64_t num_graphs, const uint32_t *kind, const uint64_t *nreq,
                            uint64_t *keep, uint64_t *peak) {
  if (!kind || !nreq || num_graphs == 0 || num_graphs > 0xFFFFFFF0ull) return PSC_ERR_INVALID;
  GenPlan P;
  const int pr = gen_plan(num_graphs, kind, nreq, 0, nullptr, &P);
  if (pr) return pr;
  const uint64_t G = num_graphs, ntot = P.vb[G], mg = std::max<uint64_t>(P.max_group, 1);
  auto r = [](uint64_t b) { return (std::max<uint64_t>(b, 8) + (2u << 20) - 1) & ~(uint64_t)((2u << 20) - 1); };
  size_t tb = 0, tb2 = 0;
  unsigned long long *nk = nullptr;
  if (rocprim::radix_sort_keys(nullptr, tb, nk, nk, mg, 0, 64, hipStream_t(0)) != hipSuccess ||
      rocprim::unique(nullptr, tb2, nk, nk, nk, mg, rocprim::equal_to<unsigned long long>(),
                      hipStream_t(0)) != hipSuccess) {
    psc::set_error("rocprim temporary size query failed");
    return PSC_ERR_HIP;
  }
  const uint64_t k = 2 * r((ntot + 1) * 8) + 2 * r(std::max<uint64_t>(P.total_keys, 1) * 4);
  const uint64_t t = 2 * r(mg * 8) + r(8) + r(G * sizeof(GenGraph)) + 2 * r((G + 1) * 8) +
                     r(std::max(tb, tb2)) + (P.max_cells ? r((P.max_cells + 1) * 8) : 0);
  if (keep) *keep = k;
  if (peak) *peak = k + t;
  return PSC_OK;
}

int psc_dev_gen_batch(uint64_t num_graphs, const uint32_t *kind,
                      const uint64_t *nreq, uint64_t seed,
                      const uint64_t *seeds, psc_dev_batch **out) {
  if (!out || !kind || !nreq || num_graphs == 0 || num_graphs > 0xFFFFFFF0ull)
    return PSC_ERR_INVALID;
  *out = nullptr;
  GenPlan P;
  {
    const int pr = gen_plan(num_graphs, kind, nreq, seed, seeds, &P);
    if (pr) return pr;
  }
  const uint32_t G = (uint32_t)num_graphs;
  const uint32_t KNN_K = GEN_KNN_K;
  const uint64_t GROUP_KEYS = GEN_GROUP_KEYS;
  std::vector<GenGraph> &gg = P.gg;
  std::vector<uint64_t> &vb = P.vb, &nkeys_g = P.nkeys;
  const uint64_t max_group = P.max_group, total_keys = P.total_keys;
  const uint64_t ntot = vb[G];
  int rc = PSC_OK;
  hipStream_t st = nullptr;
  GenGraph *d_g = nullptr;
  unsigned long long *keys = nullptr, *tmp = nullptr, *cnt = nullptr;
  uint64_t *cstart = nullptr, *off = nullptr, *ioff = nullptr;
  uint32_t *E = nullptr, *iE = nullptr;
  void *temp = nullptr;
  size_t tb = 0, tb2 = 0;
  uint64_t m_done = 0;
  psc_dev_batch *db = nullptr;
  std::vector<uint64_t> eb(G + 1, 0);
  uint64_t *d_vb = nullptr, *d_eb = nullptr;
  psc::pool_setup();
  {
    psc::GateGuard gate;
    B_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
  }
  B_HIP(pool_alloc(&keys, std::max<uint64_t>(max_group, 1) * 8, st));
  B_HIP(pool_alloc(&tmp, std::max<uint64_t>(max_group, 1) * 8, st));
  B_HIP(pool_alloc(&cnt, 8, st));
  B_HIP(pool_alloc(&d_g, G * sizeof(GenGraph), st));
  B_HIP(pool_alloc(&d_vb, (G + 1) * 8, st));
  B_HIP(pool_alloc(&d_eb, (G + 1) * 8, st));
  B_HIP(pool_alloc(&off, (ntot + 1) * 8, st));
  B_HIP(pool_alloc(&ioff, (ntot + 1) * 8, st));
  B_HIP(pool_alloc(&E, std::max<uint64_t>(total_keys, 1) * 4, st));
  B_HIP(pool_alloc(&iE, std::max<uint64_t>(total_keys, 1) * 4, st));
  B_HIP(hipMemcpyAsync(d_vb, vb.data(), (G + 1) * 8, hipMemcpyHostToDevice, st));
  B_HIP(rocprim::radix_sort_keys(nullptr, tb, keys, tmp, std::max<uint64_t>(max_group, 1),
                                 0, 64, st));
  B_HIP(rocprim::unique(nullptr, tb2, tmp, keys, cnt, std::max<uint64_t>(max_group, 1),
                        rocprim::equal_to<unsigned long long>(), st));
  tb = std::max(tb, tb2);
  B_HIP(pool_alloc(&temp, std::max<size_t>(tb, 8), st));
  for (uint32_t g0 = 0; g0 < G;) {
    // ---- one group of graphs [g0, g1)
    uint32_t g1 = g0;
    uint64_t gkeys = 0;
    while (g1 < G && (g1 == g0 || gkeys + nkeys_g[g1] <= GROUP_KEYS)) gkeys += nkeys_g[g1++];
    std::vector<GenGraph> rg, kgv;
    uint64_t items = 0, kitems = 0, ncells = 0, kb = 0;
    for (uint32_t g = g0; g < g1; ++g) {   // RMAT / grid keys first, k-NN after
      GenGraph x = gg[g];
      if (x.kind == KIND_KNN) continue;
      x.ibase = items;
      x.kbase = kb;
      items += x.nitems;
      kb += nkeys_g[g];
      rg.push_back(x);
    }
    for (uint32_t g = g0; g < g1; ++g) {
      GenGraph x = gg[g];
      if (x.kind != KIND_KNN) continue;
      x.ibase = kitems;
      x.kbase = kb;
      x.cbase = ncells;
      ncells += (uint64_t)x.cells1 * x.cells1 * x.cells1;
      kitems += x.nitems;
      kb += nkeys_g[g];
      kgv.push_back(x);
    }
    if (!rg.empty()) {
      B_HIP(hipMemcpyAsync(d_g, rg.data(), rg.size() * sizeof(GenGraph),
                           hipMemcpyHostToDevice, st));
      k_batch_keys<<<blocks_for(items), 256, 0, st>>>(
          d_g, (uint32_t)rg.size(), items, psc::prob_threshold(0.57),
          psc::prob_threshold(0.76), psc::prob_threshold(0.95),
          psc::prob_threshold(0.6), keys);
      B_HIP(hipStreamSynchronize(st));
    }
    if (!kgv.empty()) {
      B_HIP(hipMemcpyAsync(d_g, kgv.data(), kgv.size() * sizeof(GenGraph),
                           hipMemcpyHostToDevice, st));
      if (cstart) {
        psc::GateGuard gate;
        (void)hipFreeAsync(cstart, st);
      }
      cstart = nullptr;
      B_HIP(pool_alloc(&cstart, (ncells + 1) * 8, st));
      // cell keys in the k-NN key region (it holds 5 slots per point)
      unsigned long long *ck = keys + kgv[0].kbase;
      k_knn_cellkeys<<<blocks_for(kitems), 256, 0, st>>>(d_g, (uint32_t)kgv.size(),
                                                          kitems, ck);
      B_HIP(rocprim::radix_sort_keys(temp, tb, ck, tmp, kitems, 0, 64, st));
      k_cell_starts<<<blocks_for(ncells + 1), 256, 0, st>>>(tmp, kitems, ncells, cstart);
      // k-NN keys are written relative to the start of the key buffer
      k_knn_search<<<blocks_for(kitems), 256, 0, st>>>(d_g, (uint32_t)kgv.size(), kitems,
                                                       tmp, cstart, KNN_K, keys);
      B_HIP(hipStreamSynchronize(st));
    }
    B_HIP(hipGetLastError());
    // sort + unique the group's keys; DROP sorts last
    unsigned long long hcnt = 0, last = 0;
    uint64_t m = 0;
    // ids of the group are < vb[g1]: keys only use bits [0, 32 + idb)
    // (DROP keeps all of them set, so it still sorts last)
    unsigned idb = 1;
    while (idb < 32 && (1ull << idb) < vb[g1]) ++idb;
    B_HIP(rocprim::radix_sort_keys(temp, tb, keys, tmp, gkeys, 0, 32 + idb, st));
    B_HIP(rocprim::unique(temp, tb, tmp, keys, cnt, gkeys,
                          rocprim::equal_to<unsigned long long>(), st));
    B_HIP(hipMemcpyAsync(&hcnt, cnt, 8, hipMemcpyDeviceToHost, st));
    B_HIP(hipStreamSynchronize(st));
    if (hcnt) {
      B_HIP(hipMemcpyAsync(&last, keys + hcnt - 1, 8, hipMemcpyDeviceToHost, st));
      B_HIP(hipStreamSynchronize(st));
    }
    m = hcnt - ((hcnt && last == DROP) ? 1 : 0);
    const uint64_t v0 = vb[g0], gn = vb[g1] - v0;
    // out-CSR of the group: offsets by binary search over keys, rebased
    k_low32<<<blocks_for(m), 256, 0, st>>>(keys, E + m_done, m);
    k_offsets_from_keys<<<blocks_for(gn + 1), 256, 0, st>>>(keys, m, gn, off + v0,
                                                            v0, m_done);
    // in-CSR: the keys are in (src, dst) order, so a stable sort of the
    // swapped keys on the dst bits alone yields (dst, src) order
    k_swap32<<<blocks_for(m), 256, 0, st>>>(keys, tmp, m);
    if (m) B_HIP(rocprim::radix_sort_keys(temp, tb, tmp, keys, m, 32, 32 + idb, st));
    k_low32<<<blocks_for(m), 256, 0, st>>>(keys, iE + m_done, m);
    k_offsets_from_keys<<<blocks_for(gn + 1), 256, 0, st>>>(keys, m, gn, ioff + v0,
                                                            v0, m_done);
    // per-graph edge prefix of the group: off at each graph's first vertex
    k_gather_at<<<(g1 - g0 + 255) / 256, 256, 0, st>>>(off, d_vb + g0, g1 - g0,
                                                       d_eb + g0);
    B_HIP(hipGetLastError());
    B_HIP(hipMemcpyAsync(eb.data() + g0, d_eb + g0, (g1 - g0) * 8,
                         hipMemcpyDeviceToHost, st));
    B_HIP(hipStreamSynchronize(st));
    m_done += m;
    g0 = g1;
  }
  eb[G] = m_done;
  db = new psc_dev_batch;
  db->g = psc_graph{ntot, m_done, off, E, ioff, iE};
  db->num_graphs = G;
  db->vbase = vb;
  db->ebase = eb;
  db->kind.assign(kind, kind + G);
  *out = db;
  db = nullptr;
  off = ioff = nullptr;
  E = iE = nullptr;
done:
  delete db;
  if (st) (void)hipStreamSynchronize(st);
  {
    psc::GateGuard gate;
    for (void *p : {(void *)d_g, (void *)keys, (void *)tmp, (void *)cnt, (void *)cstart,
                    (void *)d_vb, (void *)d_eb, (void *)off, (void *)ioff, (void *)E,
                    (void *)iE, temp})
      if (p) (void)hipFreeAsync(p, st);
    if (st) {
      (void)hipStreamSynchronize(st);
      (void)hipStreamDestroy(st);
    }
  }
  return rc;
}

// The caller guarantees that no work using the batch is still in flight
// (every handle that read it has returned from its run / rebind).
void psc_dev_batch_free(psc_dev_batch *db) {
  psc::GateGuard gate;
  if (!db) return;
  pool_free((void *)db->g.offset);
  pool_free((void *)db->g.edges);
  pool_free((void *)db->g.in_offset);
  pool_free((void *)db->g.in_edges);
  delete db;
}

int psc_dev_batch_info(const psc_dev_batch *db, uint64_t *num_graphs,
                       uint64_t *vbase, uint64_t *ebase, uint32_t *kind) {
  if (!db) return PSC_ERR_INVALID;
  if (num_graphs) *num_graphs = db->num_graphs;
  if (vbase) std::memcpy(vbase, db->vbase.data(), db->vbase.size() * 8);
  if (ebase) std::memcpy(ebase, db->ebase.data(), db->ebase.size() * 8);
  if (kind) std::memcpy(kind, db->kind.data(), db->kind.size() * 4);
  return PSC_OK;
}

// Handle over a device batch in place (the union is borrowed).
int psc_batch_create_dev(const psc_dev_batch *db, psc_handle **out) {
  psc::GateGuard gate;
  if (!db || !out) return PSC_ERR_INVALID;
  *out = nullptr;
  auto *h = new (std::nothrow) psc_handle;
  if (!h) return PSC_ERR_NOMEM;
  int rc = h->eng.init(&db->g, false, (uint32_t)db->num_graphs, db->vbase.data());
  if (rc) {
    delete h;
    return rc;
  }
  *out = h;
  return PSC_OK;
}

// Graph g of a device batch as a stand-alone device graph with local ids
// (for the one-graph-per-launch path).
int psc_dev_batch_extract(const psc_dev_batch *db, uint64_t g, psc_dev_graph **out) {
  psc::GateGuard gate;
  if (!db || !out || g >= db->num_graphs) return PSC_ERR_INVALID;
  *out = nullptr;
  const uint64_t v0 = db->vbase[g], n = db->vbase[g + 1] - v0;
  const uint64_t e0 = db->ebase[g], m = db->ebase[g + 1] - e0;
  int rc = PSC_OK;
  uint64_t *off = nullptr, *ioff = nullptr;
  uint32_t *E = nullptr, *iE = nullptr;
  B_HIP(hipMalloc(&off, (n + 1) * 8));
  B_HIP(hipMalloc(&ioff, (n + 1) * 8));
  B_HIP(hipMalloc(&E, std::max<uint64_t>(m, 1) * 4));
  B_HIP(hipMalloc(&iE, std::max<uint64_t>(m, 1) * 4));
  k_extract<<<blocks_for(std::max(n + 1, m)), 256>>>(db->g.offset, db->g.edges, v0, n,
                                                     e0, m, off, E);
  k_extract<<<blocks_for(std::max(n + 1, m)), 256>>>(db->g.in_offset, db->g.in_edges,
                                                     v0, n, e0, m, ioff, iE);
  B_HIP(hipGetLastError());
  B_HIP(hipDeviceSynchronize());
  {
    auto *dg = new psc_dev_graph;
    dg->g = psc_graph{n, m, off, E, ioff, iE};
    *out = dg;
    off = ioff = nullptr;
    E = iE = nullptr;
  }
done:
  for (void *p : {(void *)off, (void *)ioff, (void *)E, (void *)iE})
    if (p) (void)hipFree(p);
  return rc;
}

}  // extern "C"
