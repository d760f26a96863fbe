// Connected components of the undirected graph under an out-CSR (every arc
// joins its endpoints): the partition Connectivity computes
// (src/Connectivity.h:37-119) through LDD (src/ldd.hpp) and union-find
// (src/union_find_rules.h).
//
// On the GPU the union-find runs over the arcs directly, lock-free: a root
// is hooked only under a smaller root, by one CAS on its own parent slot,
// and finds halve their paths.  Parent pointers therefore only ever point
// to smaller ids (no cycles, every pass ends), and a component's last root
// is its minimum vertex id: the labels are the minimum vertex id of each
// component whatever order the CASes land in.  The reference needs LDD to
// keep its CPU union-find short; here the arcs go to the union-find at once,
// after Afforest's sampling (Sutton et al., IPDPS 2018): every vertex first
// links CC_SAMPLE out-arcs, the most frequent root among CC_PROBES spread
// vertices is taken as the giant component, and -- when the graph carries
// its in-CSR -- vertices already in it skip their remaining arcs, which
// their other ends link from the in-CSR or the out-CSR.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>

#include "../../include/psc.h"
#include "device_common.hpp"
#include "engine.hpp"

namespace psc {
namespace {

constexpr uint32_t CC_HEAVY = 256;   // longer lists are shared by a wave

__global__ void k_cc_init(uint32_t *p, uint64_t n) {
  for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < n;
       v += (uint64_t)gridDim.x * blockDim.x)
    p[v] = (uint32_t)v;
}

// root of x; every visited slot is re-pointed at its grandparent (path
// halving: the new parent is an ancestor, so parents keep decreasing)
__device__ __forceinline__ uint32_t cc_find(uint32_t *p, uint32_t x) {
  uint32_t px = __atomic_load_n(&p[x], __ATOMIC_RELAXED);
  while (px != x) {
    const uint32_t gp = __atomic_load_n(&p[px], __ATOMIC_RELAXED);
    if (gp != px) __atomic_store_n(&p[x], gp, __ATOMIC_RELAXED);
    x = px;
    px = gp;
  }
  return x;
}

// union of the components of u and v: the larger root goes under the
// smaller one if it is still a root, else both roots are found again.
// With a forest buffer, the arc (u, v) of every hook is recorded: each
// hook joins two sets, so the recorded arcs form a spanning forest
// (Connectivity's spanning_forest, src/Connectivity.h:100-118).
struct CcForest {
  uint32_t *arcs;             // 2 per hook, at most n - 1 hooks
  unsigned long long *count;
};

__device__ __forceinline__ void cc_unite(uint32_t *p, uint32_t u0, uint32_t v0,
                                         CcForest f) {
  uint32_t u = u0, v = v0;
  while (true) {
    u = cc_find(p, u);
    v = cc_find(p, v);
    if (u == v) return;
    const uint32_t hi = u > v ? u : v, lo = u > v ? v : u;
    if (atomicCAS(&p[hi], hi, lo) == hi) {
      if (f.arcs) {
        const unsigned long long at = atomicAdd(f.count, 1ull);
        f.arcs[2 * at] = u0;
        f.arcs[2 * at + 1] = v0;
      }
      return;
    }
    u = hi;
    v = lo;
  }
}

constexpr uint32_t CC_SAMPLE = 2;      // out-arcs per vertex linked first
constexpr uint32_t CC_PROBES = 1024;   // vertices whose roots elect the giant component
constexpr uint32_t CC_NONE = 0xFFFFFFFFu;

// sampling (Afforest): every vertex links its first CC_SAMPLE out-arcs
__global__ void k_cc_sample(const uint64_t *__restrict__ off, const uint32_t *__restrict__ adj,
                            uint64_t n, uint32_t *p, CcForest f) {
  for (uint64_t u = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; u < n;
       u += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t b = off[u], e = off[u + 1];
    for (uint64_t j = b; j < e && j < b + CC_SAMPLE; ++j) cc_unite(p, (uint32_t)u, adj[j], f);
  }
}

// roots of CC_PROBES spread vertices, for the host to elect the most
// frequent one
__global__ void k_cc_probe(uint32_t *p, uint64_t n, uint32_t *roots) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < CC_PROBES) roots[i] = cc_find(p, (uint32_t)(hash32(i) % n));
}

// The remaining arcs, a thread per vertex for lists up to CC_HEAVY arcs
// (longer ones are listed for k_cc_heavy): out-arcs after the sampled ones
// and, with an in-CSR, every in-arc.  A vertex whose root is `giant` (its
// set holds the elected vertex) is skipped: an arc with both ends in that
// set joins nothing new, and any other arc is also an in- or out-arc of
// its end outside the set, which links it.  Skipping only ever happens
// for a vertex already joined to the giant root, so it is exact whatever
// the order of hooks.
__device__ __forceinline__ bool cc_skip(uint32_t *p, uint32_t u, uint32_t giant) {
  return giant != CC_NONE && cc_find(p, u) == giant;
}

__global__ __launch_bounds__(BLOCK) void k_cc_light(const uint64_t *__restrict__ off,
                                                    const uint32_t *__restrict__ adj,
                                                    const uint64_t *__restrict__ ioff,
                                                    const uint32_t *__restrict__ iadj,
                                                    uint64_t n, uint32_t skip0, uint32_t giant,
                                                    uint32_t *p, uint32_t *heavy,
                                                    uint32_t *nheavy, CcForest f) {
  for (uint64_t u = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; u < n;
       u += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t e = off[u + 1], b = off[u] + skip0 < e ? off[u] + skip0 : e;
    const uint64_t ib = ioff ? ioff[u] : 0, ie = ioff ? ioff[u + 1] : 0;
    const bool any = e > b || ie > ib;
    const bool run = any && !cc_skip(p, (uint32_t)u, giant);
    const bool big = run && (e - b) + (ie - ib) > CC_HEAVY;
    const uint32_t at = wave_append(nheavy, big);
    if (big) {
      heavy[at] = (uint32_t)u;
      continue;
    }
    if (!run) continue;
    for (uint64_t j = b; j < e; ++j) cc_unite(p, (uint32_t)u, adj[j], f);
    for (uint64_t j = ib; j < ie; ++j) cc_unite(p, iadj[j], (uint32_t)u, f);   // arc iadj[j] -> u
  }
}

// one wave per listed vertex, lanes over its remaining arcs
__global__ __launch_bounds__(BLOCK) void k_cc_heavy(const uint64_t *__restrict__ off,
                                                    const uint32_t *__restrict__ adj,
                                                    const uint64_t *__restrict__ ioff,
                                                    const uint32_t *__restrict__ iadj,
                                                    uint32_t skip0, const uint32_t *heavy,
                                                    const uint32_t *nheavy, uint32_t *p,
                                                    CcForest f) {
  const uint32_t nh = *nheavy;
  const uint64_t waves = (uint64_t)gridDim.x * (blockDim.x / WAVE);
  for (uint64_t h = (blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) / WAVE; h < nh;
       h += waves) {
    const uint32_t u = heavy[h];
    const uint64_t e = off[u + 1], b = off[u] + skip0 < e ? off[u] + skip0 : e;
    for (uint64_t j = b + lane(); j < e; j += WAVE) cc_unite(p, u, adj[j], f);
    if (ioff)
      for (uint64_t j = ioff[u] + lane(); j < ioff[u + 1]; j += WAVE) cc_unite(p, iadj[j], u, f);
  }
}

// *bad = 1 if an offset falls or leaves [0, m], or an endpoint is >= n
__global__ void k_cc_check(const uint64_t *off, const uint32_t *adj, uint64_t n, uint64_t m,
                           uint32_t *bad) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  const uint64_t i0 = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  for (uint64_t v = i0; v < n; v += stride)
    if (off[v + 1] < off[v] || off[v + 1] > m) *bad = 1;
  if (i0 == 0 && (off[0] != 0 || off[n] != m)) *bad = 1;
  for (uint64_t j = i0; j < m; j += stride)
    if (adj[j] >= n) *bad = 1;
}

// Final labels: every slot is written only by its own thread, with a root
// walk that writes nothing (a halving store from another thread could
// otherwise replace a written root by an older ancestor).  No hook happens
// in this kernel, so every walk ends at the same root.
__global__ void k_cc_labels(uint32_t *p, uint64_t n) {
  for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < n;
       v += (uint64_t)gridDim.x * blockDim.x) {
    uint32_t x = (uint32_t)v, px = __atomic_load_n(&p[x], __ATOMIC_RELAXED);
    while (px != x) {
      x = px;
      px = __atomic_load_n(&p[x], __ATOMIC_RELAXED);
    }
    __atomic_store_n(&p[v], x, __ATOMIC_RELAXED);
  }
}

unsigned grid_for_cc(uint64_t n) {
  return (unsigned)std::max<uint64_t>(1, std::min<uint64_t>((n + BLOCK - 1) / BLOCK, 65536));
}

#define CC_HIP(call)                                                      \
  do {                                                                    \
    hipError_t e_ = (call);                                               \
    if (e_ != hipSuccess) {                                               \
      set_error(std::string(#call) + ": " + hipGetErrorString(e_));       \
      rc = PSC_ERR_HIP;                                                   \
      goto done;                                                          \
    }                                                                     \
  } while (0)

}  // namespace
}  // namespace psc

extern "C" {

static int cc_device(const psc_graph *dev_g, uint32_t *d_labels, uint32_t *d_forest,
                     uint64_t *forest_count, double *ms) {
  using namespace psc;
  if (!dev_g || (dev_g->n && (!d_labels || !dev_g->offset)) ||
      (dev_g->m && !dev_g->edges) || dev_g->n >= 0xFFFFFFFFull) {
    set_error("psc_connectivity_device: bad arguments");
    return PSC_ERR_INVALID;
  }
  const uint64_t n = dev_g->n;
  int rc = PSC_OK;
  hipStream_t st = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  uint32_t *heavy = nullptr, *nheavy = nullptr, *roots = nullptr;
  unsigned long long *fcount = nullptr, hfc = 0;
  CcForest f{nullptr, nullptr};
  uint32_t hbad = 0, giant = CC_NONE;
  uint32_t hroots[CC_PROBES];
  float t = 0;
  // with the in-CSR (the transpose of the out-CSR) the giant component's
  // vertices can be skipped; without it every arc is linked
  const bool has_in = dev_g->m && dev_g->in_offset && dev_g->in_edges;
  if (forest_count) *forest_count = 0;
  if (n == 0) {
    if (ms) *ms = 0;
    return PSC_OK;
  }
  CC_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
  if (d_forest) {
    CC_HIP(hipMalloc(&fcount, 8));
    CC_HIP(hipMemsetAsync(fcount, 0, 8, st));
    f = CcForest{d_forest, fcount};
  }
  CC_HIP(hipEventCreate(&e0));
  CC_HIP(hipEventCreate(&e1));
  CC_HIP(hipMalloc(&heavy, n * 4));
  CC_HIP(hipMalloc(&nheavy, 4));
  CC_HIP(hipMalloc(&roots, CC_PROBES * 4));
  // the input is checked before any kernel indexes with it
  k_cc_init<<<1, 64, 0, st>>>(nheavy, 1);   // used as the flag: 0
  k_cc_check<<<grid_for_cc(std::max<uint64_t>(n, dev_g->m)), BLOCK, 0, st>>>(
      dev_g->offset, dev_g->edges, n, dev_g->m, nheavy);
  if (has_in)
    k_cc_check<<<grid_for_cc(std::max<uint64_t>(n, dev_g->m)), BLOCK, 0, st>>>(
        dev_g->in_offset, dev_g->in_edges, n, dev_g->m, nheavy);
  CC_HIP(hipGetLastError());
  CC_HIP(hipMemcpyAsync(&hbad, nheavy, 4, hipMemcpyDeviceToHost, st));
  CC_HIP(hipStreamSynchronize(st));
  if (hbad) {
    set_error("psc_connectivity: offsets must run from 0 to m without falling, "
              "endpoints must be < n");
    rc = PSC_ERR_INVALID;
    goto done;
  }
  CC_HIP(hipEventRecord(e0, st));
  k_cc_init<<<grid_for_cc(n), BLOCK, 0, st>>>(d_labels, n);
  k_cc_init<<<1, 64, 0, st>>>(nheavy, 1);   // nheavy = 0
  k_cc_sample<<<grid_for_cc(n), BLOCK, 0, st>>>(dev_g->offset, dev_g->edges, n, d_labels, f);
  if (has_in) {   // elect the most frequent root of the probes
    k_cc_probe<<<(CC_PROBES + BLOCK - 1) / BLOCK, BLOCK, 0, st>>>(d_labels, n, roots);
    CC_HIP(hipGetLastError());
    CC_HIP(hipMemcpyAsync(hroots, roots, sizeof hroots, hipMemcpyDeviceToHost, st));
    CC_HIP(hipStreamSynchronize(st));
    std::sort(hroots, hroots + CC_PROBES);
    uint32_t best = 0;
    for (uint32_t i = 0, j; i < CC_PROBES; i = j) {
      for (j = i; j < CC_PROBES && hroots[j] == hroots[i]; ++j) {}
      if (j - i > best) {
        best = j - i;
        giant = hroots[i];
      }
    }
  }
  k_cc_light<<<grid_for_cc(n), BLOCK, 0, st>>>(
      dev_g->offset, dev_g->edges, has_in ? dev_g->in_offset : nullptr,
      has_in ? dev_g->in_edges : nullptr, n, CC_SAMPLE, giant, d_labels, heavy, nheavy, f);
  k_cc_heavy<<<grid_for_cc(n * WAVE / 16), BLOCK, 0, st>>>(
      dev_g->offset, dev_g->edges, has_in ? dev_g->in_offset : nullptr,
      has_in ? dev_g->in_edges : nullptr, CC_SAMPLE, heavy, nheavy, d_labels, f);
  k_cc_labels<<<grid_for_cc(n), BLOCK, 0, st>>>(d_labels, n);
  CC_HIP(hipGetLastError());
  CC_HIP(hipEventRecord(e1, st));
  CC_HIP(hipStreamSynchronize(st));
  CC_HIP(hipEventElapsedTime(&t, e0, e1));
  if (ms) *ms = t;
  if (fcount) {
    CC_HIP(hipMemcpy(&hfc, fcount, 8, hipMemcpyDeviceToHost));
    *forest_count = hfc;
  }
done:
  if (fcount) (void)hipFree(fcount);
  if (heavy) (void)hipFree(heavy);
  if (nheavy) (void)hipFree(nheavy);
  if (roots) (void)hipFree(roots);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (st) (void)hipStreamDestroy(st);
  return rc;
}

int psc_connectivity_device(const psc_graph *dev_g, uint32_t *d_labels, double *ms) {
  psc::GateGuard gate;
  return cc_device(dev_g, d_labels, nullptr, nullptr, ms);
}

int psc_spanning_forest_device(const psc_graph *dev_g, uint32_t *d_labels, uint32_t *d_forest,
                               uint64_t *count, double *ms) {
  psc::GateGuard gate;
  if (!count || (dev_g && dev_g->n > 1 && !d_forest)) {
    psc::set_error("psc_spanning_forest_device: bad arguments");
    return PSC_ERR_INVALID;
  }
  return cc_device(dev_g, d_labels, dev_g && dev_g->n > 1 ? d_forest : nullptr, count, ms);
}

static int cc_host(const psc_graph *g, uint32_t *labels, uint32_t *forest, uint64_t *count) {
  using namespace psc;
  if (!g || (g->n && (!labels || !g->offset)) || (g->m && !g->edges) ||
      (forest && !count)) {
    set_error("psc_connectivity: bad arguments");
    return PSC_ERR_INVALID;
  }
  const uint64_t n = g->n, m = g->m;
  int rc = PSC_OK;
  uint64_t *off = nullptr, *ioff = nullptr;
  uint32_t *adj = nullptr, *iadj = nullptr, *lab = nullptr, *fst = nullptr;
  psc_graph dg{};
  const bool want_forest = forest && g->n > 1;
  if (count) *count = 0;
  const bool has_in = m && g->in_offset && g->in_edges;
  if (n == 0) return PSC_OK;
  CC_HIP(hipMalloc(&off, (n + 1) * 8));
  CC_HIP(hipMalloc(&adj, std::max<uint64_t>(m, 1) * 4));
  CC_HIP(hipMalloc(&lab, n * 4));
  CC_HIP(hipMemcpy(off, g->offset, (n + 1) * 8, hipMemcpyHostToDevice));
  if (m) CC_HIP(hipMemcpy(adj, g->edges, m * 4, hipMemcpyHostToDevice));
  if (has_in) {
    CC_HIP(hipMalloc(&ioff, (n + 1) * 8));
    CC_HIP(hipMalloc(&iadj, m * 4));
    CC_HIP(hipMemcpy(ioff, g->in_offset, (n + 1) * 8, hipMemcpyHostToDevice));
    CC_HIP(hipMemcpy(iadj, g->in_edges, m * 4, hipMemcpyHostToDevice));
  }
  if (want_forest) CC_HIP(hipMalloc(&fst, (n - 1) * 8));
  dg = psc_graph{n, m, off, adj, ioff, iadj};
  rc = cc_device(&dg, lab, fst, count, nullptr);
  if (rc == PSC_OK) CC_HIP(hipMemcpy(labels, lab, n * 4, hipMemcpyDeviceToHost));
  if (rc == PSC_OK && want_forest && *count)
    CC_HIP(hipMemcpy(forest, fst, *count * 8, hipMemcpyDeviceToHost));
done:
  if (off) (void)hipFree(off);
  if (adj) (void)hipFree(adj);
  if (ioff) (void)hipFree(ioff);
  if (iadj) (void)hipFree(iadj);
  if (lab) (void)hipFree(lab);
  if (fst) (void)hipFree(fst);
  return rc;
}

int psc_connectivity(const psc_graph *g, uint32_t *labels) {
  psc::GateGuard gate;
  return cc_host(g, labels, nullptr, nullptr);
}

int psc_spanning_forest(const psc_graph *g, uint32_t *labels, uint32_t *forest,
                        uint64_t *count) {
  psc::GateGuard gate;
  if (!count || (g && g->n > 1 && !forest)) {
    psc::set_error("psc_spanning_forest: bad arguments");
    return PSC_ERR_INVALID;
  }
  return cc_host(g, labels, forest, count);
}

}  // extern "C"
