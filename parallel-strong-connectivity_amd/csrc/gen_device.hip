// Device graph builders: the seeded RMAT and grid generators of
// graph_host.cpp, run on the GPU so multi-GB inputs are produced directly in
// HBM.  Edge keys (u << 32 | v) are generated with the same counter-based
// draws as the host (gen_common.hpp), radix-sorted, de-duplicated, and both
// CSR directions are built by binary search over the sorted keys, so the
// result is bit-identical to the host generator.
#include <hip/hip_runtime.h>

#include <cstring>
#include <string>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>
#include <rocprim/device/device_select.hpp>

#include "../../include/psc.h"
#include "engine.hpp"
#include "gen_common.hpp"

namespace psc {
void set_error(const std::string &msg);
}

namespace {

constexpr unsigned long long DROP = ~0ull;   // self-loop / absent edge

#define GEN_HIP(call)                                                    \
  do {                                                                   \
    hipError_t e_ = (call);                                              \
    if (e_ != hipSuccess) {                                              \
      psc::set_error(std::string(#call) + ": " + hipGetErrorString(e_)); \
      rc = PSC_ERR_HIP;                                                  \
      goto done;                                                         \
    }                                                                    \
  } while (0)

__global__ void k_rmat_keys(unsigned long long *keys, uint64_t samples,
                            uint64_t seed, int scale, uint32_t t0, uint32_t t1,
                            uint32_t t2, int permute) {
  const psc::IdPerm perm(scale, seed);
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
       i < samples; i += (uint64_t)gridDim.x * blockDim.x) {
    uint32_t u, v;
    psc::rmat_sample(seed, i, scale, t0, t1, t2, &u, &v);
    if (permute) {
      u = (uint32_t)perm(u);
      v = (uint32_t)perm(v);
    }
    keys[i] = (u == v) ? DROP : (((unsigned long long)u << 32) | v);
  }
}

__global__ void k_grid_keys(unsigned long long *keys, uint32_t rows,
                            uint32_t cols, uint64_t seed, uint32_t thr) {
  const uint64_t nh = (uint64_t)rows * (cols - 1);
  const uint64_t npairs = nh + (uint64_t)(rows - 1) * cols;
  for (uint64_t p = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; p < npairs;
       p += (uint64_t)gridDim.x * blockDim.x) {
    uint64_t x, y;
    if (p < nh) {
      const uint64_t r = p / (cols - 1), c = p % (cols - 1);
      x = r * cols + c;
      y = x + 1;
    } else {
      const uint64_t q = p - nh, r = q / cols, c = q % cols;
      x = r * cols + c;
      y = x + cols;
    }
    const int kind = psc::grid_pair_kind(seed, p, thr);
    keys[2 * p] = (kind == 2) ? DROP : ((x << 32) | y);
    keys[2 * p + 1] = (kind == 1) ? DROP : ((y << 32) | x);
  }
}

__global__ void k_swap_halves(const unsigned long long *in,
                              unsigned long long *out, uint64_t m) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < m;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const unsigned long long k = in[i];
    out[i] = (k << 32) | (k >> 32);
  }
}

__global__ void k_low_halves(const unsigned long long *keys, uint32_t *E,
                             uint64_t m) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < m;
       i += (uint64_t)gridDim.x * blockDim.x)
    E[i] = (uint32_t)keys[i];
}

// off[v] = first index with key >= v << 32, off[n] = m
__global__ void k_offsets(const unsigned long long *keys, uint64_t m,
                          uint64_t n, uint64_t *off) {
  for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v <= n;
       v += (uint64_t)gridDim.x * blockDim.x) {
    const unsigned long long t = (unsigned long long)v << 32;
    uint64_t lo = 0, hi = m;
    while (lo < hi) {
      const uint64_t mid = (lo + hi) >> 1;
      if (keys[mid] < t) lo = mid + 1; else hi = mid;
    }
    off[v] = lo;
  }
}

// make_directed on the device.  A vertex with out-edges marks the first
// slot of its list with its id (every other slot holds 0); a max-scan over
// the slots then gives each edge its source.  Offsets that fall or leave
// [0, m] set *bad.
__global__ void k_src_marks(const uint64_t *off, uint64_t n, uint64_t m, uint32_t *mark,
                            unsigned long long *bad) {
  for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < n;
       v += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t b = off[v], e = off[v + 1];
    if (e < b || e > m) {
      *bad = 1;
      continue;
    }
    if (e > b) mark[b] = (uint32_t)v;
  }
}

// keys[i] = dst << 32 | src: sorted, they list every in-edge by target and,
// within a target, by source, as make_directed's sort of (v, u) pairs does.
// Endpoints >= n set *bad.
__global__ void k_dst_src_keys(const uint32_t *E, const uint32_t *src, uint64_t m,
                               uint64_t n, unsigned long long *keys,
                               unsigned long long *bad) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < m;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t d = E[i];
    if (d >= n) *bad = 1;
    keys[i] = ((unsigned long long)d << 32) | src[i];
  }
}

int grid_blocks(uint64_t n) {
  uint64_t g = (n + 255) / 256;
  if (g < 1) g = 1;
  if (g > 65536) g = 65536;
  return (int)g;
}

// Sort + unique `keys` (len entries, DROP marks holes), then build both CSR
// directions of an n-vertex graph into *out.
int build_csr(unsigned long long *keys, uint64_t len, uint64_t n,
              psc_dev_graph **out) {
  int rc = PSC_OK;
  hipStream_t st = nullptr;
  unsigned long long *tmpk = nullptr, *cnt = nullptr;
  void *temp = nullptr;
  size_t temp_bytes = 0, tb2 = 0;
  uint64_t m = 0;
  unsigned long long hcnt = 0, last = 0;
  psc_dev_graph *dg = nullptr;
  uint64_t *off = nullptr, *ioff = nullptr;
  uint32_t *E = nullptr, *iE = nullptr;
  GEN_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
  GEN_HIP(hipMalloc(&tmpk, std::max<uint64_t>(len, 1) * 8));
  GEN_HIP(hipMalloc(&cnt, 8));
  GEN_HIP(rocprim::radix_sort_keys(nullptr, temp_bytes, keys, tmpk, len, 0, 64, st));
  GEN_HIP(rocprim::unique(nullptr, tb2, tmpk, keys, cnt, len,
                          rocprim::equal_to<unsigned long long>(), st));
  temp_bytes = std::max(temp_bytes, tb2);
  GEN_HIP(hipMalloc(&temp, std::max<size_t>(temp_bytes, 8)));
  if (len) {
    GEN_HIP(rocprim::radix_sort_keys(temp, temp_bytes, keys, tmpk, len, 0, 64, st));
    GEN_HIP(rocprim::unique(temp, temp_bytes, tmpk, keys, cnt, len,
                            rocprim::equal_to<unsigned long long>(), st));
    GEN_HIP(hipMemcpyAsync(&hcnt, cnt, 8, hipMemcpyDeviceToHost, st));
    GEN_HIP(hipStreamSynchronize(st));
    if (hcnt) {
      GEN_HIP(hipMemcpyAsync(&last, keys + hcnt - 1, 8, hipMemcpyDeviceToHost, st));
      GEN_HIP(hipStreamSynchronize(st));
    }
    m = hcnt - ((hcnt && last == DROP) ? 1 : 0);
  }
  GEN_HIP(hipMalloc(&off, (n + 1) * 8));
  GEN_HIP(hipMalloc(&ioff, (n + 1) * 8));
  GEN_HIP(hipMalloc(&E, std::max<uint64_t>(m, 1) * 4));
  GEN_HIP(hipMalloc(&iE, std::max<uint64_t>(m, 1) * 4));
  k_low_halves<<<grid_blocks(m), 256, 0, st>>>(keys, E, m);
  k_offsets<<<grid_blocks(n + 1), 256, 0, st>>>(keys, m, n, off);
  k_swap_halves<<<grid_blocks(m), 256, 0, st>>>(keys, tmpk, m);
  if (m) GEN_HIP(rocprim::radix_sort_keys(temp, temp_bytes, tmpk, keys, m, 0, 64, st));
  k_low_halves<<<grid_blocks(m), 256, 0, st>>>(keys, iE, m);
  k_offsets<<<grid_blocks(n + 1), 256, 0, st>>>(keys, m, n, ioff);
  GEN_HIP(hipGetLastError());
  GEN_HIP(hipStreamSynchronize(st));
  dg = new psc_dev_graph;
  dg->g = psc_graph{n, m, off, E, ioff, iE};
  *out = dg;
  off = ioff = nullptr;
  E = iE = nullptr;
done:
  if (temp) (void)hipFree(temp);
  if (tmpk) (void)hipFree(tmpk);
  if (cnt) (void)hipFree(cnt);
  if (off) (void)hipFree(off);
  if (ioff) (void)hipFree(ioff);
  if (E) (void)hipFree(E);
  if (iE) (void)hipFree(iE);
  if (st) (void)hipStreamDestroy(st);
  return rc;
}

}  // namespace

extern "C" {

int psc_dev_gen_rmat(uint32_t scale, uint64_t samples, double a, double b,
                     double c, uint64_t seed, int permute, psc_dev_graph **out) {
  psc::GateGuard gate;
  if (!out || scale < 1 || scale > 31 || a < 0 || b < 0 || c < 0 ||
      a + b + c > 1.0) {
    psc::set_error("psc_dev_gen_rmat: bad arguments");
    return PSC_ERR_INVALID;
  }
  *out = nullptr;
  int rc = PSC_OK;
  unsigned long long *keys = nullptr;
  GEN_HIP(hipMalloc(&keys, std::max<uint64_t>(samples, 1) * 8));
  k_rmat_keys<<<grid_blocks(samples), 256>>>(
      keys, samples, seed, (int)scale, psc::prob_threshold(a),
      psc::prob_threshold(a + b), psc::prob_threshold(a + b + c), permute);
  GEN_HIP(hipGetLastError());
  GEN_HIP(hipDeviceSynchronize());
  rc = build_csr(keys, samples, 1ull << scale, out);
done:
  if (keys) (void)hipFree(keys);
  return rc;
}

int psc_dev_gen_grid(uint32_t rows, uint32_t cols, double p_both,
                     uint64_t seed, psc_dev_graph **out) {
  psc::GateGuard gate;
  if (!out || rows == 0 || cols == 0 ||
      (uint64_t)rows * cols >= 0xFFFFFFFFull) {
    psc::set_error("psc_dev_gen_grid: bad arguments");
    return PSC_ERR_INVALID;
  }
  *out = nullptr;
  int rc = PSC_OK;
  const uint64_t npairs = (uint64_t)rows * (cols - 1) + (uint64_t)(rows - 1) * cols;
  unsigned long long *keys = nullptr;
  GEN_HIP(hipMalloc(&keys, std::max<uint64_t>(2 * npairs, 1) * 8));
  if (npairs)
    k_grid_keys<<<grid_blocks(npairs), 256>>>(keys, rows, cols, seed,
                                              psc::prob_threshold(p_both));
  GEN_HIP(hipGetLastError());
  GEN_HIP(hipDeviceSynchronize());
  rc = build_csr(keys, 2 * npairs, (uint64_t)rows * cols, out);
done:
  if (keys) (void)hipFree(keys);
  return rc;
}

int psc_dev_make_directed(uint64_t n, uint64_t m, const uint64_t *d_offset,
                          const uint32_t *d_edges, psc_dev_graph **out) {
  psc::GateGuard gate;
  if (!out || !d_offset || (m && !d_edges) || n >= 0xFFFFFFFFull) {
    psc::set_error("psc_dev_make_directed: bad arguments");
    return PSC_ERR_INVALID;
  }
  *out = nullptr;
  int rc = PSC_OK;
  hipStream_t st = nullptr;
  uint64_t *off = nullptr, *ioff = nullptr;
  uint32_t *E = nullptr, *iE = nullptr, *mark = nullptr, *src = nullptr;
  unsigned long long *keys = nullptr, *tmpk = nullptr, *bad = nullptr;
  void *temp = nullptr;
  size_t tb = 0, tb2 = 0;
  uint64_t ends[2] = {0, 0};
  unsigned long long hbad = 0;
  psc_dev_graph *dg = nullptr;
  GEN_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
  GEN_HIP(hipMalloc(&off, (n + 1) * 8));
  GEN_HIP(hipMalloc(&ioff, (n + 1) * 8));
  GEN_HIP(hipMalloc(&E, std::max<uint64_t>(m, 1) * 4));
  GEN_HIP(hipMalloc(&iE, std::max<uint64_t>(m, 1) * 4));
  GEN_HIP(hipMalloc(&bad, 8));
  GEN_HIP(hipMemcpyAsync(off, d_offset, (n + 1) * 8, hipMemcpyDeviceToDevice, st));
  if (m) GEN_HIP(hipMemcpyAsync(E, d_edges, m * 4, hipMemcpyDeviceToDevice, st));
  GEN_HIP(hipMemcpyAsync(&ends[0], off, 8, hipMemcpyDeviceToHost, st));
  GEN_HIP(hipMemcpyAsync(&ends[1], off + n, 8, hipMemcpyDeviceToHost, st));
  GEN_HIP(hipMemsetAsync(bad, 0, 8, st));
  GEN_HIP(hipStreamSynchronize(st));
  if (ends[0] != 0 || ends[1] != m) {
    psc::set_error("offsets must start at 0 and end at m");
    rc = PSC_ERR_INVALID;
    goto done;
  }
  if (m) {
    GEN_HIP(hipMalloc(&mark, m * 4));
    GEN_HIP(hipMalloc(&src, m * 4));
    GEN_HIP(hipMalloc(&keys, m * 8));
    GEN_HIP(hipMalloc(&tmpk, m * 8));
    GEN_HIP(rocprim::inclusive_scan(nullptr, tb, mark, src, m, rocprim::maximum<uint32_t>(), st));
    GEN_HIP(rocprim::radix_sort_keys(nullptr, tb2, keys, tmpk, m, 0, 64, st));
    tb = std::max(tb, tb2);
    GEN_HIP(hipMalloc(&temp, std::max<size_t>(tb, 8)));
    GEN_HIP(hipMemsetAsync(mark, 0, m * 4, st));
    k_src_marks<<<grid_blocks(n), 256, 0, st>>>(off, n, m, mark, bad);
    GEN_HIP(hipGetLastError());
    tb2 = tb;
    GEN_HIP(rocprim::inclusive_scan(temp, tb2, mark, src, m, rocprim::maximum<uint32_t>(), st));
    k_dst_src_keys<<<grid_blocks(m), 256, 0, st>>>(E, src, m, n, keys, bad);
    GEN_HIP(hipGetLastError());
    GEN_HIP(hipMemcpyAsync(&hbad, bad, 8, hipMemcpyDeviceToHost, st));
    GEN_HIP(hipStreamSynchronize(st));
    if (hbad) {
      psc::set_error("edge endpoint out of range or offsets not increasing");
      rc = PSC_ERR_INVALID;
      goto done;
    }
    tb2 = tb;
    GEN_HIP(rocprim::radix_sort_keys(temp, tb2, keys, tmpk, m, 0, 64, st));
    k_low_halves<<<grid_blocks(m), 256, 0, st>>>(tmpk, iE, m);
  }
  k_offsets<<<grid_blocks(n + 1), 256, 0, st>>>(tmpk, m, n, ioff);
  GEN_HIP(hipGetLastError());
  GEN_HIP(hipStreamSynchronize(st));
  dg = new psc_dev_graph;
  dg->g = psc_graph{n, m, off, E, ioff, iE};
  *out = dg;
  off = ioff = nullptr;
  E = iE = nullptr;
done:
  if (temp) (void)hipFree(temp);
  if (mark) (void)hipFree(mark);
  if (src) (void)hipFree(src);
  if (keys) (void)hipFree(keys);
  if (tmpk) (void)hipFree(tmpk);
  if (bad) (void)hipFree(bad);
  if (off) (void)hipFree(off);
  if (ioff) (void)hipFree(ioff);
  if (E) (void)hipFree(E);
  if (iE) (void)hipFree(iE);
  if (st) (void)hipStreamDestroy(st);
  return rc;
}

void psc_dev_graph_free(psc_dev_graph *dg) {
  psc::GateGuard gate;
  if (!dg) return;
  (void)hipFree((void *)dg->g.offset);
  (void)hipFree((void *)dg->g.edges);
  (void)hipFree((void *)dg->g.in_offset);
  (void)hipFree((void *)dg->g.in_edges);
  delete dg;
}

int psc_dev_graph_download(const psc_dev_graph *dg, psc_host_graph **out) {
  psc::GateGuard gate;
  if (!dg || !out) return PSC_ERR_INVALID;
  *out = nullptr;
  const uint64_t n = dg->g.n, m = dg->g.m;
  auto *hg = (psc_host_graph *)calloc(1, sizeof(psc_host_graph));
  if (!hg) return PSC_ERR_NOMEM;
  hg->offset = (uint64_t *)malloc((n + 1) * 8);
  hg->in_offset = (uint64_t *)malloc((n + 1) * 8);
  hg->edges = (uint32_t *)malloc(std::max<uint64_t>(m, 1) * 4);
  hg->in_edges = (uint32_t *)malloc(std::max<uint64_t>(m, 1) * 4);
  if (!hg->offset || !hg->in_offset || !hg->edges || !hg->in_edges) {
    psc_host_graph_free(hg);
    return PSC_ERR_NOMEM;
  }
  hipError_t e = hipMemcpy(hg->offset, dg->g.offset, (n + 1) * 8, hipMemcpyDeviceToHost);
  if (e == hipSuccess)
    e = hipMemcpy(hg->in_offset, dg->g.in_offset, (n + 1) * 8, hipMemcpyDeviceToHost);
  if (e == hipSuccess && m)
    e = hipMemcpy(hg->edges, dg->g.edges, m * 4, hipMemcpyDeviceToHost);
  if (e == hipSuccess && m)
    e = hipMemcpy(hg->in_edges, dg->g.in_edges, m * 4, hipMemcpyDeviceToHost);
  if (e != hipSuccess) {
    psc::set_error(std::string("download: ") + hipGetErrorString(e));
    psc_host_graph_free(hg);
    return PSC_ERR_HIP;
  }
  hg->g = psc_graph{n, m, hg->offset, hg->edges, hg->in_offset, hg->in_edges};
  *out = hg;
  return PSC_OK;
}

}  // extern "C"
