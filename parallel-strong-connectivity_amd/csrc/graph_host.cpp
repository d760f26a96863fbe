// Host graph I/O and seeded generators (the L3 layer the reference keeps on
// the host, src/graph.hpp).  GBBS .bin layout (src/graph.hpp:120-185):
//   uint64 n, uint64 m, uint64 sizes = (n+1)*8 + m*4 + 3*8,
//   uint64 offset[n+1], uint32 edges[m]
// and the -large layout (src/graph.hpp:279-315) stores the out-CSR block
// followed by an in-CSR block of the same shape.
#include <omp.h>
#include <parallel/algorithm>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "../../include/psc.h"
#include "gen_common.hpp"

namespace psc {
void set_error(const std::string &msg);
}

namespace {

psc_host_graph *alloc_graph(uint64_t n, uint64_t m) {
  auto *hg = (psc_host_graph *)calloc(1, sizeof(psc_host_graph));
  if (!hg) return nullptr;
  hg->offset = (uint64_t *)malloc((n + 1) * 8);
  hg->in_offset = (uint64_t *)malloc((n + 1) * 8);
  hg->edges = (uint32_t *)malloc(std::max<uint64_t>(m, 1) * 4);
  hg->in_edges = (uint32_t *)malloc(std::max<uint64_t>(m, 1) * 4);
  if (!hg->offset || !hg->in_offset || !hg->edges || !hg->in_edges) {
    psc_host_graph_free(hg);
    return nullptr;
  }
  hg->g.n = n;
  hg->g.m = m;
  hg->g.offset = hg->offset;
  hg->g.edges = hg->edges;
  hg->g.in_offset = hg->in_offset;
  hg->g.in_edges = hg->in_edges;
  return hg;
}

// In-CSR by a stable counting transpose: in-lists come out sorted by
// source, as make_directed's sort of (v, u) pairs (src/graph.hpp:222-245).
void transpose(uint64_t n, uint64_t m, const uint64_t *off, const uint32_t *E,
               uint64_t *ioff, uint32_t *iE) {
  std::vector<uint64_t> cnt(n + 1, 0);
  for (uint64_t i = 0; i < m; ++i) cnt[E[i] + 1]++;
  for (uint64_t v = 0; v < n; ++v) cnt[v + 1] += cnt[v];
  memcpy(ioff, cnt.data(), (n + 1) * 8);
  for (uint64_t u = 0; u < n; ++u)
    for (uint64_t j = off[u]; j < off[u + 1]; ++j) iE[cnt[E[j]]++] = (uint32_t)u;
}

// CSR from (u << 32 | v) keys: sort, drop duplicates and self-loops.
psc_host_graph *csr_from_keys(uint64_t n, std::vector<uint64_t> &keys) {
  __gnu_parallel::sort(keys.begin(), keys.end());
  uint64_t w = 0;
  for (uint64_t i = 0; i < keys.size(); ++i) {
    const uint64_t k = keys[i];
    if ((k >> 32) == (k & 0xFFFFFFFFull)) continue;
    if (w && keys[w - 1] == k) continue;
    keys[w++] = k;
  }
  const uint64_t m = w;
  psc_host_graph *hg = alloc_graph(n, m);
  if (!hg) return nullptr;
  std::vector<uint64_t> cnt(n + 1, 0);
  for (uint64_t i = 0; i < m; ++i) {
    cnt[(keys[i] >> 32) + 1]++;
    hg->edges[i] = (uint32_t)keys[i];
  }
  for (uint64_t v = 0; v < n; ++v) cnt[v + 1] += cnt[v];
  memcpy(hg->offset, cnt.data(), (n + 1) * 8);
  transpose(n, m, hg->offset, hg->edges, hg->in_offset, hg->in_edges);
  return hg;
}

bool read_exact(FILE *f, void *p, size_t bytes) {
  return bytes == 0 || fread(p, 1, bytes, f) == bytes;
}
bool write_exact(FILE *f, const void *p, size_t bytes) {
  return bytes == 0 || fwrite(p, 1, bytes, f) == bytes;
}

}  // namespace

extern "C" {

void psc_host_graph_free(psc_host_graph *hg) {
  if (!hg) return;
  free(hg->offset);
  free(hg->edges);
  free(hg->in_offset);
  free(hg->in_edges);
  free(hg);
}

int psc_make_directed(uint64_t n, uint64_t m, const uint64_t *offset,
                      const uint32_t *edges, psc_host_graph **out) {
  if (!out || (n && !offset) || (m && !edges)) return PSC_ERR_INVALID;
  *out = nullptr;
  if (offset[0] != 0 || offset[n] != m) {
    psc::set_error("offsets must start at 0 and end at m");
    return PSC_ERR_INVALID;
  }
  for (uint64_t i = 0; i < m; ++i)
    if (edges[i] >= n) {
      psc::set_error("edge endpoint out of range");
      return PSC_ERR_INVALID;
    }
  psc_host_graph *hg = alloc_graph(n, m);
  if (!hg) return PSC_ERR_NOMEM;
  memcpy(hg->offset, offset, (n + 1) * 8);
  if (m) memcpy(hg->edges, edges, m * 4);
  transpose(n, m, hg->offset, hg->edges, hg->in_offset, hg->in_edges);
  *out = hg;
  return PSC_OK;
}

int psc_read_bin(const char *path, int large, psc_host_graph **out) {
  if (!path || !out) return PSC_ERR_INVALID;
  *out = nullptr;
  FILE *f = fopen(path, "rb");
  if (!f) {
    psc::set_error(std::string("cannot open ") + path);
    return PSC_ERR_IO;
  }
  uint64_t hdr[3];
  if (!read_exact(f, hdr, sizeof(hdr))) {
    fclose(f);
    psc::set_error("truncated header");
    return PSC_ERR_IO;
  }
  const uint64_t n = hdr[0], m = hdr[1];
  if (hdr[2] != (n + 1) * 8 + m * 4 + 3 * 8 || n >= 0xFFFFFFFFull) {
    fclose(f);
    psc::set_error("bad .bin header (sizes field mismatch)");
    return PSC_ERR_IO;
  }
  psc_host_graph *hg = alloc_graph(n, m);
  if (!hg) {
    fclose(f);
    return PSC_ERR_NOMEM;
  }
  bool ok = read_exact(f, hg->offset, (n + 1) * 8) && read_exact(f, hg->edges, m * 4);
  if (ok && large) {
    uint64_t h2[3];
    ok = read_exact(f, h2, sizeof(h2)) && h2[0] == n && h2[1] == m &&
         read_exact(f, hg->in_offset, (n + 1) * 8) &&
         read_exact(f, hg->in_edges, m * 4);
  }
  if (ok) ok = fgetc(f) == EOF;   // "Bad data" check, src/graph.hpp:178-181
  fclose(f);
  if (ok) ok = hg->offset[0] == 0 && hg->offset[n] == m;
  if (ok)
    for (uint64_t i = 0; i < m && ok; ++i) ok = hg->edges[i] < n;
  if (ok && large) {
    ok = hg->in_offset[0] == 0 && hg->in_offset[n] == m;
    for (uint64_t i = 0; i < m && ok; ++i) ok = hg->in_edges[i] < n;
  }
  if (!ok) {
    psc_host_graph_free(hg);
    psc::set_error("malformed .bin graph");
    return PSC_ERR_IO;
  }
  if (!large) {
    if (strstr(path, "sym")) {   // symmetric input, src/graph.hpp:265-268
      memcpy(hg->in_offset, hg->offset, (n + 1) * 8);
      if (m) memcpy(hg->in_edges, hg->edges, m * 4);
    } else {
      transpose(n, m, hg->offset, hg->edges, hg->in_offset, hg->in_edges);
    }
  }
  *out = hg;
  return PSC_OK;
}

int psc_read_bin_device(const char *path, int large, psc_dev_graph **out) {
  if (!path || !out) return PSC_ERR_INVALID;
  *out = nullptr;
  if (large || strstr(path, "sym")) {
    // both CSRs come from the file: read, check and upload them
    psc_host_graph *hg = nullptr;
    int rc = psc_read_bin(path, large, &hg);
    if (rc) return rc;
    const uint64_t n = hg->g.n, m = hg->g.m;
    void *p[4] = {nullptr, nullptr, nullptr, nullptr};
    const void *src[4] = {hg->offset, hg->edges, hg->in_offset, hg->in_edges};
    const size_t bytes[4] = {(n + 1) * 8, m * 4, (n + 1) * 8, m * 4};
    for (int i = 0; i < 4 && rc == PSC_OK; ++i) {
      rc = psc_device_alloc(bytes[i] ? bytes[i] : 4, &p[i]);
      if (rc == PSC_OK && bytes[i]) rc = psc_device_copy(p[i], src[i], bytes[i]);
    }
    psc_host_graph_free(hg);
    if (rc) {
      for (void *q : p)
        if (q) psc_device_free(q);
      return rc;
    }
    auto *dg = new psc_dev_graph;
    dg->g = psc_graph{n, m, (const uint64_t *)p[0], (const uint32_t *)p[1],
                      (const uint64_t *)p[2], (const uint32_t *)p[3]};
    *out = dg;
    return PSC_OK;
  }
  // one CSR in the file: read it, upload it, transpose on the device
  FILE *f = fopen(path, "rb");
  if (!f) {
    psc::set_error(std::string("cannot open ") + path);
    return PSC_ERR_IO;
  }
  uint64_t hdr[3];
  if (!read_exact(f, hdr, sizeof(hdr))) {
    fclose(f);
    psc::set_error("truncated header");
    return PSC_ERR_IO;
  }
  const uint64_t n = hdr[0], m = hdr[1];
  if (hdr[2] != (n + 1) * 8 + m * 4 + 3 * 8 || n >= 0xFFFFFFFFull) {
    fclose(f);
    psc::set_error("bad .bin header (sizes field mismatch)");
    return PSC_ERR_IO;
  }
  std::vector<uint64_t> off(n + 1);
  std::vector<uint32_t> E(m);
  bool ok = read_exact(f, off.data(), (n + 1) * 8) && read_exact(f, E.data(), m * 4);
  if (ok) ok = fgetc(f) == EOF;   // "Bad data" check, src/graph.hpp:178-181
  fclose(f);
  if (!ok) {
    psc::set_error("malformed .bin graph");
    return PSC_ERR_IO;
  }
  void *d_off = nullptr, *d_e = nullptr;
  int rc = psc_device_alloc((n + 1) * 8, &d_off);
  if (rc == PSC_OK) rc = psc_device_alloc(m ? m * 4 : 4, &d_e);
  if (rc == PSC_OK) rc = psc_device_copy(d_off, off.data(), (n + 1) * 8);
  if (rc == PSC_OK && m) rc = psc_device_copy(d_e, E.data(), m * 4);
  // offsets and endpoints are checked on the device
  if (rc == PSC_OK)
    rc = psc_dev_make_directed(n, m, (const uint64_t *)d_off, (const uint32_t *)d_e, out);
  if (d_off) psc_device_free(d_off);
  if (d_e) psc_device_free(d_e);
  return rc;
}

int psc_write_bin(const char *path, const psc_graph *g, int large) {
  if (!path || !g) return PSC_ERR_INVALID;
  FILE *f = fopen(path, "wb");
  if (!f) {
    psc::set_error(std::string("cannot create ") + path);
    return PSC_ERR_IO;
  }
  const uint64_t hdr[3] = {g->n, g->m, (g->n + 1) * 8 + g->m * 4 + 3 * 8};
  bool ok = write_exact(f, hdr, sizeof(hdr)) &&
            write_exact(f, g->offset, (g->n + 1) * 8) &&
            write_exact(f, g->edges, g->m * 4);
  if (ok && large)
    ok = write_exact(f, hdr, sizeof(hdr)) &&
         write_exact(f, g->in_offset, (g->n + 1) * 8) &&
         write_exact(f, g->in_edges, g->m * 4);
  ok = (fclose(f) == 0) && ok;
  if (!ok) {
    psc::set_error(std::string("write failed: ") + path);
    return PSC_ERR_IO;
  }
  return PSC_OK;
}

int psc_gen_rmat(uint32_t scale, uint64_t samples, double a, double b,
                 double c, uint64_t seed, int permute, psc_host_graph **out) {
  if (!out || scale < 1 || scale > 31 || a < 0 || b < 0 || c < 0 ||
      a + b + c > 1.0) {
    psc::set_error("psc_gen_rmat: bad arguments");
    return PSC_ERR_INVALID;
  }
  *out = nullptr;
  const uint64_t n = 1ull << scale;
  const uint32_t t0 = psc::prob_threshold(a), t1 = psc::prob_threshold(a + b),
                 t2 = psc::prob_threshold(a + b + c);
  const psc::IdPerm perm((int)scale, seed);
  std::vector<uint64_t> keys(samples);
#pragma omp parallel for schedule(static)
  for (uint64_t i = 0; i < samples; ++i) {
    uint32_t u, v;
    psc::rmat_sample(seed, i, (int)scale, t0, t1, t2, &u, &v);
    if (permute) {
      u = (uint32_t)perm(u);
      v = (uint32_t)perm(v);
    }
    keys[i] = ((uint64_t)u << 32) | v;
  }
  psc_host_graph *hg = csr_from_keys(n, keys);
  if (!hg) return PSC_ERR_NOMEM;
  *out = hg;
  return PSC_OK;
}

int psc_gen_grid(uint32_t rows, uint32_t cols, double p_both, uint64_t seed,
                 psc_host_graph **out) {
  if (!out || rows == 0 || cols == 0 ||
      (uint64_t)rows * cols >= 0xFFFFFFFFull) {
    psc::set_error("psc_gen_grid: bad arguments");
    return PSC_ERR_INVALID;
  }
  *out = nullptr;
  const uint64_t n = (uint64_t)rows * cols;
  const uint64_t nh = (uint64_t)rows * (cols - 1);
  const uint64_t npairs = nh + (uint64_t)(rows - 1) * cols;
  const uint32_t thr = psc::prob_threshold(p_both);
  std::vector<uint64_t> keys(2 * npairs);
#pragma omp parallel for schedule(static)
  for (uint64_t p = 0; p < npairs; ++p) {
    uint64_t x, y;   // pair endpoints, x < y
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
    const uint64_t none = (x << 32) | x;   // self-loop marker, dropped later
    keys[2 * p] = (kind == 2) ? none : ((x << 32) | y);
    keys[2 * p + 1] = (kind == 1) ? none : ((y << 32) | x);
  }
  psc_host_graph *hg = csr_from_keys(n, keys);
  if (!hg) return PSC_ERR_NOMEM;
  *out = hg;
  return PSC_OK;
}

int psc_gen_knn3d(uint64_t n, uint32_t k, uint64_t seed, psc_host_graph **out) {
  if (!out || n < 2 || k == 0 || k >= n || n >= 0xFFFFFFFFull || k > 64) {
    psc::set_error("psc_gen_knn3d: bad arguments");
    return PSC_ERR_INVALID;
  }
  *out = nullptr;
  std::vector<uint32_t> X(n), Y(n), Z(n);
#pragma omp parallel for schedule(static)
  for (uint64_t i = 0; i < n; ++i) {
    X[i] = psc::knn_coord(seed, i, 0);
    Y[i] = psc::knn_coord(seed, i, 1);
    Z[i] = psc::knn_coord(seed, i, 2);
  }
  // uniform cell grid, ~2 points per cell
  uint32_t G = (uint32_t)std::max(1.0, std::floor(std::cbrt((double)n / 2.0)));
  const uint64_t cells = (uint64_t)G * G * G;
  const double cw = 2097152.0 / G;   // 2^21 / G
  auto cell_of = [&](uint32_t x) {
    uint32_t c = (uint32_t)(x / cw);
    return c >= G ? G - 1 : c;
  };
  std::vector<uint64_t> cstart(cells + 1, 0);
  std::vector<uint32_t> cid(n);
  for (uint64_t i = 0; i < n; ++i) {
    const uint64_t c = ((uint64_t)cell_of(X[i]) * G + cell_of(Y[i])) * G + cell_of(Z[i]);
    cid[i] = (uint32_t)c;
    cstart[c + 1]++;
  }
  for (uint64_t c = 0; c < cells; ++c) cstart[c + 1] += cstart[c];
  std::vector<uint32_t> pts(n);
  {
    std::vector<uint64_t> fill(cstart.begin(), cstart.end() - 1);
    for (uint64_t i = 0; i < n; ++i) pts[fill[cid[i]]++] = (uint32_t)i;
  }
  std::vector<uint64_t> keys(n * k);
#pragma omp parallel
  {
    std::vector<std::pair<uint64_t, uint32_t>> best;
    best.reserve(k + 1);
#pragma omp for schedule(dynamic, 256)
    for (uint64_t i = 0; i < n; ++i) {
      best.clear();
      const int cx = (int)cell_of(X[i]), cy = (int)cell_of(Y[i]), cz = (int)cell_of(Z[i]);
      for (int r = 0;; ++r) {
        // visit the shell of Chebyshev radius r around (cx, cy, cz)
        for (int dx = -r; dx <= r; ++dx)
          for (int dy = -r; dy <= r; ++dy)
            for (int dz = -r; dz <= r; ++dz) {
              if (std::max(std::abs(dx), std::max(std::abs(dy), std::abs(dz))) != r) continue;
              const int x = cx + dx, y = cy + dy, z = cz + dz;
              if (x < 0 || y < 0 || z < 0 || x >= (int)G || y >= (int)G || z >= (int)G) continue;
              const uint64_t c = ((uint64_t)x * G + y) * G + z;
              for (uint64_t t = cstart[c]; t < cstart[c + 1]; ++t) {
                const uint32_t j = pts[t];
                if (j == i) continue;
                const int64_t ddx = (int64_t)X[i] - X[j], ddy = (int64_t)Y[i] - Y[j],
                              ddz = (int64_t)Z[i] - Z[j];
                const uint64_t d = (uint64_t)(ddx * ddx + ddy * ddy + ddz * ddz);
                const std::pair<uint64_t, uint32_t> cand(d, j);
                if (best.size() < k) {
                  best.insert(std::upper_bound(best.begin(), best.end(), cand), cand);
                } else if (cand < best.back()) {
                  best.pop_back();
                  best.insert(std::upper_bound(best.begin(), best.end(), cand), cand);
                }
              }
            }
        // every point outside the searched cube is farther than r*cw from i
        if (best.size() == k) {
          const double reach = (double)r * cw;
          if (reach * reach > (double)best.back().first) break;
        }
        if (r > (int)G) break;
      }
      for (uint32_t t = 0; t < k; ++t)
        keys[i * k + t] = ((uint64_t)i << 32) | best[t].second;
    }
  }
  psc_host_graph *hg = csr_from_keys(n, keys);
  if (!hg) return PSC_ERR_NOMEM;
  *out = hg;
  return PSC_OK;
}

}  // extern "C"
