/* ORACLE (test infrastructure only; never linked into the product).
 *
 *   psc_oracle_cc   connected components of the undirected graph underlying
 *                   an out-CSR (every arc joins its endpoints), the
 *                   partition that Connectivity (src/Connectivity.h:37-119)
 *                   computes through LDD (src/ldd.hpp) and union-find
 *                   (src/union_find_rules.h).  The reference's label values
 *                   depend on the order of its concurrent CAS operations;
 *                   the partition does not, so the oracle returns each
 *                   vertex's minimum vertex id in its component.
 *                   Sequential union-find: union by smaller root id, full
 *                   path compression.  Pinned in tests/test_oracle.py
 *                   against breadth-first search on small graphs and
 *                   scipy.sparse.csgraph.connected_components.
 */
#include <stdint.h>
#include <stdlib.h>

static uint32_t find_root(uint32_t *p, uint32_t x) {
  uint32_t r = x;
  while (p[r] != r) r = p[r];
  while (p[x] != r) { /* full path compression */
    const uint32_t next = p[x];
    p[x] = r;
    x = next;
  }
  return r;
}

int psc_oracle_cc(uint64_t n, const uint64_t *off, const uint32_t *adj, uint32_t *label) {
  if (n >= 0xFFFFFFFFull) return -1;
  for (uint64_t v = 0; v < n; ++v) label[v] = (uint32_t)v;
  for (uint64_t u = 0; u < n; ++u)
    for (uint64_t j = off[u]; j < off[u + 1]; ++j) {
      const uint32_t a = find_root(label, (uint32_t)u), b = find_root(label, adj[j]);
      if (a == b) continue;
      if (a < b) label[b] = a; else label[a] = b; /* smaller id stays root */
    }
  for (uint64_t v = 0; v < n; ++v) label[v] = find_root(label, (uint32_t)v);
  return 0;
}
