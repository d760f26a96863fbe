/*
 * ORACLE — test infrastructure only.  Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load this code, and only as the
 * checker.  The product path never links it.
 *
 * Iterative Tarjan SCC, canonicaliser and -status summary.
 *
 *   psc_oracle_tarjan      follows baselines/tarjan_scc/tarjan_scc.hpp:39-75
 *                          (dfs :39-60, tarjan :72-75).  The recursion of the
 *                          reference is replaced by an explicit stack of
 *                          (vertex, next-edge) frames so grids with ~n-deep
 *                          DFS paths do not need `ulimit -s unlimited`
 *                          (scripts/Experiment1.py:51).  Labels are the SCC
 *                          completion order 1..cnt, as the reference's
 *                          scc[] (tarjan_scc.hpp:53-57).
 *   psc_oracle_canonical   maps any labelling to "minimum vertex id of the
 *                          component"; the SCC partition is unique, so the
 *                          canonical labelling is a function of the graph
 *                          alone (src/scc.hpp:475-476 only promises
 *                          equal <=> same SCC).
 *   psc_oracle_status      n_scc and largest SCC, as SCC_status in
 *                          src/scc.cpp:13-29 / tarjan_scc.cpp:10-20.
 *
 * Pinned by: networkx / scipy cross-checks on the committed fixtures,
 * brute-force transitive closure on n <= 64 and closed-form families
 * (tests/test_oracle.py).
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* returns number of SCCs, or -1 on allocation failure.
 * comp[v] in 1..cnt, the order in which Tarjan completes components. */
int64_t psc_oracle_tarjan(uint64_t n, const uint64_t *off, const uint32_t *adj,
                          uint32_t *comp) {
  if (n == 0) return 0;
  uint32_t *dfn = (uint32_t *)calloc(n, sizeof(uint32_t));
  uint32_t *low = (uint32_t *)malloc(n * sizeof(uint32_t));
  uint32_t *stk = (uint32_t *)malloc(n * sizeof(uint32_t));   /* Tarjan stack */
  uint32_t *cs_v = (uint32_t *)malloc(n * sizeof(uint32_t));  /* call stack */
  uint64_t *cs_e = (uint64_t *)malloc(n * sizeof(uint64_t));
  if (!dfn || !low || !stk || !cs_v || !cs_e) {
    free(dfn); free(low); free(stk); free(cs_v); free(cs_e);
    return -1;
  }
  memset(comp, 0, n * sizeof(uint32_t));
  uint64_t sp = 0, idx = 0;
  int64_t cnt = 0;
  for (uint64_t r = 0; r < n; r++) {
    if (dfn[r]) continue;
    /* enter r: dfs(u) prologue, tarjan_scc.hpp:40-41 */
    uint64_t top = 0;
    cs_v[0] = (uint32_t)r;
    cs_e[0] = off[r];
    dfn[r] = low[r] = (uint32_t)++idx;
    stk[sp++] = (uint32_t)r;
    while (1) {
      uint32_t u = cs_v[top];
      uint64_t e = cs_e[top];
      if (e < off[u + 1]) {
        uint32_t v = adj[e];
        cs_e[top] = e + 1;
        if (!dfn[v]) {               /* tree edge: recurse (:44-46) */
          dfn[v] = low[v] = (uint32_t)++idx;
          stk[sp++] = v;
          ++top;
          cs_v[top] = v;
          cs_e[top] = off[v];
        } else if (!comp[v]) {       /* back/cross edge inside stack (:48-49) */
          if (dfn[v] < low[u]) low[u] = dfn[v];
        }
        continue;
      }
      /* all edges of u done: pop component if u is a root (:52-58) */
      if (low[u] == dfn[u]) {
        cnt++;
        while (1) {
          uint32_t x = stk[--sp];
          comp[x] = (uint32_t)cnt;
          if (x == u) break;
        }
      }
      if (top == 0) break;
      --top;
      /* return to parent: low[p] = min(low[p], low[u]) (:47) */
      uint32_t p = cs_v[top];
      if (low[u] < low[p]) low[p] = low[u];
    }
  }
  free(dfn); free(low); free(stk); free(cs_v); free(cs_e);
  return cnt;
}

/* Canonicalise labels: out[v] = min { w : labels[w] == labels[v] }.
 * Labels are arbitrary 64-bit values; a small open-addressing map keyed by
 * label keeps this O(n).  Returns 0, or -1 on allocation failure. */
int psc_oracle_canonical(uint64_t n, const uint64_t *labels, uint32_t *out) {
  if (n == 0) return 0;
  uint64_t cap = 1;
  while (cap < 2 * n) cap <<= 1;
  uint64_t *keys = (uint64_t *)malloc(cap * sizeof(uint64_t));
  uint32_t *vals = (uint32_t *)malloc(cap * sizeof(uint32_t));
  uint8_t *used = (uint8_t *)calloc(cap, 1);
  if (!keys || !vals || !used) {
    free(keys); free(vals); free(used);
    return -1;
  }
  for (uint64_t v = 0; v < n; v++) {
    uint64_t k = labels[v];
    uint64_t h = k * 0x9E3779B97F4A7C15ull;
    h ^= h >> 29;
    h &= cap - 1;
    while (used[h] && keys[h] != k) h = (h + 1) & (cap - 1);
    if (!used[h]) {          /* first (= smallest) vertex with this label */
      used[h] = 1;
      keys[h] = k;
      vals[h] = (uint32_t)v;
    }
    out[v] = vals[h];
  }
  free(keys); free(vals); free(used);
  return 0;
}

/* -status summary of a canonical labelling (src/scc.cpp:13-29):
 * number of SCCs and size of the largest. */
int psc_oracle_status(uint64_t n, const uint32_t *canon, uint64_t *n_scc,
                      uint64_t *largest) {
  uint32_t *cnt = (uint32_t *)calloc(n ? n : 1, sizeof(uint32_t));
  if (!cnt) return -1;
  uint64_t ns = 0, big = 0;
  for (uint64_t v = 0; v < n; v++) {
    if (cnt[canon[v]]++ == 0) ns++;
    if (cnt[canon[v]] > big) big = cnt[canon[v]];
  }
  free(cnt);
  *n_scc = ns;
  *largest = big;
  return 0;
}
