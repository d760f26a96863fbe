/* Oracle (test infrastructure only): least-element lists, the sequential
 * algorithm of src/LeList.hpp:76-115 (LE_LIST::seq_le_list).
 *
 * For the sources in the given order (the reference's `centers`), a BFS
 * over the out-arcs from each source enters a vertex only when its distance
 * from this source is below dist[], the smallest distance from any earlier
 * source (dist is kept across sources, :77-79,94-101).  Every vertex it
 * enters gives an entry (target, priority, distance).  The entries are
 * sorted by (target, priority) and the priority is replaced by the source
 * vertex (:103-114).
 *
 * The reference sorts with std::sort; here a stable counting sort by
 * target is enough, because the entries of one target are appended in
 * increasing priority (sources are processed in order).  Pinned in
 * tests/test_oracle.py against unpruned BFS from every source followed by
 * the LE-list definition, and against closed forms on paths. */
#include <stdint.h>
#include <stdlib.h>

#define LE_INF 0xFFFFFFFFu

/* Returns the number of entries; *out receives a malloc'ed array of
 * 3 * count uint32 (target, source vertex, distance), freed with
 * psc_oracle_free.  Returns (uint64_t)-1 on allocation failure or if
 * order is not a permutation of 0..n-1. */
uint64_t psc_oracle_le_lists(uint64_t n, const uint64_t *off, const uint32_t *adj,
                             const uint32_t *order, uint32_t **out) {
  *out = NULL;
  if (n == 0) return 0;
  uint32_t *dist = malloc(n * sizeof *dist);
  uint32_t *queue = malloc(n * sizeof *queue);
  uint8_t *seen = calloc(n, 1);
  uint64_t cap = 4 * n, cnt = 0;
  uint32_t *ent = malloc(cap * 3 * sizeof *ent); /* (target, priority, distance) */
  uint32_t *res = NULL;
  uint64_t *start = NULL;
  uint64_t ret = (uint64_t)-1;
  if (!dist || !queue || !seen || !ent) goto done;
  for (uint64_t i = 0; i < n; ++i) {
    if (order[i] >= n || seen[order[i]]) goto done;
    seen[order[i]] = 1;
    dist[i] = LE_INF;
  }
  for (uint64_t i = 0; i < n; ++i) {
    const uint32_t s = order[i];
    uint64_t head = 0, tail = 0;
    dist[s] = 0;
    queue[tail++] = s;
    while (head < tail) {
      const uint32_t u = queue[head++];
      if (cnt == cap) {
        uint32_t *t = realloc(ent, cap * 2 * 3 * sizeof *ent);
        if (!t) goto done;
        ent = t;
        cap *= 2;
      }
      ent[3 * cnt] = u;
      ent[3 * cnt + 1] = (uint32_t)i;
      ent[3 * cnt + 2] = dist[u];
      ++cnt;
      /* a vertex enters the queue once per source: its first discovery
       * sets dist to the BFS distance, later ones cannot beat it */
      for (uint64_t j = off[u]; j < off[u + 1]; ++j) {
        const uint32_t w = adj[j];
        if (dist[u] + 1 < dist[w]) {
          dist[w] = dist[u] + 1;
          queue[tail++] = w;
        }
      }
    }
  }
  /* stable counting sort by target */
  start = calloc(n + 1, sizeof *start);
  res = malloc((cnt ? cnt : 1) * 3 * sizeof *res);
  if (!start || !res) goto done;
  for (uint64_t e = 0; e < cnt; ++e) start[ent[3 * e] + 1]++;
  for (uint64_t v = 0; v < n; ++v) start[v + 1] += start[v];
  for (uint64_t e = 0; e < cnt; ++e) {
    const uint64_t o = start[ent[3 * e]]++;
    res[3 * o] = ent[3 * e];
    res[3 * o + 1] = order[ent[3 * e + 1]];
    res[3 * o + 2] = ent[3 * e + 2];
  }
  *out = res;
  res = NULL;
  ret = cnt;
done:
  free(dist);
  free(queue);
  free(seen);
  free(ent);
  free(res);
  free(start);
  return ret;
}

void psc_oracle_free(void *p) { free(p); }
