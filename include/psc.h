/*
 * psc.h — C ABI of the MI355X strongly-connected-components path.
 *
 * Plain pointers and sizes only; no C++ or torch types cross this boundary.
 * Every entry point returns PSC_OK (0) or a negative PSC_ERR_* code and never
 * aborts; psc_last_error() gives the message of the last failure on the
 * calling thread.
 *
 * Reference interfaces replaced (paths are into ucrparlay/
 * Parallel-Strong-Connectivity):
 *   struct Graph                       src/graph.hpp:35-43      -> psc_graph
 *   SCC::SCC(Graph&)                   src/scc.hpp:81-92        -> psc_create
 *   SCC::scc(labels, beta, local_reach, local_scc)
 *                                      src/scc.hpp:79,277-479   -> psc_run
 *   globals tau (src/reach.hpp:19), _min_bag_size_ (src/scc.hpp:24), set by
 *   src/scc.cpp:45-46                                           -> psc_params
 *   SCC::breakdown_report              src/scc.hpp:106-114      -> psc_stats
 *   SCC_status                         src/scc.cpp:13-29        -> psc_status
 *   read_graph / read_large_graph      src/graph.hpp:248-315    -> psc_read_bin
 *   make_directed                      src/graph.hpp:222-245    -> psc_make_directed
 */
#ifndef PSC_H
#define PSC_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PSC_OK 0
#define PSC_ERR_INVALID (-1) /* bad argument or malformed graph          */
#define PSC_ERR_HIP (-2)     /* a HIP runtime call failed                */
#define PSC_ERR_NOMEM (-3)   /* device or host allocation failed         */
#define PSC_ERR_IO (-4)      /* file could not be read or written        */
#define PSC_ERR_BUDGET (-5)  /* request exceeds the serving HBM budget   */

/* Directed graph in CSR form, both directions (src/graph.hpp:35-43).
 * offset/in_offset hold n+1 entries, edges/in_edges hold m entries.
 * Vertex ids are < n < 2^32-1 (UINT_N_MAX is the empty key,
 * src/utilities.h:26).  Duplicate edges and self-loops are tolerated. */
typedef struct psc_graph {
  uint64_t n;
  uint64_t m;
  const uint64_t *offset;
  const uint32_t *edges;
  const uint64_t *in_offset;
  const uint32_t *in_edges;
} psc_graph;

/* Run parameters.  beta, local_reach, local_scc mirror the SCC::scc
 * arguments; tau and min_bag the globals of src/scc.cpp:45-46; dense_div
 * the divisor of the sparse/dense switch threshold m/20 (src/reach.hpp:49);
 * seed drives the pivot permutation that replaces parlay::random_shuffle
 * (src/scc.hpp:303); table_init, when non-zero, overrides the initial
 * label-table slot count of src/scc.hpp:398-400,412-414 (used to force the
 * overflow-and-double path of src/scc.hpp:256-275). */
typedef struct psc_params {
  double beta;          /* pivot batch growth, default 1.5          */
  int32_t local_reach;  /* VGC in the first-SCC reach (-local_reach) */
  int32_t local_scc;    /* VGC in multi-search (-local_scc)          */
  uint64_t tau;         /* VGC local budget per seed, default 512    */
  uint64_t min_bag;     /* -lambda, default 1024                     */
  uint64_t dense_div;   /* default 20                                */
  uint64_t seed;        /* pivot permutation seed, default 1         */
  uint64_t table_init;  /* 0 = reference sizing                      */
  int32_t canonical;    /* 1: labels_out = min vertex id of the SCC  */
  int32_t timing;       /* 1: fill the phase times of psc_stats;
                           2: also per-kernel events (psc_kernel_profile);
                           3: also one first-reach level per host poll,
                              reported per level on stderr            */
  uint64_t tail_max;    /* a graph with at most this many unfinished
                           vertices after the first SCC, or after a
                           light host-driven round (and at most 2^20
                           remainder vertices), runs its remaining
                           multi-pivot rounds in one launch; 0 =
                           host-driven rounds only; default 65536     */
  int32_t pivot_reach;  /* 1 (default): a host-driven round whose
                           windows hold one pivot position each runs
                           as bitmap reaches (same labels as the
                           multi-search); 0: always multi-search      */
  int32_t capture;      /* 1 (default): the level batches enqueued
                           between host polls are captured once into
                           hipGraphs and replayed; 0: plain launches.
                           A capture excludes the library's allocating
                           calls in every other thread for its
                           duration (the labels are the same either
                           way)                                       */
  uint64_t vgc_batches; /* serial 64-vertex batches one wave's local
                           (VGC) search may run per level launch
                           before its new vertices go to the shared
                           frontier; 0 = default (64).  Bounds how much
                           of tau a lone wave spends serially         */
} psc_params;

/* Breakdown of one psc_run (src/scc.hpp:94-114 timers plus counters).
 * Times are device times between HIP events on the handle's stream, in ms. */
typedef struct psc_stats {
  double ms_total;     /* labels-init to labels-ready               */
  double ms_init;      /* init + trim          (init_timer)         */
  double ms_first;     /* first-SCC reaches    (first_round_timer)  */
  double ms_multi;     /* multi-search rounds  (multi_search_timer) */
  double ms_resize;    /* searches rerun after table overflow       */
  double ms_finalize;  /* final label rewrite (+ canonicalisation)  */
  uint64_t n_trimmed;
  uint64_t first_scc_size;
  uint64_t rounds;
  uint64_t sub_rounds;
  uint64_t levels_fwd;
  uint64_t levels_bwd;
  uint64_t resizes;
  uint64_t host_syncs;
  /* algorithmic traffic counters (bytes), counted on the device */
  uint64_t bytes_csr;    /* offsets + adjacency read                 */
  uint64_t bytes_label;  /* label / visited reads and writes         */
  uint64_t bytes_table;  /* table probes and writes                  */
  uint64_t bytes_front;  /* frontier / log appends and reads         */
  uint64_t edges_scanned;
  uint64_t table_inserts;
} psc_stats;

typedef struct psc_handle psc_handle;

/* Fill p with the reference defaults (src/scc.cpp:40-46). */
void psc_params_default(psc_params *p);

/* Upload both CSRs once and allocate the device working set
 * (SCC::SCC, src/scc.hpp:81-92).  Host arrays are only read.  The set-up
 * also picks, for each vertex, one in- and one out-neighbour of large
 * degree (read by the pull levels). */
int psc_create(const psc_graph *g, psc_handle **out);

/* Same, for a graph whose four arrays already live in device memory of the
 * current HIP device; the handle borrows them (caller keeps them alive, and
 * unchanged while the handle is bound to them: the gathered neighbours
 * would go stale; re-bind with psc_batch_rebind or re-create instead). */
int psc_create_device(const psc_graph *dev_g, psc_handle **out);

/* SCC::scc (src/scc.hpp:277-479): labels_out[n] on the host receives
 * (internal label & 0xffffffff) - 1, equal iff same SCC, or the minimum
 * vertex id of the SCC when p->canonical.  stats may be NULL. */
int psc_run(psc_handle *h, const psc_params *p, uint64_t *labels_out,
            psc_stats *stats);

/* Same, but labels stay in device memory: d_labels_out is a device pointer
 * with room for n uint64 values. */
int psc_run_device(psc_handle *h, const psc_params *p, uint64_t *d_labels_out,
                   psc_stats *stats);

/* Allocation-free SCC.  No counterpart in the reference: for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  psc_create_device + psc_run_device
 * in one call, with no handle: the graph (device pointers) and d_labels_out
 * (n uint64, device) are the caller's; every working buffer comes from a
 * fixed per-device scratch array of PSC_NOALLOC_SCRATCH_BYTES, held from the
 * device's first call until the process ends.  The call makes no device or
 * pinned-host allocation and creates no stream or event: it runs on `stream`
 * (a hipStream_t; NULL = the null stream) and returns after synchronising
 * it.  Labels are those psc_run_device returns for the same graph and
 * params; p->timing and p->capture are ignored (ms_total is host time).
 * Graphs with n > PSC_NOALLOC_MAX_N or m > PSC_NOALLOC_MAX_M are rejected
 * with PSC_ERR_INVALID, a run whose tables outgrow the scratch returns
 * PSC_ERR_NOMEM; both leave a message in psc_last_error().  Calls on one
 * device share the scratch array: calls that overlap on one device race. */
#define PSC_NOALLOC_SCRATCH_BYTES (256ull << 20)
#define PSC_NOALLOC_MAX_N (1ull << 20)
#define PSC_NOALLOC_MAX_M (1ull << 24)
int psc_run_noalloc(const psc_graph *dev_g, const psc_params *p, uint64_t *d_labels_out,
                    void *stream, psc_stats *stats);
/* Size of the scratch array, and the most of it any call on the current
 * device has used so far. */
uint64_t psc_noalloc_scratch_bytes(void);
uint64_t psc_noalloc_scratch_peak(void);

/* SCC_status (src/scc.cpp:13-29) computed on the device from labels in
 * [0, n] held in device memory: number of SCCs and the largest size. */
int psc_status_device(psc_handle *h, const uint64_t *d_labels,
                      uint64_t *n_scc, uint64_t *largest);

/* Per-kernel breakdown of the last psc_run made with timing >= 2: one HIP
 * event pair around every launch, summed per kernel, plus the algorithmic
 * bytes each kernel moved (CSR, labels, table, frontier; counted on the
 * device).  Returns the number of entries written. */
typedef struct psc_kernel_prof {
  char name[32];
  uint64_t launches;
  double total_ms;
  double bytes;
} psc_kernel_prof;

int psc_kernel_profile(psc_handle *h, psc_kernel_prof *out, int cap);

void psc_destroy(psc_handle *h);

/* Device-memory plumbing for callers without a HIP runtime of their own
 * (no reference counterpart): allocate / free device memory of the current
 * device and copy in either direction (hipMemcpyDefault). */
int psc_device_alloc(size_t bytes, void **ptr);
int psc_device_free(void *ptr);
int psc_device_copy(void *dst, const void *src, size_t bytes);
/* Page-locked host memory (hipHostMalloc): label buffers that downloads
 * fill at full link speed. */
int psc_host_alloc(size_t bytes, void **ptr);
int psc_host_free(void *ptr);

/* The seeded pivot order of the first round, the stand-in for
 * parlay::random_shuffle of the non-trimmed vertices (src/scc.hpp:303):
 * element i of a permutation of [0, n), or with inverse != 0 the position of
 * value i (host function, i < n). */
uint64_t psc_pivot_perm(uint64_t n, uint64_t seed, uint64_t i, int inverse);

/* Message of the last error on this thread ("" if none). */
const char *psc_last_error(void);

/* HIP stream a handle runs on (as void*, a hipStream_t). */
void *psc_handle_stream(psc_handle *h);

/* ------------------------------------------------------------------ I/O
 * Host-side graph container (owned arrays). */
typedef struct psc_host_graph {
  psc_graph g;      /* views into the arrays below */
  uint64_t *offset;
  uint32_t *edges;
  uint64_t *in_offset;
  uint32_t *in_edges;
} psc_host_graph;

void psc_host_graph_free(psc_host_graph *hg);

/* GBBS .bin reader: large=0 reads one CSR and builds the in-CSR by
 * transposition (read_graph -> read_binary -> make_directed,
 * src/graph.hpp:120-185,222-245,248-277); large=1 reads the out-CSR
 * followed by the in-CSR (read_large_graph, src/graph.hpp:279-315). */
int psc_read_bin(const char *path, int large, psc_host_graph **out);

/* Writer for the same two layouts. */
int psc_write_bin(const char *path, const psc_graph *g, int large);

/* Build a host graph from an out-CSR, computing the in-CSR as
 * make_directed does (src/graph.hpp:222-245): in-lists sorted by source. */
int psc_make_directed(uint64_t n, uint64_t m, const uint64_t *offset,
                      const uint32_t *edges, psc_host_graph **out);

/* --------------------------------------------------------- generators
 * Seeded synthetic inputs (no reference counterpart: the workloads the
 * benchmark configurations name).  All remove self-loops and duplicate
 * edges and return sorted adjacency lists in both directions.  The device
 * versions write the identical graph straight into HBM. */

/* RMAT: 2^scale vertices, `samples` edge samples with quadrant
 * probabilities (a, b, c, 1-a-b-c); permute=1 relabels vertex ids by a
 * seeded bijection. */
int psc_gen_rmat(uint32_t scale, uint64_t samples, double a, double b,
                 double c, uint64_t seed, int permute, psc_host_graph **out);

/* rows x cols lattice, id = r*cols + c; each 4-neighbour pair is kept in
 * both directions with probability p_both, else in one random direction. */
int psc_gen_grid(uint32_t rows, uint32_t cols, double p_both, uint64_t seed,
                 psc_host_graph **out);

/* n points uniform in the unit cube (21-bit integer coordinates); edge
 * i -> j for the k nearest j != i (ties broken by smaller id). */
int psc_gen_knn3d(uint64_t n, uint32_t k, uint64_t seed, psc_host_graph **out);

/* Device-resident graph produced by a device generator. */
typedef struct psc_dev_graph {
  psc_graph g; /* device pointers */
} psc_dev_graph;

int psc_dev_gen_rmat(uint32_t scale, uint64_t samples, double a, double b,
                     double c, uint64_t seed, int permute,
                     psc_dev_graph **out);
int psc_dev_gen_grid(uint32_t rows, uint32_t cols, double p_both,
                     uint64_t seed, psc_dev_graph **out);
void psc_dev_graph_free(psc_dev_graph *dg);

/* make_directed (src/graph.hpp:222-245) on the device: copies the out-CSR
 * given in device memory (offsets u64[n+1], edges u32[m], kept as given:
 * duplicates and self-loops stay) and builds the in-CSR on the GPU, in-lists
 * sorted by source as the reference's sort of (v, u) pairs leaves them.
 * The result is identical to psc_make_directed; free it with
 * psc_dev_graph_free.  Rejects offsets that do not run from 0 to m without
 * falling, and endpoints >= n. */
int psc_dev_make_directed(uint64_t n, uint64_t m, const uint64_t *d_offset,
                          const uint32_t *d_edges, psc_dev_graph **out);

/* The .bin reader of psc_read_bin with the graph left in device memory:
 * for one CSR in the file (large=0, not symmetric) the out-CSR is uploaded
 * and the in-CSR built on the device (psc_dev_make_directed) instead of by
 * a host transpose.  Free with psc_dev_graph_free. */
int psc_read_bin_device(const char *path, int large, psc_dev_graph **out);

/* ------------------------------------------------------ connectivity
 * Connected components of the undirected graph under an out-CSR (every arc
 * joins its endpoints): the partition of Connectivity
 * (src/Connectivity.h:37-119, LDD + union-find).  The reference's label
 * values depend on the order of its concurrent CASes; here labels[v] is
 * the minimum vertex id of v's component, the same on every run.  The
 * in-CSR is optional (in_offset / in_edges may be NULL); when given it must
 * be the transpose of the out-CSR, and vertices of the largest component
 * found by sampling then skip their remaining arcs.  The device version
 * takes device pointers and returns the kernels' time in *ms (may be
 * NULL). */
int psc_connectivity(const psc_graph *g, uint32_t *labels);
int psc_connectivity_device(const psc_graph *dev_g, uint32_t *d_labels, double *ms);

/* Connectivity with spanning_forest = true (src/Connectivity.h:37-40,
 * 100-118): the labels as above, and in forest[2 i], forest[2 i + 1]
 * (i < *count = n - number of components) the arcs (u, v) of a spanning
 * forest -- the arcs whose union-find hook joined two components.  The
 * forest holds up to n - 1 arcs (2 (n - 1) uint32).  Which arcs are
 * chosen depends on the order of concurrent hooks, as in the reference;
 * the partition they span does not. */
int psc_spanning_forest(const psc_graph *g, uint32_t *labels, uint32_t *forest,
                        uint64_t *count);
int psc_spanning_forest_device(const psc_graph *dev_g, uint32_t *d_labels, uint32_t *d_forest,
                               uint64_t *count, double *ms);

/* ------------------------------------------------------------ LE-lists
 * Least-element lists (LE_LIST::le_list, src/LeList.hpp:127-328; the
 * LeList app, src/LeList.cpp:5-58).  `order` is the reference's `centers`
 * (a permutation of 0..n-1; the reference draws it at random): for each
 * target t, in order of the sources, the rows (t, s, d(s, t)) whose
 * distance along the out-arcs is below that of every earlier source.
 * Rows are 3 uint32 (target, source vertex, distance) sorted by (target,
 * position of the source in order) -- the reference's `list` after the
 * filter.  beta (>= 1) grows the source batches as -beta does (default
 * 1.5); the rows do not depend on it.  Only the out-CSR is read.
 *
 * psc_le_lists takes host arrays and returns *rows in host memory (free
 * with psc_le_free); psc_le_lists_device takes device pointers, returns
 * *d_rows in device memory (free with psc_le_free_device) and the kernels'
 * time in *ms (may be NULL).  *count receives the number of rows. */
int psc_le_lists(const psc_graph *g, const uint32_t *order, double beta, uint32_t **rows,
                 uint64_t *count);
int psc_le_lists_device(const psc_graph *dev_g, const uint32_t *d_order, double beta,
                        uint32_t **d_rows, uint64_t *count, double *ms);
void psc_le_free(uint32_t *rows);
int psc_le_free_device(uint32_t *d_rows);

/* Copy a device graph back to host arrays (for checking). */
int psc_dev_graph_download(const psc_dev_graph *dg, psc_host_graph **out);

/* ------------------------------------------------------------- batches
 * Many graphs in one handle (no reference counterpart; the reference runs
 * one graph per process).  Packed form: graph g has n_g = vbase[g+1] -
 * vbase[g] vertices and m_g = ebase[g+1] - ebase[g] edges; its n_g + 1
 * local offsets (0..m_g) sit at offset[vbase[g] + g], its uint32 local
 * neighbour ids at edges[ebase[g]]; the in-CSR likewise.  The batch is held
 * on the device as one disjoint union; each graph keeps its own trim
 * numbering, pivot permutation (seed), label namespace and pivot-batch
 * schedule, so its labels equal a stand-alone psc_run of that graph. */
typedef struct psc_batch_desc {
  uint64_t num_graphs;
  const uint64_t *vbase;      /* num_graphs + 1 */
  const uint64_t *ebase;      /* num_graphs + 1 */
  const uint64_t *offset;     /* vbase[G] + G   */
  const uint32_t *edges;      /* ebase[G]       */
  const uint64_t *in_offset;  /* vbase[G] + G   */
  const uint32_t *in_edges;   /* ebase[G]       */
} psc_batch_desc;

/* on_device = 1: the desc arrays are device pointers. */
int psc_batch_create(const psc_batch_desc *b, int on_device, psc_handle **out);

/* Fold device-resident graphs (arrays in device memory) into one batch
 * handle; the graphs are copied, the caller may free them after.  Graph g's
 * ids and offsets are shifted by id_base[g] / edge_base[g] (NULL means 0:
 * local ids). */
int psc_batch_create_from_graphs(const psc_graph *dev_graphs, uint64_t count,
                                 const uint64_t *id_base,
                                 const uint64_t *edge_base, psc_handle **out);

/* Re-bind a handle made by psc_batch_create_from_graphs (or psc_create) to
 * a new set of device graphs, reusing its stream and device buffers (they
 * only grow).  Used by serving lanes.  Graph g's ids and offsets may be
 * shifted by id_base[g] / edge_base[g] (views into a device batch, see
 * psc_dev_batch_view); NULL means 0. */
int psc_batch_rebind(psc_handle *h, const psc_graph *dev_graphs, uint64_t count,
                     const uint64_t *id_base, const uint64_t *edge_base);

/* Serving lanes (no reference counterpart).  psc_lane_create makes a
 * re-bindable handle that carves every device buffer -- the union copy of
 * the graphs psc_batch_rebind gathers, the working set, the tables -- from
 * one region of arena_bytes allocated here, so the lane holds exactly that
 * much device memory whatever it runs.  A rebind or run that does not fit
 * returns PSC_ERR_NOMEM ("scratch exhausted") and the lane stays usable for
 * a smaller binding.  psc_batch_layout_bytes gives the arena bytes a
 * binding of num_graphs graphs with n vertices and m edges in all takes
 * before its first run (the run adds its round buffers). */
int psc_lane_create(uint64_t arena_bytes, psc_handle **out);
uint64_t psc_batch_layout_bytes(uint64_t n, uint64_t m, uint64_t num_graphs);

/* Device bytes a handle holds (its arena, or its buffers), and the most of
 * its arena ever in use (= held for a handle without arena). */
int psc_handle_device_bytes(psc_handle *h, uint64_t *held, uint64_t *peak);

/* Free and total device memory as the driver reports them. */
int psc_device_mem_info(uint64_t *free_bytes, uint64_t *total_bytes);

/* The current device's stream-ordered memory pool, which handles and device
 * batches allocate from: bytes it holds from the driver and bytes of live
 * allocations; psc_pool_trim gives unused memory back, keeping keep_bytes. */
int psc_pool_usage(uint64_t *reserved, uint64_t *used);
int psc_pool_trim(uint64_t keep_bytes);

/* Number of graphs of a handle (1 for psc_create) and their vertex prefix
 * (num_graphs + 1 entries). */
uint64_t psc_handle_graphs(psc_handle *h);
int psc_handle_vbase(psc_handle *h, uint64_t *vbase_out);

/* SCC of every graph of the batch.  labels_out (host, vbase[G] entries,
 * may be NULL) receives each graph's labels over its local ids, laid out
 * like the vertices; beta_per_graph (may be NULL) overrides p->beta per
 * graph.  tau and dense_div of p apply to the whole batch. */
int psc_batch_run(psc_handle *h, const psc_params *p,
                  const double *beta_per_graph, uint64_t *labels_out,
                  psc_stats *stats);

/* Seeded description of a mixed batch: kind_out[g] in {0 RMAT, 1 grid,
 * 2 3D k-NN} among the kinds set in kinds_mask (bit k = kind k), vertex
 * count log-uniform in [2^lo_log2, 2^hi_log2] (RMAT rounded to a power of
 * two, grids to a square). */
int psc_batch_spec(uint64_t num_graphs, double lo_log2, double hi_log2,
                   uint64_t seed, uint32_t kinds_mask, uint32_t *kind_out,
                   uint64_t *n_out);

typedef struct psc_dev_batch psc_dev_batch;

/* Generate a batch straight into HBM as a disjoint union: RMAT graphs get
 * 16 n samples, (a,b,c,d) = (.57,.19,.19,.05), scrambled ids; grids keep a
 * neighbour pair in both directions with probability 0.6; k-NN uses k = 5.
 * Graph g uses seeds[g], or splitmix64(seed * 0x9E3779B97F4A7C15 + g + 1)
 * when seeds is NULL; with that seed it equals psc_gen_rmat(log2 n, 16 n,
 * .57, .19, .19, seed_g, 1), psc_gen_grid(side, side, 0.6, seed_g) or
 * psc_gen_knn3d(n, 5, seed_g). */
int psc_dev_gen_batch(uint64_t num_graphs, const uint32_t *kind,
                      const uint64_t *n, uint64_t seed, const uint64_t *seeds,
                      psc_dev_batch **out);
void psc_dev_batch_free(psc_dev_batch *db);
/* Device bytes psc_dev_gen_batch takes for the same kinds and sizes:
 * *keep = what the returned batch holds, *peak = all it allocates at once
 * while generating (keep included; pool-granular). */
int psc_dev_gen_batch_bytes(uint64_t num_graphs, const uint32_t *kind, const uint64_t *n,
                            uint64_t *keep, uint64_t *peak);
/* The union graph of a device batch (device pointers, global ids; for a
 * one-graph batch these are the graph's local ids). */
int psc_dev_batch_union(const psc_dev_batch *db, psc_graph *out);
int psc_dev_batch_info(const psc_dev_batch *db, uint64_t *num_graphs,
                       uint64_t *vbase, uint64_t *ebase, uint32_t *kind);
/* Handle over a device batch (borrowed; keep the batch alive). */
int psc_batch_create_dev(const psc_dev_batch *db, psc_handle **out);
/* Graph g of a device batch as a view into the union (no copy) and the
 * id / offset bases to pass to psc_batch_rebind. */
int psc_dev_batch_view(const psc_dev_batch *db, uint64_t g, psc_graph *out,
                       uint64_t *id_base, uint64_t *edge_base);
/* Graph g of a device batch as its own device graph with local ids. */
int psc_dev_batch_extract(const psc_dev_batch *db, uint64_t g,
                          psc_dev_graph **out);

#ifdef __cplusplus
}
#endif

#endif /* PSC_H */
