// Host-side owner of one graph's device working set (the SCC object of
// src/scc.hpp:38-115): CSR views, labels, reach bitmaps and frontiers,
// pivot order, label tables, logs and pair frontiers.  Buffers grow on
// demand and are reused across runs.
#pragma once
#include <functional>
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/psc.h"

namespace psc {

struct ReachSearch;
struct MSSearch;

void set_error(const std::string &msg);
extern thread_local std::string g_last_error;

// Error propagation of the host code: a failed HIP call records its message
// and returns PSC_ERR_HIP; PSC_TRY forwards a non-zero status.
#define PSC_HIP(call)                                                     \
  do {                                                                    \
    hipError_t e_ = (call);                                               \
    if (e_ != hipSuccess) {                                               \
      set_error(std::string(#call) + ": " + hipGetErrorString(e_));       \
      return PSC_ERR_HIP;                                                 \
    }                                                                     \
  } while (0)

#define PSC_TRY(expr)            \
  do {                           \
    int r_ = (expr);             \
    if (r_ != PSC_OK) return r_; \
  } while (0)


// Capture gate.  A stream in capture mode must not see the runtime's
// allocation, free, stream / event creation and destruction, graph
// instantiation or device-wide synchronisation happen in other threads of
// the process at the same time (a serving process whose lanes create, bind
// and free while one lane captured its level batches ended in a host
// segmentation fault inside the runtime).  Every such call of this library
// holds the gate shared (GateGuard, re-entrant per thread); capturing and
// instantiating a level batch holds it exclusively (CaptureGuard), after
// the calls in flight have left and before new ones enter.
void gate_enter();
void gate_leave();
struct GateGuard {
  GateGuard() { gate_enter(); }
  ~GateGuard() { gate_leave(); }
  GateGuard(const GateGuard &) = delete;
  GateGuard &operator=(const GateGuard &) = delete;
};
struct CaptureGuard {
  CaptureGuard();
  ~CaptureGuard();
  CaptureGuard(const CaptureGuard &) = delete;
  CaptureGuard &operator=(const CaptureGuard &) = delete;
};

// Device memory comes from the device's stream-ordered pool, allocated
// and freed in the order of the owning handle's stream: growing a buffer
// never waits for other streams (a plain hipFree waits for the device).
void pool_setup();
hipStream_t pool_free_stream();

// A fixed device region that an engine's buffers are carved from instead of
// the pool (bump allocation, 256-byte aligned; emptied when the engine is
// bound to a graph).  A buffer that grows takes a new piece; the old one is
// not reused before the next binding.
struct Arena {
  char *base = nullptr;
  size_t cap = 0, used = 0, peak = 0;
  void *take(size_t bytes);
};

struct DevBuf {
  void *ptr = nullptr;
  size_t cap = 0;
  hipStream_t *owner = nullptr;   // the stream that orders its (re)allocation
  Arena *const *arena = nullptr;  // the owning engine's arena (null: pool)
  bool carved = false;            // a piece of the engine's set-up block
  int ensure(size_t bytes);
  void release();
  template <class T> T *as() const { return static_cast<T *>(ptr); }
};

// A remainder stays in (or returns to) the single-workgroup rounds of the
// tail while a round scans at most this many edges times beta.
constexpr double TAIL_ROUND_EDGES = 262144.0;

// Pinned staging layout (bytes): poll buffer, reach descriptors,
// multi-search descriptors, flags.
constexpr size_t PIN_POLL = 0, PIN_REACH = 1024, PIN_MS = 2048, PIN_FLAGS = 3072;

// kernel kinds timed by the per-launch profiler (timing >= 2)
enum : int {
  KK_STEP, KK_PULL, KK_HEAVY, KK_MS_EXPAND, KK_MS_HEAVY, KK_PARTITION, KK_SELECT,
  KK_SEED, KK_TAIL, KK_MISC, KK_N
};
extern const char *const KK_NAMES[KK_N];

struct KProf {
  uint64_t launches = 0;
  double ms = 0;
  double bytes = 0;
};

// A handle's stream with its timing events and small pinned staging
// buffer; destroyed handles return theirs to a process-wide cache.
struct StreamKit {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev[8] = {};
  void *pinned = nullptr;
  void *gpin = nullptr;    // per-graph staging and its capacity (u64 entries)
  size_t gpin_cap = 0;
  void *tpin = nullptr;    // remainder descriptors and their capacity (bytes)
  size_t tpin_cap = 0;
};
bool kit_put(const StreamKit &k);
bool kit_take(int device, StreamKit *k);

struct Engine {
  uint64_t n = 0, m = 0;
  int device = 0, num_cu = 1;
  hipStream_t stream = nullptr;
  bool ready = false;       // setup() done
  // Arena mode: every device buffer comes from `arena` (re-carved at each
  // binding, arena_reset).  A serving lane owns its arena (lane_arena, one
  // allocation made when the lane is created); the allocation-free entry
  // point uses the library's static scratch and also sets host_static: the
  // stream is then the caller's, host staging is the engine's own (pageable)
  // memory and no event is created, so the engine makes no device,
  // pinned-host, stream or event allocation at all.
  Arena *arena = nullptr;
  bool host_static = false;
  Arena lane_arena;
  bool lane = false;        // a serving lane: re-bindable, graphs copied in
  unsigned long long *labels_ext = nullptr;   // allocation-free: the output labels
  void arena_reset();
  static constexpr size_t HOST_STAGE = 8192;
  alignas(64) unsigned char host_stage[HOST_STAGE];
  hipEvent_t ev[8] = {};
  void *h_pinned = nullptr;
  bool owned_graph = false;
  const uint64_t *d_off = nullptr, *d_ioff = nullptr;
  const uint32_t *d_adj = nullptr, *d_iadj = nullptr;
  DevBuf g_off, g_adj, g_ioff, g_iadj;   // owned copies (psc_create)
  DevBuf labels, vis[2], rq[2][2], hv[2], hc[2];
  DevBuf rhead[2];   // pull head of each vertex, per search (k_pull_heads)
  DevBuf hubs;       // degree histogram and hub table of the head selection
  DevBuf centers, front, tiles, tile_off;
  DevBuf trim_off;   // exclusive non-trimmed count per tile (kept past selects)
  DevBuf trim_bm;   // exclusions of a single-pivot round (per-run scratch)
  DevBuf trim_static;   // trimmed vertices (zero in- or out-degree), one bit each
  // dead ends of the two searches (no out-edge / no in-edge), two bitmaps:
  // a run's visited bitmaps start as copies of them
  DevBuf dead;
  // per-graph non-trimmed / trimmed vertices before each graph (nstart,
  // then zstart, G + 1 each), fixed by the binding like the bitmaps above
  std::vector<uint64_t> starts_h;
  DevBuf pbits, sort_tmp;    // remaining vertices by pivot position; scan scratch
  // single-launch rounds of small remainders (tail.hip)
  DevBuf tail_cnt, tail_lab, tail_off, tail_adj, tail_table, tail_lists, tail_meta,
      tail_spec, tail_trace;
  void *tail_pin = nullptr;   // pinned staging of the tail descriptors
  size_t tail_pin_cap = 0;
  int tail_pin_ensure(size_t bytes);
  DevBuf reach_desc, reach_ctl, ms_desc, ms_ctl, misc, scratch32;
  DevBuf table[2], log[2], mq[2][2], mhp[2], mhc[2];
  int grid_step = 1, grid_ms = 1, grid_heavy = 1, grid_pull = 1;
  int prof_level = 0;
  std::vector<hipEvent_t> pev;
  size_t pev_used = 0;
  std::vector<int> pkind;
  KProf kprof[KK_N];
  double kbytes[KK_N] = {};
  // graphs held by this handle: graph g owns vertices [voff[g], voff[g+1])
  uint32_t G = 1;
  std::vector<uint64_t> voff_h;
  std::vector<double> beta_g;     // per-graph pivot-batch growth (batches)
  DevBuf voff_d, gA, gB, gC, gD, gid_d;
  DevBuf rnd;   // per-round windows: prefix (A+1), sources (A), graphs (A)
  DevBuf rr_plab;   // per-graph pivot label of a single-pivot round
  unsigned long long *h_gpin = nullptr;   // pinned per-graph staging
  size_t h_gpin_cap = 0;
  uint64_t log_n[2] = {0, 0};
  uint32_t ms_levels_hint = 2;   // sub-rounds of the previous multi-search
  uint32_t reach_levels_hint = 0; // levels of the previous first reach
  uint32_t round_levels_hint = 0; // levels of the previous single-pivot round
  uint32_t reach_levels_prev = 0, round_levels_prev = 0;   // the last run's depth
  // set while the labelling is enqueued ahead of the poll that shows the
  // first reach over: k_first_label checks level label_gate_level itself
  bool label_gate = false;
  uint32_t label_gate_level = 0;
  uint64_t nnz_total = 0;         // non-trimmed vertices of the current run
  uint64_t syncs = 0;
  unsigned char h_reach[512];
  unsigned char h_ms[256];
  unsigned char h_ms_dev[256];   // descriptors last copied to the device
  bool ms_dev_valid = false;

  ~Engine();
  Engine();
  std::vector<DevBuf *> all_bufs();
  // set-up blocks: the buffers a binding needs, allocated together (one
  // pool allocation instead of one per buffer) and freed with the engine
  std::vector<void *> setup_blocks;
  uint64_t setup_bytes = 0;
  int ensure_all(const std::vector<std::pair<DevBuf *, size_t>> &req);
  std::vector<std::pair<DevBuf *, size_t>> init_requests(uint64_t n, uint64_t m, uint64_t G);
  // one-time device set-up: stream, events, pinned staging, launch shapes
  int setup();
  int init(const psc_graph *dg, bool owned, uint32_t ngraphs = 1,
           const uint64_t *voff = nullptr);
  static constexpr int NREG = 6;
  int gpin(size_t count);
  unsigned long long *region(int r);
  int upload(DevBuf &dst, const uint64_t *src, size_t count, int reg);
  int upload32(DevBuf &dst, const uint32_t *src, size_t count, int reg);
  int upload_round(const std::vector<uint64_t> &wpre, const std::vector<uint64_t> &wsrc,
                   const std::vector<uint32_t> &wgid);
  int run(const psc_params &p, unsigned long long *d_out, psc_stats *st);
  int status(const unsigned long long *d_labels, uint64_t *n_scc,
             uint64_t *largest);

  template <class F>
  int select(F f, uint64_t len, typename F::T *sel, typename F::T *rej,
             uint64_t *count, unsigned long long *count_dev = nullptr);
  int poll(void *dst, const void *src, size_t bytes);
  int scan_tiles(uint64_t nt, unsigned long long *total,
                 unsigned long long *out = nullptr);
  // copy the n labels to host memory on this handle's stream
  int download_labels(uint64_t *dst);
  int prof_reset(int level);
  void prof_begin(int kind);
  void prof_end();
  int prof_collect();
  // rr != nullptr: the reaches of a single-pivot round (round_reach)
  struct ReachRound {
    const uint32_t *front;             // the round's pivots (device)
    const unsigned long long *nfront;  // their count (device)
    uint32_t A;                        // windows, their graphs (device)
    const uint32_t *wgid;
    uint64_t n_rem;                    // remainder positions
    uint64_t reached[2];               // out: vertices reached per direction
  };
  // `after_levels` (may be empty) enqueues work behind a batch of levels,
  // before the poll that tells whether the reach has ended, once the batch
  // reaches the previous run's level count; *after_ran tells whether it ran
  // behind the last batch
  int first_reach(const psc_params &p, psc_stats *st, ReachRound *rr = nullptr,
                  const std::function<int()> &after_levels = {},
                  bool *after_ran = nullptr);
  int round_reach(const psc_params &p, uint32_t A, const uint32_t *wgid_d,
                  uint64_t n_rem, psc_stats *st);
  int ms_prepare(int s, uint64_t slots, MSSearch *d);
  // labels -> (label & VAL_MASK) - 1, skipped on the device while `go`
  // (may be null) reads zero
  int finalize_labels(const unsigned long long *go);
  bool finalized = false;   // this run's finalize is already enqueued
  // the first labelling wrote final values for trimmed and first-SCC
  // vertices (all but one per graph): finalize covers the rest only
  bool direct_final = false;
  uint64_t final_rem = 0;   // remaining vertices (centers) of this run
  DevBuf first_trim;        // per graph: its trimmed vertex of rank 0
  int ms_tail(const psc_params &p, const uint64_t *h_cstart, uint64_t n_rem,
              std::vector<uint64_t> &cur, std::vector<uint64_t> &stepg, psc_stats *st,
              bool *used);
  int multi_search(const psc_params &p, uint64_t nfront, uint32_t A,
                   const uint32_t *wgid_d, uint64_t slots[2], psc_stats *st);

  // Captured level batches: the K levels enqueued between two host polls
  // (K x the level's launches) are captured once into a hipGraph per
  // distinct launch sequence and replayed afterwards.  The key holds every
  // kernel argument of the batch; kernels read everything else from
  // device descriptors.
  struct BatchKey {
    int kind;                 // 0 = reach levels, 1 = multi-search levels
    uint32_t L0, K, tau, waves;
    int vgc, g0, g1;
    const void *S, *C, *lab;
    bool operator==(const BatchKey &o) const {
      return kind == o.kind && L0 == o.L0 && K == o.K && tau == o.tau &&
             waves == o.waves && vgc == o.vgc && g0 == o.g0 && g1 == o.g1 && S == o.S &&
             C == o.C && lab == o.lab;
    }
  };
  std::vector<std::pair<BatchKey, hipGraphExec_t>> gcache;
  uint64_t graph_replays = 0, graph_captures = 0;
  template <class F>
  int replay(const BatchKey &k, F enqueue);
};

}  // namespace psc

// The opaque handle of the C ABI.
struct psc_handle {
  psc::Engine eng;
};
