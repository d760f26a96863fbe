// Device-side state layouts and kernel declarations of the single-graph SCC
// path.  One ReachSearch / MSSearch describes one direction of a search; the
// forward and backward searches of src/scc.hpp run fused in the same
// launches (search index 0 = out-CSR, 1 = in-CSR).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "device_common.hpp"

namespace psc {

constexpr int REACH_RING = 512;      // per-wave VGC queue (vertices, LDS)
constexpr int MS_RING = 512;         // per-wave VGC queue (pairs, LDS)
constexpr int OUT_CAP = 512;         // per-wave staged appends (vertices)
constexpr int OUT_CAP64 = 256;       // per-wave staged appends (pairs)
constexpr uint32_t PULL_HEAVY = 64;  // pull remainder above which it is chunked
constexpr uint32_t HEAVY_DEG = 512;  // degree above which edges are chunked
constexpr uint32_t CHUNK = 256;     // edges per heavy work item
constexpr uint32_t PULL_PROBE = 8;   // in-edges a pull lane tests first
constexpr uint32_t VGC_MAX_DEG = 8;  // only vertices this sparse stay local
// A wave's local search also stops taking vertices once it has scanned
// VGC_EDGE_MULT * tau edges per seed: a lone wave walking thousands of
// edges serially costs more than the levels it saves on low-diameter graphs,
// while degree-<=4 lattices never reach the cap.
constexpr uint32_t VGC_EDGE_MULT = 4;
// A wave keeps at most VGC_WAVE_MULT * tau vertices local per launch, over
// all its seed batches: with a large frontier every wave has work anyway,
// and long serial local searches would only leave the launch waiting on a
// few waves.
constexpr uint32_t VGC_WAVE_MULT = 4;
// ... and at most VGC_MAX_BATCHES serial 64-vertex batches per launch: the
// cost of a local search is its number of batches (each a chain of
// dependent loads), and a launch waits for its slowest wave.
constexpr uint32_t VGC_MAX_BATCHES = 64;
// Pull cost estimate per unvisited non-trimmed vertex, in pushed-edge
// units (measured: a pulled vertex costs about as much as a pushed edge).
constexpr uint64_t PULL_VERTEX_COST = 1;
// Fixed cost of a pull level's bitmap walk, in pushed-edge units: n / this.
constexpr uint64_t PULL_WALK_DIV = 64;
constexpr unsigned long long H_CNT_SHIFT = 40;
constexpr unsigned long long H_CHUNK_MASK = (1ull << 40) - 1;

enum : int { MODE_NONE = 0, MODE_PUSH = 1, MODE_PULL = 2 };

// ---------------------------------------------------------------- reach
struct ReachSearch {
  const uint64_t *off;   // push direction CSR
  const uint32_t *adj;
  const uint64_t *roff;  // pull direction CSR (transpose)
  const uint32_t *radj;
  const uint32_t *rhead; // pull head of each non-trimmed vertex, by rank
  uint32_t *vis;         // visited bitmap, ceil(n/32) words
  const uint32_t *trim;  // vertices a pull level skips: the trimmed ones, or
                         // the excluded ones of a single-pivot round
  const uint32_t *tstat; // trimmed vertices (zero in- or out-degree); the
                         // rank of a non-trimmed v is nzoff[v / 2048] + its
                         // clear bits before v in its 2048-vertex tile
  const unsigned long long *nzoff;   // non-trimmed vertices before each tile
  const uint32_t *excl;  // vertices outside the searched subproblems (push
                         // skips them), or nullptr
  uint32_t *q[2];        // frontier ping-pong, n entries each
  uint32_t *hv;          // heavy vertices of the current level
  unsigned long long *hc;  // exclusive chunk start of each heavy vertex
  uint64_t thresh;       // m / dense_div (src/reach.hpp:49)
  double avg_deg;        // m / n, out-edge estimate after a pull level
  uint32_t n;
  uint32_t pad;
  uint64_t nnz;          // non-trimmed vertices (all graphs)
};

constexpr uint32_t PULL_UNSTORED = 2;
struct ReachSlot {
  uint32_t qsize;            // frontier size of this level
  uint32_t grab;             // seeds handed out
  unsigned long long degsum; // out-degree sum of the frontier
  unsigned long long heavy;  // (count << 40) | chunks
  unsigned long long hgrab;  // heavy chunks handed out
  // frontier produced by a pull level: 1 stored (up to the cap), 2
  // counted only (PULL_UNSTORED), so this level pulls unless it is empty
  uint32_t from_pull;
  // nonzero: the previous (pull) level left in-list remainders to this one,
  // which pulls, completely, whatever its frontier
  uint32_t deferred;
  unsigned long long cum_vis;  // vertices visited before this level
};

// per-search counters (algorithmic traffic accounting)
enum : int {
  R_PUSH_EXP, R_PUSH_EDGES, R_PUSH_VIS, R_PUSH_APP,
  R_HEAVY_EDGES, R_HEAVY_VIS, R_HEAVY_APP, R_HEAVY_ITEMS,
  R_PULL_SCANS, R_PULL_EDGES, R_PULL_VIS, R_PULL_UNITS,
  R_HEAD_TESTS, R_HEAD_VIS, R_HEAD_UNITS,   // head-only pull levels
  NRC
};

struct ReachCtl {
  ReachSlot slot[2][3];      // [search][level % 3]
  uint32_t done[2];
  uint32_t levels[2];
  unsigned long long cnt[2][NRC];
};

// ------------------------------------------------------------ multi-search
struct MSSearch {
  const uint64_t *off;
  const uint32_t *adj;
  unsigned long long *table;    // (vertex << 32 | label) slots
  uint64_t mask;                // logical size - 1
  unsigned long long *log;      // every inserted pair, in insert order
  unsigned long long log_cap;
  unsigned long long *q[2];     // pair frontiers
  unsigned long long q_cap;
  unsigned long long *hp;       // heavy pairs of the current level
  unsigned long long *hc;       // their exclusive chunk starts
  uint32_t probe_limit;
  uint32_t pad;
};

struct MSSlot {
  unsigned long long qsize, grab, heavy, hgrab;
};

enum : int {
  M_EXP, M_EDGES, M_INS, M_APP, M_PROBES,
  M_HEAVY_EDGES, M_HEAVY_INS, M_HEAVY_APP, M_HEAVY_PROBES, M_HEAVY_ITEMS,
  NMC
};

struct MSCtl {
  MSSlot slot[2][3];           // [search][level % 3]
  unsigned long long log_n[2];
  uint32_t full[2];
  uint32_t active[2];
  unsigned long long cnt[2][NMC];
};

}  // namespace psc
