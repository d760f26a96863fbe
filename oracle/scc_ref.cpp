// ORACLE — test infrastructure only.  Only tests/, __graft_entry__.smoke()
// and bench.py's cpu_baseline leg may load this code, and only as the
// checker or as the timed CPU baseline.  The product path never links it.
//
// CPU restatement (C++17 + OpenMP) of the reference's SCC hot path:
//   hash bag            src/hash_bag.h:16-218
//   resizable table     src/resizable_table.h:40-148,196-240,283-312
//   single reach + VGC  src/reach.hpp:44-243
//   multi-search, partition, driver   src/scc.hpp:22-34,64-74,117-479
// ParlayLib primitives (parallel_for, filter, pack, reduce, random_shuffle)
// are restated with OpenMP loops and prefix sums.  parlay::random_shuffle is
// replaced by a seeded bijection (pivot_perm) shared with the GPU path, so
// both pick the same first pivot; raw label values depend on it, the
// canonical partition does not.
//
// Parity: canonical labels of psc_oracle_scc_run are checked against the
// iterative Tarjan (tarjan.c), networkx and scipy on every fixture;
// psc_oracle_table_* and psc_oracle_bag_* are checked against Python sets
// (tests/test_oracle.py).
#include <omp.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <vector>

namespace {

using NodeId = uint32_t;   // src/utilities.h:14-18
using EdgeId = uint64_t;   // src/utilities.h:20-24
constexpr NodeId UINT_N_MAX = 0xFFFFFFFFu;
constexpr uint64_t TOP_BIT = uint64_t(UINT_N_MAX) + 1;  // src/scc.hpp:22
constexpr uint64_t VAL_MASK = UINT_N_MAX;               // src/scc.hpp:23

// ---------------------------------------------------------------- helpers
// 32-bit integer mixer (any good 32-bit hash serves; slot placement is not
// part of the result, src/utilities.h:108-118 uses parlay::hash32).
inline uint32_t hash32(uint32_t x) {
  x ^= x >> 16; x *= 0x7feb352du;
  x ^= x >> 15; x *= 0x846ca68bu;
  x ^= x >> 16;
  return x;
}

template <class T> inline bool cas(T* a, T o, T n) {
  return __atomic_compare_exchange_n(a, &o, n, false, __ATOMIC_SEQ_CST,
                                     __ATOMIC_SEQ_CST);
}
// compare_and_swap: read first, then CAS (src/utilities.h:60-63)
template <class T> inline bool test_cas(T* a, T o, T n) {
  return __atomic_load_n(a, __ATOMIC_RELAXED) == o && cas(a, o, n);
}
// write_max with CAS loop (src/utilities.h:94-101)
inline void write_max(uint64_t* a, uint64_t b) {
  uint64_t c = __atomic_load_n(a, __ATOMIC_RELAXED);
  while (c < b && !__atomic_compare_exchange_n(a, &c, b, false,
                                               __ATOMIC_SEQ_CST,
                                               __ATOMIC_RELAXED)) {
  }
}
template <class T> inline T ld(const T* a) {
  return __atomic_load_n(a, __ATOMIC_RELAXED);
}

// Seeded bijection on [0, N) by cycle-walking a multiply/xorshift bijection
// of the enclosing power of two.  Stands in for parlay::random_shuffle
// (src/scc.hpp:303); the GPU path uses the same function.
struct PivotPerm {
  uint64_t N, mask, a, b;
  int k;
  PivotPerm(uint64_t n, uint64_t seed) : N(n) {
    k = 0;
    while ((1ull << k) < n) k++;
    mask = (k >= 64) ? ~0ull : ((1ull << k) - 1);
    uint64_t s = seed * 0x9E3779B97F4A7C15ull + 0x632BE59BD9B4E019ull;
    a = (s ^ (s >> 31)) & mask;
    s = s * 0xD1B54A32D192ED03ull + 1;
    b = (s ^ (s >> 29)) & mask;
  }
  uint64_t step(uint64_t x) const {
    int sh = k / 2 + 1;
    x = (x * 0x9E3779B1ull + a) & mask;
    x ^= x >> sh;
    x = (x * 0x85EBCA77ull + b) & mask;
    x ^= x >> sh;
    return x;
  }
  uint64_t operator()(uint64_t i) const {
    uint64_t y = step(i);
    while (y >= N) y = step(y);
    return y;
  }
};

// Stable parallel pack of idx in [0,len) with pred(idx) into out; returns
// count.  Restates parlay::filter / pack_into_uninitialized.
template <class Pred, class Val>
uint64_t par_pack(uint64_t len, Pred pred, Val val, NodeId* out) {
  int nt = omp_get_max_threads();
  std::vector<uint64_t> cnt(nt + 1, 0);
#pragma omp parallel num_threads(nt)
  {
    int t = omp_get_thread_num();
    uint64_t lo = len * t / nt, hi = len * (t + 1) / nt, c = 0;
    for (uint64_t i = lo; i < hi; i++) c += pred(i) ? 1 : 0;
    cnt[t + 1] = c;
#pragma omp barrier
#pragma omp single
    for (int j = 0; j < nt; j++) cnt[j + 1] += cnt[j];
    uint64_t o = cnt[t];
    for (uint64_t i = lo; i < hi; i++)
      if (pred(i)) out[o++] = val(i);
  }
  return cnt[nt];
}

// ---------------------------------------------------------------- hash bag
// src/hash_bag.h:16-218.  Chunks of size lambda*2^i; an insert lands in the
// current chunk at hash(u); one in `rate` inserts samples a per-chunk
// counter, and after EXP_SAMPLES samples the chunk pointer advances, so the
// active chunk stays about half full.
constexpr int MAX_PROBES = 100;   // src/hash_bag.h:17
constexpr int EXP_SAMPLES = 64;   // src/hash_bag.h:18

struct HashBag {
  uint64_t n;
  double load_factor = 0.5;
  int pointer = 1;
  int num_hash_bag;
  uint64_t pool_size = 0;
  NodeId empty = UINT_N_MAX;
  std::vector<uint64_t> bag_size;
  std::vector<uint32_t> counter;
  std::vector<NodeId> pool;

  HashBag(uint64_t _n, uint64_t min_bag) : n(_n) {   // src/hash_bag.h:61-76
    uint64_t cur = min_bag;
    bag_size.push_back(0);
    while (pool_size < n / load_factor) {
      bag_size.push_back(cur + bag_size.back());
      pool_size += cur;
      cur *= 2;
    }
    num_hash_bag = (int)bag_size.size();
    counter.assign(num_hash_bag, 0);
    pool.assign(pool_size, empty);
  }
  void set_values(NodeId u, int& lp, uint64_t& len, uint64_t& idx,
                  uint64_t& rate) {                    // src/hash_bag.h:32-41
    lp = ld(&pointer);
    len = bag_size[lp] - bag_size[lp - 1];
    idx = (hash32(u) & (len - 1)) + bag_size[lp - 1];
    rate = std::max(1.0, std::floor(len * load_factor / EXP_SAMPLES));
  }
  // Sampling step shared by both insert flavours (src/hash_bag.h:82-99).
  void sample(NodeId u, int& lp, uint64_t& len, uint64_t& idx,
              uint64_t& rate) {
    set_values(u, lp, len, idx, rate);
    while (idx % rate == 0) {
      bool ok = false;
      uint32_t ret = ld(&counter[lp]);
      while (ret < (uint32_t)EXP_SAMPLES && !ok) {
        ok = cas(&counter[lp], ret, ret + 1);
        ret = ld(&counter[lp]);
      }
      if (ret >= (uint32_t)EXP_SAMPLES && lp + 1 < num_hash_bag)
        cas(&pointer, lp, lp + 1);
      if (lp != ld(&pointer)) set_values(u, lp, len, idx, rate);
      if (ok || lp + 1 == num_hash_bag) break;
    }
  }
  void insert(NodeId u) {                              // src/hash_bag.h:78-120
    int lp;
    uint64_t len, idx, rate;
    sample(u, lp, len, idx, rate);
    for (uint64_t i = 0;; i++) {
      if (cas(&pool[idx], empty, u)) break;
      idx = (idx == bag_size[lp] - 1) ? bag_size[lp - 1] : idx + 1;
      if (i % MAX_PROBES == 0 && lp != ld(&pointer))
        set_values(u, lp, len, idx, rate);
      if (i == len) {
        if (lp == ld(&pointer)) cas(&pointer, lp, lp + 1);
        set_values(u, lp, len, idx, rate);
      }
    }
  }
  void insert(NodeId u, uint32_t* stamp, uint32_t cur) { // src/hash_bag.h:122-172
    int lp;
    uint64_t len, idx, rate;
    sample(u, lp, len, idx, rate);
    uint32_t ks = ld(&stamp[idx]);
    for (uint64_t i = 0;; i++) {
      if (ks != cur && cas(&stamp[idx], ks, cur)) break;
      idx = (idx == bag_size[lp] - 1) ? bag_size[lp - 1] : idx + 1;
      ks = ld(&stamp[idx]);
      if (i % MAX_PROBES == 0 && lp != ld(&pointer))
        set_values(u, lp, len, idx, rate);
      if (i == len) {
        if (lp == ld(&pointer)) cas(&pointer, lp, lp + 1);
        set_values(u, lp, len, idx, rate);
      }
    }
    __atomic_store_n(&pool[idx], u, __ATOMIC_RELAXED);
  }
  uint64_t pack(NodeId* out) {                          // src/hash_bag.h:193-202
    uint64_t len = bag_size[pointer];
    uint64_t r = par_pack(
        len, [&](uint64_t i) { return pool[i] != empty; },
        [&](uint64_t i) { return pool[i]; }, out);
#pragma omp parallel for schedule(static)
    for (uint64_t i = 0; i < len; i++) pool[i] = empty;   // clear() :42-56
    for (int i = 0; i <= pointer; i++) counter[i] = 0;
    pointer = 1;
    return r;
  }
  uint64_t pack(NodeId* out, const uint32_t* stamp, uint32_t cur) { // :204-218
    uint64_t len = bag_size[pointer];
    uint64_t r = par_pack(
        len, [&](uint64_t i) { return stamp[i] == cur; },
        [&](uint64_t i) { return pool[i]; }, out);
    for (int i = 0; i <= pointer; i++) counter[i] = 0;
    pointer = 1;
    return r;
  }
};

// ---------------------------------------------------------------- table
// gbbs::resizable_table<NodeId,NodeId> (src/resizable_table.h:99-313):
// linear probing on hash32(k) & mask, one 64-bit CAS per insert, empty slot
// {~0,~0}.  A slot packs (k << 32) | v.
struct Table {
  uint64_t m = 0, mask = 0;
  std::atomic<uint64_t> ne{0};
  bool full = false;
  std::vector<uint64_t> slots;
  static constexpr uint64_t EMPTY = ~0ull;
  static uint32_t key(uint64_t s) { return (uint32_t)(s >> 32); }
  static uint32_t val(uint64_t s) { return (uint32_t)s; }

  void init(uint64_t want) {                    // src/resizable_table.h:136-148
    m = 1;
    while (m < want) m <<= 1;                   // 1 << log2_up(_m)
    mask = m - 1;
    slots.resize(m);
    clear();
  }
  void clear() {                                // :310-312
#pragma omp parallel for schedule(static)
    for (uint64_t i = 0; i < m; i++) slots[i] = EMPTY;
    ne = 0;
    full = false;
  }
  void double_size() { init(2 * m); }           // :196-204
  // insert: true iff (k,v) is new; gives up after 0.8m probes (:214-240)
  bool insert(uint32_t k, uint32_t v) {
    uint64_t kv = (uint64_t(k) << 32) | v;
    uint64_t h = hash32(k) & mask;
    for (uint64_t c = 0; c < (uint64_t)(0.8 * (double)m); c++) {
      if (ld(&full)) return false;
      uint64_t s = ld(&slots[h]);
      if (s == EMPTY && cas(&slots[h], EMPTY, kv)) {
        ne.fetch_add(1, std::memory_order_relaxed);
        return true;
      }
      if (ld(&slots[h]) == kv) return false;
      h = (h + 1) & mask;
    }
    __atomic_store_n(&full, true, __ATOMIC_RELAXED);
    return false;
  }
  bool contains(uint32_t k, uint32_t v) const {  // :283-294
    uint64_t kv = (uint64_t(k) << 32) | v;
    uint64_t h = hash32(k) & mask;
    for (uint64_t c = 0; c <= mask; c++) {
      uint64_t s = slots[h];
      if (s == kv) return true;
      if (s == EMPTY) return false;
      h = (h + 1) & mask;
    }
    return false;
  }
  // iter_kv init/has_next/next (:49-97): visit every value stored under k.
  template <class F> bool for_each_value(uint32_t k, F f) const {
    uint64_t h = hash32(k) & mask;
    for (uint64_t c = 0; c <= mask; c++) {
      uint64_t s = ld(&slots[h]);
      if (s == EMPTY) return true;
      if (key(s) == k && !f(val(s))) return false;
      h = (h + 1) & mask;
    }
    return true;
  }
};

// ---------------------------------------------------------------- graph
struct GraphView {
  uint64_t n, m;
  const EdgeId* off[2];   // [0] out-CSR, [1] in-CSR  (src/graph.hpp:35-43)
  const NodeId* adj[2];
};

// ---------------------------------------------------------------- REACH
// src/reach.hpp:21-243.  `dir` selects the CSR instead of swap_graph
// (:51-54) so the caller's graph is never mutated.
struct Reach {
  const GraphView& g;
  uint64_t n;
  std::vector<NodeId> front;
  HashBag bag;
  std::vector<uint8_t> dense_front, new_dense;
  bool sparse_pre = true;
  uint64_t n_front = 0, threshold, tau;
  uint64_t num_round = 0;
  int dir = 0;
  std::vector<std::vector<NodeId>> lq;   // per-thread VGC queues (:127)
  std::vector<NodeId> heavy;

  Reach(const GraphView& G, uint64_t _tau, uint64_t dense_div)
      : g(G), n(G.n), front(G.n ? G.n : 1), bag(G.n, 1 << 10),
        dense_front(G.n), new_dense(G.n), tau(_tau) {
    threshold = g.m / dense_div;             // :49
    lq.resize(omp_get_max_threads());
    for (auto& q : lq) q.resize(tau + 1);
    heavy.resize(G.n ? G.n : 1);
  }
  const EdgeId* O() const { return g.off[dir]; }
  const NodeId* A() const { return g.adj[dir]; }
  const EdgeId* IO() const { return g.off[dir ^ 1]; }
  const NodeId* IA() const { return g.adj[dir ^ 1]; }

  void front_sparse2dense() {                 // :76-79
#pragma omp parallel for schedule(static)
    for (uint64_t i = 0; i < n; i++) dense_front[i] = 0;
#pragma omp parallel for schedule(static)
    for (uint64_t i = 0; i < n_front; i++) dense_front[front[i]] = 1;
  }
  void front_dense2sparse() {                 // :81-84
    n_front = par_pack(
        n, [&](uint64_t i) { return dense_front[i] != 0; },
        [&](uint64_t i) { return (NodeId)i; }, front.data());
  }
  uint64_t dense_sample(uint32_t seed) {      // :192-206
    int count = 0;
    uint64_t out_edges = 0;
    for (uint32_t i = 1; count < 50; i++) {
      uint64_t idx = hash32(seed + i) % n;
      if (dense_front[idx]) {
        count++;
        out_edges += O()[idx + 1] - O()[idx];
      }
    }
    return n_front * (out_edges / count);
  }
  bool judge() {                              // :86-112
    uint64_t out_edges = 0;
    bool both = false;
    if (!sparse_pre && n_front < (1u << 14)) {
      front_dense2sparse();
      both = true;
    }
    if (sparse_pre || both) {
#pragma omp parallel for reduction(+ : out_edges) schedule(static)
      for (uint64_t i = 0; i < n_front; i++)
        out_edges += O()[front[i] + 1] - O()[front[i]];
    } else {
      out_edges = dense_sample(hash32((uint32_t)num_round));
    }
    bool sparse_now = (n_front + out_edges) < threshold;
    if (!both && sparse_pre != sparse_now) {
      if (!sparse_now) front_sparse2dense();
      else front_dense2sparse();
    }
    return sparse_now;
  }
  uint64_t sparse_update(uint8_t* dst, bool local) {   // :114-171
    const EdgeId* off = O();
    const NodeId* E = A();
    uint64_t qs = tau;
    std::atomic<uint64_t> nheavy{0};
#pragma omp parallel for schedule(dynamic, 16)
    for (uint64_t i = 0; i < n_front; i++) {
      NodeId node = front[i];
      uint64_t deg_f = off[node + 1] - off[node];
      if (local && deg_f < qs && deg_f > 0) {
        NodeId* Q = lq[omp_get_thread_num()].data();
        uint64_t head = 0, tail = 0, explored = 0;
        Q[tail++] = node;
        while (head < tail && tail < qs) {
          NodeId u = Q[head++];
          uint64_t deg_u = off[u + 1] - off[u];
          if (deg_u > qs) {
            bag.insert(u);
            break;
          }
          for (uint64_t j = off[u]; j < off[u + 1]; j++) {
            NodeId v = E[j];
            if (!ld(&dst[v]) && test_cas(&dst[v], (uint8_t)0, (uint8_t)1)) {
              explored++;
              if (tail < qs && explored < qs) Q[tail++] = v;
              else bag.insert(v);
            }
          }
        }
        for (uint64_t j = head; j < tail; j++) bag.insert(Q[j]);
      } else if (deg_f > 0) {
        if (deg_f > 2048) {   // edge-parallel below (granularity 2048, :156-166)
          heavy[nheavy.fetch_add(1)] = node;
          continue;
        }
        for (uint64_t j = off[node]; j < off[node + 1]; j++) {
          NodeId v = E[j];
          if (!ld(&dst[v]) && test_cas(&dst[v], (uint8_t)0, (uint8_t)1))
            bag.insert(v);
        }
      }
    }
    for (uint64_t h = 0; h < nheavy; h++) {
      NodeId node = heavy[h];
#pragma omp parallel for schedule(dynamic, 2048)
      for (uint64_t j = off[node]; j < off[node + 1]; j++) {
        NodeId v = E[j];
        if (!ld(&dst[v]) && test_cas(&dst[v], (uint8_t)0, (uint8_t)1))
          bag.insert(v);
      }
    }
    return bag.pack(front.data());
  }
  uint64_t dense_update(uint8_t* dst) {       // :173-190
    const EdgeId* ioff = IO();
    const NodeId* iE = IA();
    uint64_t cnt = 0;
#pragma omp parallel for schedule(dynamic, 1024) reduction(+ : cnt)
    for (uint64_t i = 0; i < n; i++) {
      new_dense[i] = 0;
      if (!ld(&dst[i])) {
        for (uint64_t j = ioff[i]; j < ioff[i + 1]; j++) {
          if (dense_front[iE[j]]) {
            dst[i] = 1;
            new_dense[i] = 1;
            cnt++;
            break;
          }
        }
      }
    }
    std::swap(dense_front, new_dense);
    return cnt;   // parlay::count(dense_front, true)
  }
  uint64_t reach(NodeId src, uint8_t* dst, bool local) {   // :208-243
#pragma omp parallel for schedule(static)
    for (uint64_t i = 0; i < n; i++) dst[i] = 0;
    front[0] = src;
    dst[src] = 1;
    n_front = 1;
    sparse_pre = true;
    num_round = 0;
    while (n_front > 0) {
      num_round++;
      bool sparse_now = judge();
      n_front = sparse_now ? sparse_update(dst, local) : dense_update(dst);
      sparse_pre = sparse_now;
    }
    return num_round;
  }
};

// ---------------------------------------------------------------- SCC
struct Params {
  double beta;
  int local_reach, local_scc;
  uint64_t tau, min_bag, dense_div, seed, table_init;
  int threads;
};

struct Stats {
  double t_total, t_init, t_first, t_multi, t_multi_safe;
  uint64_t n_trimmed, first_scc_size, rounds, resizes, reach_levels_fwd,
      reach_levels_bwd, sub_rounds;
};

struct SCC {
  GraphView g;
  uint64_t n;
  Params p;
  HashBag bag;
  std::vector<NodeId> front, substep_front, centers, P;
  std::vector<uint8_t> bits;
  std::vector<uint32_t> bag_stamp;
  uint32_t current_stamp = 1;
  uint64_t n_front = 0, label_offset = 0;
  Table table_forw, table_back;
  std::vector<std::vector<NodeId>> lq;
  Stats st{};

  SCC(const GraphView& G, const Params& P_, uint64_t min_bag)   // :81-92
      : g(G), n(G.n), p(P_), bag(G.n, min_bag) {
    front.resize(n ? n : 1);
    substep_front.resize(n ? n : 1);
    bits.resize(n ? n : 1);
    bag_stamp.resize(bag.pool_size);
    lq.resize(omp_get_max_threads());
    for (auto& q : lq) q.resize(p.tau + 1);
  }
  void time_stamp_update() {                   // :64-74
    current_stamp++;
    if (current_stamp == UINT32_MAX - 1) {
#pragma omp parallel for schedule(static)
      for (uint64_t i = 0; i < bag_stamp.size(); i++)
        bag_stamp[i] = (bag_stamp[i] == current_stamp);
      current_stamp = 1;
    }
  }
  // edge_map: copy every label of cur onto ngb (:117-136)
  bool edge_map(NodeId cur, NodeId ngb, Table& t) {
    bool changed = false;
    t.for_each_value(cur, [&](uint32_t lab) {
      changed |= t.insert(ngb, lab);
      return !ld(&t.full);
    });
    return changed && !ld(&t.full);
  }
  int multi_search(uint64_t* labels, Table& table, bool forward, bool local) {
    // :144-254
#pragma omp parallel for schedule(static)
    for (uint64_t i = 0; i < n_front; i++) {
      table.insert(front[i], (uint32_t)(label_offset + i));
      substep_front[i] = front[i];
    }
    uint64_t sub_n = n_front;
    const EdgeId* off = g.off[forward ? 0 : 1];
    const NodeId* E = g.adj[forward ? 0 : 1];
    int round = 0, sub_round = 0;
    uint64_t qs = p.tau;
    while (sub_n > 0) {
      sub_round++;
      time_stamp_update();
      round++;
#pragma omp parallel for schedule(static)
      for (uint64_t i = 0; i < sub_n; i++) bits[substep_front[i]] = 0;
#pragma omp parallel for schedule(dynamic, 8)
      for (uint64_t i = 0; i < sub_n; i++) {
        NodeId node = substep_front[i];
        uint64_t degree = off[node + 1] - off[node];
        uint64_t cnt = 0;
        if (local && degree < qs) {          // VGC local queue (:177-219)
          NodeId* Q = lq[omp_get_thread_num()].data();
          uint64_t head = 0, tail = 0;
          Q[tail++] = node;
          while (head < tail && cnt < qs) {
            NodeId u = Q[head];
            uint64_t deg = off[u + 1] - off[u];
            if (deg >= qs) break;
            head++;
            for (uint64_t j = off[u]; j < off[u + 1]; j++) {
              NodeId v = E[j];
              uint64_t lv = ld(&labels[v]);
              if (!(lv & TOP_BIT) && lv == ld(&labels[u])) {
                cnt++;
                bool ok = (sub_round == 1 && local)
                              ? table.insert(v, (uint32_t)(label_offset + i))
                              : edge_map(node, v, table);
                if (ok) {
                  if (cnt < qs) Q[tail++] = v;
                  else if (test_cas(&bits[v], (uint8_t)0, (uint8_t)1))
                    bag.insert(v, bag_stamp.data(), current_stamp);
                }
              }
            }
          }
          for (uint64_t j = head; j < tail; j++) {
            NodeId u = Q[j];
            if (test_cas(&bits[u], (uint8_t)0, (uint8_t)1))
              bag.insert(u, bag_stamp.data(), current_stamp);
          }
        } else if (degree > 0) {
          for (uint64_t j = off[node]; j < off[node + 1]; j++) {
            NodeId v = E[j];
            uint64_t lv = ld(&labels[v]);
            if (!(lv & TOP_BIT) && lv == ld(&labels[node])) {
              bool ok = (sub_round == 1 && local)
                            ? table.insert(v, (uint32_t)(label_offset + i))
                            : edge_map(node, v, table);
              if (ok && test_cas(&bits[v], (uint8_t)0, (uint8_t)1))
                bag.insert(v, bag_stamp.data(), current_stamp);
            }
          }
        }
      }
      if (table.full) return -1;
      sub_n = bag.pack(substep_front.data(), bag_stamp.data(), current_stamp);
    }
    st.sub_rounds += round;
    return round;
  }
  int multi_search_safe(uint64_t* labels, Table& t, bool fwd, bool local) {
    // :256-275 — on overflow clear bits, double the table, rerun.
    while (true) {
      int r = multi_search(labels, t, fwd, local);
      if (r != -1) return r;
      st.resizes++;
#pragma omp parallel for schedule(static)
      for (uint64_t i = 0; i < n; i++) bits[i] = 0;
      t.double_size();
    }
  }
  void run(uint64_t* labels) {                  // :277-479
    using clk = std::chrono::steady_clock;
    auto sec = [](clk::time_point a, clk::time_point b) {
      return std::chrono::duration<double>(b - a).count();
    };
    st = Stats{};
    auto t0 = clk::now();
    current_stamp = 1;
    std::vector<uint8_t> dist_1(n), dist_2(n);
#pragma omp parallel for schedule(static)
    for (uint64_t i = 0; i < n; i++) {
      labels[i] = 0;
      bits[i] = 0;
    }
#pragma omp parallel for schedule(static)
    for (uint64_t i = 0; i < bag_stamp.size(); i++) bag_stamp[i] = 0;
    // trim: one pass, zero in- or out-degree (:295-305)
    auto nonzero = [&](uint64_t v) {
      return (g.off[0][v + 1] != g.off[0][v]) && (g.off[1][v + 1] != g.off[1][v]);
    };
    std::vector<NodeId> NZ(n ? n : 1), Z(n ? n : 1);
    uint64_t nnz = par_pack(n, nonzero, [](uint64_t i) { return (NodeId)i; },
                            NZ.data());
    uint64_t nz = par_pack(
        n, [&](uint64_t v) { return !nonzero(v); },
        [](uint64_t i) { return (NodeId)i; }, Z.data());
    P.resize(nnz ? nnz : 1);
    PivotPerm perm(nnz, p.seed);
#pragma omp parallel for schedule(static)
    for (uint64_t i = 0; i < nnz; i++) P[i] = NZ[perm(i)];
#pragma omp parallel for schedule(static)
    for (uint64_t i = 0; i < nz; i++) labels[Z[i]] = 1 + (i | TOP_BIT);
    st.n_trimmed = nz;
    auto t1 = clk::now();
    st.t_init = sec(t0, t1);
    label_offset = nz + 1;
    if (nnz > 0) {
      // first round: forward and backward reach from P[0] (:316-355)
      NodeId source = P[0];
      Reach R(g, p.tau, p.dense_div);
      R.dir = 0;
      st.reach_levels_fwd = R.reach(source, dist_1.data(), p.local_reach);
      R.dir = 1;
      st.reach_levels_bwd = R.reach(source, dist_2.data(), p.local_reach);
      uint64_t label = label_offset;
      uint64_t big = 0;
#pragma omp parallel for schedule(static) reduction(+ : big)
      for (uint64_t i = 0; i < nnz; i++) {
        NodeId v = P[i];
        if (dist_1[v] && dist_2[v]) {
          labels[v] = label | TOP_BIT;
          big++;
        } else if (dist_1[v] || dist_2[v]) {
          labels[v] = label;
        }
      }
      st.first_scc_size = big;
      centers.resize(nnz);
      uint64_t nc = par_pack(
          nnz, [&](uint64_t i) { return !(labels[P[i]] & TOP_BIT); },
          [&](uint64_t i) { return P[i]; }, centers.data());
      centers.resize(nc);
    } else {
      centers.clear();
    }
    auto t2 = clk::now();
    st.t_first = sec(t1, t2);
    label_offset++;
    // multi-pivot rounds (:369-473)
    uint64_t step = 1, cur = 0, finished = 0;
    uint64_t n_rem = centers.size();
    uint64_t in_ne = 1, out_ne = 1;
    while (finished != n_rem) {
      uint64_t end = std::min(cur + step, n_rem);
      uint64_t vs = end - cur;
      finished += vs;
      step = (uint64_t)std::ceil(step * p.beta);
      st.rounds++;
      n_front = par_pack(
          vs, [&](uint64_t i) { return !(labels[centers[cur + i]] & TOP_BIT); },
          [&](uint64_t i) { return centers[cur + i]; }, front.data());
      cur = end;
      if (n_front == 0) continue;
      auto tm0 = clk::now();
      uint64_t base = std::max<uint64_t>(
          std::min<uint64_t>((uint64_t)std::ceil(0.3 * n_rem), 6000000),
          (uint64_t)(p.beta) * out_ne);
      table_forw.init(p.table_init ? p.table_init : 2 * base);
      multi_search_safe(labels, table_forw, true, p.local_scc);
      base = std::max<uint64_t>(
          std::min<uint64_t>((uint64_t)std::ceil(0.3 * n_rem), 6000000),
          (uint64_t)(p.beta) * in_ne);
      table_back.init(p.table_init ? p.table_init : 2 * base);
      multi_search_safe(labels, table_back, false, p.local_scc);
      st.t_multi += sec(tm0, clk::now());
      label_offset += n_front;
      // partition (:437-464)
      Table& sm = (table_forw.m <= table_back.m) ? table_forw : table_back;
      Table& lg = (table_forw.m > table_back.m) ? table_forw : table_back;
#pragma omp parallel for schedule(static)
      for (uint64_t s = 0; s < sm.m; s++) {
        uint64_t kv = sm.slots[s];
        if (kv == Table::EMPTY) continue;
        uint32_t v = Table::key(kv);
        uint64_t lab = Table::val(kv);
        write_max(&labels[v], lg.contains(v, (uint32_t)lab) ? (lab | TOP_BIT) : lab);
      }
#pragma omp parallel for schedule(static)
      for (uint64_t s = 0; s < lg.m; s++) {
        uint64_t kv = lg.slots[s];
        if (kv == Table::EMPTY) continue;
        write_max(&labels[Table::key(kv)], (uint64_t)Table::val(kv));
      }
      out_ne = table_forw.ne;
      in_ne = table_back.ne;
    }
    // finalize (:475-476)
#pragma omp parallel for schedule(static)
    for (uint64_t i = 0; i < n; i++) labels[i] = (labels[i] & VAL_MASK) - 1;
    auto t3 = clk::now();
    st.t_multi_safe = sec(t2, t3);
    st.t_total = sec(t0, t3);
  }
};

}  // namespace

// ---------------------------------------------------------------- C ABI
extern "C" {

struct psc_oracle_params {
  double beta;
  int32_t local_reach, local_scc;
  uint64_t tau, min_bag, dense_div, seed, table_init;
  int32_t threads;
};

struct psc_oracle_stats {
  double t_total, t_init, t_first, t_multi, t_rounds;
  uint64_t n_trimmed, first_scc_size, rounds, resizes, levels_fwd, levels_bwd,
      sub_rounds;
};

struct OracleHandle {
  GraphView g;
  SCC* scc;
};

void* psc_oracle_scc_create(uint64_t n, uint64_t m, const uint64_t* off,
                            const uint32_t* adj, const uint64_t* in_off,
                            const uint32_t* in_adj,
                            const psc_oracle_params* prm) {
  if (prm->threads > 0) omp_set_num_threads(prm->threads);
  auto* h = new OracleHandle;
  h->g.n = n;
  h->g.m = m;
  h->g.off[0] = off; h->g.adj[0] = adj;
  h->g.off[1] = in_off; h->g.adj[1] = in_adj;
  Params p{prm->beta, prm->local_reach, prm->local_scc, prm->tau ? prm->tau : 1,
           prm->min_bag ? prm->min_bag : 1024,
           prm->dense_div ? prm->dense_div : 20, prm->seed, prm->table_init,
           prm->threads};
  h->scc = new SCC(h->g, p, p.min_bag);
  return h;
}

int psc_oracle_scc_run(void* hv, uint64_t* labels, psc_oracle_stats* out) {
  auto* h = (OracleHandle*)hv;
  if (h->scc->p.threads > 0) omp_set_num_threads(h->scc->p.threads);
  h->scc->run(labels);
  if (out) {
    const Stats& s = h->scc->st;
    out->t_total = s.t_total; out->t_init = s.t_init; out->t_first = s.t_first;
    out->t_multi = s.t_multi; out->t_rounds = s.t_multi_safe;
    out->n_trimmed = s.n_trimmed; out->first_scc_size = s.first_scc_size;
    out->rounds = s.rounds; out->resizes = s.resizes;
    out->levels_fwd = s.reach_levels_fwd; out->levels_bwd = s.reach_levels_bwd;
    out->sub_rounds = s.sub_rounds;
  }
  return 0;
}

void psc_oracle_scc_destroy(void* hv) {
  auto* h = (OracleHandle*)hv;
  delete h->scc;
  delete h;
}

int psc_oracle_max_threads(void) { return omp_get_max_threads(); }

// single reach (src/reach.hpp:208-243) exposed for per-function tests
uint64_t psc_oracle_reach(uint64_t n, uint64_t m, const uint64_t* off,
                          const uint32_t* adj, const uint64_t* in_off,
                          const uint32_t* in_adj, uint32_t src, int dir,
                          int local, uint64_t tau, uint8_t* dst) {
  GraphView g{n, m, {off, in_off}, {adj, in_adj}};
  Reach R(g, tau ? tau : 1, 20);
  R.dir = dir;
  return R.reach(src, dst, local != 0);
}

// table hooks: insert a batch in parallel, report per-item "new" flags,
// contains and the values stored under one key (src/resizable_table.h).
void* psc_oracle_table_new(uint64_t m) {
  auto* t = new Table;
  t->init(m);
  return t;
}
void psc_oracle_table_free(void* t) { delete (Table*)t; }
int psc_oracle_table_insert(void* tv, uint64_t cnt, const uint32_t* k,
                            const uint32_t* v, uint8_t* fresh) {
  auto* t = (Table*)tv;
#pragma omp parallel for schedule(dynamic, 64)
  for (uint64_t i = 0; i < cnt; i++) fresh[i] = t->insert(k[i], v[i]);
  return t->full ? 1 : 0;
}
int psc_oracle_table_contains(void* tv, uint32_t k, uint32_t v) {
  return ((Table*)tv)->contains(k, v);
}
uint64_t psc_oracle_table_values(void* tv, uint32_t k, uint32_t* out,
                                 uint64_t cap) {
  uint64_t c = 0;
  ((Table*)tv)->for_each_value(k, [&](uint32_t x) {
    if (c < cap) out[c] = x;
    c++;
    return true;
  });
  return c;
}
uint64_t psc_oracle_table_size(void* tv) { return ((Table*)tv)->ne; }

// hash-bag hooks: insert a batch in parallel then pack (src/hash_bag.h)
uint64_t psc_oracle_bag_roundtrip(uint64_t n, uint64_t min_bag, uint64_t cnt,
                                  const uint32_t* items, uint32_t* out,
                                  int stamped) {
  HashBag b(n, min_bag);
  if (!stamped) {
#pragma omp parallel for schedule(dynamic, 64)
    for (uint64_t i = 0; i < cnt; i++) b.insert(items[i]);
    return b.pack(out);
  }
  std::vector<uint32_t> stamp(b.pool_size, 0);
#pragma omp parallel for schedule(dynamic, 64)
  for (uint64_t i = 0; i < cnt; i++) b.insert(items[i], stamp.data(), 7);
  return b.pack(out, stamp.data(), 7);
}

// the shared pivot permutation, for the product-vs-oracle pivot test
uint64_t psc_oracle_pivot_perm(uint64_t N, uint64_t seed, uint64_t i) {
  return PivotPerm(N, seed)(i);
}

}  // extern "C"
