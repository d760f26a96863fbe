/*
 * scc_gpu.hpp -- the SCC class of the reference (src/scc.hpp:38-115) backed
 * by the MI355X path through libpsc.so.
 *
 * In src/scc.cpp, line 1 `#include "scc.hpp"` becomes
 * `#include "scc_gpu.hpp"`; the rest of scc.cpp (flags, the warm-up and
 * timed runs, breakdown_report, SCC_status) compiles unchanged.  This header
 * includes the reference's src/graph.hpp (Graph, read_graph, parlay's
 * sequence) as scc.hpp does (src/scc.hpp:14), and defines the two globals
 * that src/reach.hpp:19 and src/scc.hpp:24 define (src/scc.cpp sets them,
 * src/scc.cpp:45-46).
 *
 * A program with its own Graph type (members n, m, offset, E, in_offset,
 * in_E holding uint64 offsets and uint32 ids, as src/graph.hpp:35-43)
 * defines PSC_SHIM_HAVE_GRAPH before including this header instead.
 */
#ifndef PSC_SCC_GPU_HPP
#define PSC_SCC_GPU_HPP

#ifndef PSC_SHIM_HAVE_GRAPH
#include "graph.hpp"
#endif

#include <stdint.h>

#include <iostream>
#include <stdexcept>
#include <string>

#include "psc.h"

inline size_t tau = 512;               // src/reach.hpp:19
inline size_t _min_bag_size_ = 1024;   // src/scc.hpp:24

class SCC {
  psc_handle *h_ = nullptr;
  size_t n_ = 0;
  // breakdown_report's timers (src/scc.hpp:94-114), seconds summed since
  // timer_reset
  double init_ = 0, first_ = 0, multi_ = 0, resize_ = 0;

  template <class S>
  static auto ptr(S &s) -> decltype(&s[0]) {
    return s.size() ? &s[0] : nullptr;
  }
  static void check(const char *what, int rc) {
    if (rc != PSC_OK) throw std::runtime_error(std::string(what) + ": " + psc_last_error());
  }

 public:
  SCC() = delete;
  SCC(const SCC &) = delete;
  SCC &operator=(const SCC &) = delete;
  // SCC::SCC (src/scc.hpp:81-92): both CSRs go to the device once
  template <class GraphT>
  explicit SCC(GraphT &G) : n_(G.n) {
    static_assert(sizeof(*ptr(G.offset)) == 8 && sizeof(*ptr(G.E)) == 4,
                  "offsets must be uint64 and ids uint32 (src/utilities.h:12-27)");
    psc_graph g{(uint64_t)G.n,
                (uint64_t)G.m,
                reinterpret_cast<const uint64_t *>(ptr(G.offset)),
                reinterpret_cast<const uint32_t *>(ptr(G.E)),
                reinterpret_cast<const uint64_t *>(ptr(G.in_offset)),
                reinterpret_cast<const uint32_t *>(ptr(G.in_E))};
    check("psc_create", psc_create(&g, &h_));
  }
  ~SCC() { psc_destroy(h_); }

  // SCC::scc (src/scc.hpp:277-479): labels[i] = (label & 0xffffffff) - 1
  template <class Labels>
  void scc(Labels &labels, double beta, bool local_reach, bool local_scc) {
    static_assert(sizeof(*ptr(labels)) == 8, "labels are size_t (src/scc.cpp:54)");
    if ((size_t)labels.size() < n_) throw std::runtime_error("labels shorter than n");
    psc_params p;
    psc_params_default(&p);
    p.beta = beta;
    p.local_reach = local_reach;
    p.local_scc = local_scc;
    p.tau = tau;
    p.min_bag = _min_bag_size_;
    psc_stats st;
    check("psc_run", psc_run(h_, &p, reinterpret_cast<uint64_t *>(ptr(labels)), &st));
    init_ += st.ms_init / 1e3;
    first_ += st.ms_first / 1e3;
    multi_ += st.ms_multi / 1e3;
    resize_ += st.ms_resize / 1e3;
  }

  void timer_reset() { init_ = first_ = multi_ = resize_ = 0; }
  void breakdown_report(int repeat) {   // the lines of src/scc.hpp:106-114
    const double r = repeat > 0 ? repeat : 1;
    std::cout << "average init_time " << init_ / r << std::endl;
    std::cout << "average first_round_time " << first_ / r << std::endl;
    std::cout << "average multi_search time " << multi_ / r << std::endl;
    std::cout << "average hash table resize time " << resize_ / r << std::endl;
  }
};

#endif  // PSC_SCC_GPU_HPP
