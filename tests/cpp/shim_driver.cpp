// Compiles include/scc_gpu.hpp the way src/scc.cpp uses the SCC class
// (src/scc.cpp:54-67): construct once, one warm-up scc(), timer_reset,
// timed runs, breakdown_report.  The Graph here is a std::vector stand-in
// with the member names and element types of src/graph.hpp:35-43.
//   shim_driver <graph.bin> <labels.out> [repeat]
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

#include <vector>

struct Graph {
  size_t n = 0, m = 0;
  bool symmetric = false;
  std::vector<uint64_t> offset;
  std::vector<uint32_t> E;
  std::vector<uint64_t> in_offset;
  std::vector<uint32_t> in_E;
};
#define PSC_SHIM_HAVE_GRAPH
#include "scc_gpu.hpp"

int main(int argc, char **argv) {
  if (argc < 3) {
    fprintf(stderr, "usage: shim_driver graph.bin labels.out [repeat]\n");
    return 2;
  }
  psc_host_graph *hg = nullptr;
  if (psc_read_bin(argv[1], 1, &hg) != PSC_OK) {
    fprintf(stderr, "read: %s\n", psc_last_error());
    return 1;
  }
  Graph G;
  G.n = hg->g.n;
  G.m = hg->g.m;
  G.offset.assign(hg->g.offset, hg->g.offset + G.n + 1);
  G.E.assign(hg->g.edges, hg->g.edges + G.m);
  G.in_offset.assign(hg->g.in_offset, hg->g.in_offset + G.n + 1);
  G.in_E.assign(hg->g.in_edges, hg->g.in_edges + G.m);
  psc_host_graph_free(hg);
  const int repeat = argc > 3 ? atoi(argv[3]) : 1;
  std::vector<size_t> label(G.n);
  try {
    tau = 512;
    _min_bag_size_ = 1024;
    SCC S(G);
    S.scc(label, 1.5, true, true);
    S.timer_reset();
    for (int i = 0; i < repeat; ++i) S.scc(label, 1.5, true, true);
    S.breakdown_report(repeat);
  } catch (const std::exception &e) {
    fprintf(stderr, "%s\n", e.what());
    return 1;
  }
  FILE *f = fopen(argv[2], "wb");
  if (!f || fwrite(label.data(), 8, label.size(), f) != label.size()) return 1;
  fclose(f);
  return 0;
}
