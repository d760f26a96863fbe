// `scc` command line: the counterpart of src/scc.cpp, driving the MI355X
// path through the C ABI.  Same flags and the same stdout lines as the
// reference (src/scc.cpp:31-72) so its log scrapers
// (scripts/data_collection.py) keep working:
//   ./scc <graph.bin> [-large] [-local_reach] [-local_scc] [-tau N]
//         [-beta B] [-lambda N] [-t N] [-status]
// plus -seed S, -dense_div D, -table_init M, -canonical, -out FILE
// (labels as raw uint64, n entries).
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/time.h>

#include <string>
#include <vector>

#include "psc.h"

namespace {

struct CommandLine {   // src/parseCommandLine.hpp:6-38
  int argc;
  char **argv;
  bool has(const char *o) const {
    for (int i = 1; i < argc; i++)
      if (!strcmp(argv[i], o)) return true;
    return false;
  }
  const char *value(const char *o) const {
    for (int i = 1; i + 1 < argc; i++)
      if (!strcmp(argv[i], o)) return argv[i + 1];
    return nullptr;
  }
  long geti(const char *o, long d) const {
    const char *v = value(o);
    return v ? atol(v) : d;
  }
  double getd(const char *o, double d) const {
    const char *v = value(o);
    return v ? atof(v) : d;
  }
};

double now_s() {
  timeval t;
  gettimeofday(&t, nullptr);
  return t.tv_sec + t.tv_usec / 1e6;
}

int fail(const char *what, int rc) {
  fprintf(stderr, "%s failed (%d): %s\n", what, rc, psc_last_error());
  return 1;
}

}  // namespace

int main(int argc, char **argv) {
  CommandLine P{argc, argv};
  if (argc < 2) {
    printf("usage: ./scc fileName\n");
    return 0;
  }
  const char *fileName = argv[1];
  psc_params prm;
  psc_params_default(&prm);
  prm.beta = P.getd("-beta", 1.5);
  printf("--beta %g--\n", prm.beta);
  // the graph goes to device memory; a one-CSR file is transposed there
  psc_dev_graph *dg = nullptr;
  int rc = psc_read_bin_device(fileName, P.has("-large") ? 1 : 0, &dg);
  if (rc) return fail("read graph", rc);
  const psc_graph *g = &dg->g;
  prm.local_reach = P.has("-local_reach");
  prm.local_scc = P.has("-local_scc");
  prm.min_bag = (uint64_t)P.geti("-lambda", 1 << 10);
  prm.tau = (uint64_t)P.geti("-tau", 512);
  prm.seed = (uint64_t)P.geti("-seed", 1);
  prm.dense_div = (uint64_t)P.geti("-dense_div", 20);
  prm.table_init = (uint64_t)P.geti("-table_init", 0);
  prm.canonical = P.has("-canonical");
  printf("tau %llu\n", (unsigned long long)prm.tau);
  printf("input graph: %s\n", fileName);
  printf("number of vertices: %llu, number of edges: %llu\n",
         (unsigned long long)g->n, (unsigned long long)g->m);
  printf("enable VGC1: %d, enable VGC_multi: %d\n", prm.local_reach, prm.local_scc);
  fflush(stdout);

  psc_handle *h = nullptr;
  rc = psc_create_device(g, &h);
  if (rc) return fail("psc_create_device", rc);
  std::vector<uint64_t> label(g->n);
  const int repeat = (int)P.geti("-t", 10);
  psc_stats st;
  rc = psc_run(h, &prm, label.data(), &st);   // warm-up, src/scc.cpp:58
  if (rc) return fail("psc_run", rc);
  double total = 0, init = 0, first = 0, multi = 0, resize = 0;
  for (int i = 0; i < repeat; i++) {
    const double t0 = now_s();
    rc = psc_run(h, &prm, label.data(), &st);
    const double c = now_s() - t0;
    if (rc) return fail("psc_run", rc);
    total += c;
    init += st.ms_init / 1e3;
    first += st.ms_first / 1e3;
    multi += st.ms_multi / 1e3;
    resize += st.ms_resize / 1e3;
    printf("scc cost: %g\n", c);
  }
  const int rep = repeat > 0 ? repeat : 1;
  printf("average cost %g\n", total / rep);
  printf("average init_time %g\n", init / rep);       // src/scc.hpp:106-114
  printf("average first_round_time %g\n", first / rep);
  printf("average multi_search time %g\n", multi / rep);
  printf("average hash table resize time %g\n", resize / rep);
  if (P.has("-status")) {   // src/scc.cpp:13-29
    // labels are (internal & 0xffffffff) - 1 in [0, n] (or minimum vertex
    // ids with -canonical), so they index n + 1 counters directly; the
    // reference's 31-bit mask (src/scc.cpp:17) would fold labels of graphs
    // with n >= 2^31 together
    std::vector<uint32_t> flag(g->n + 1, 0);
    for (uint64_t i = 0; i < g->n; i++) {
      if (label[i] > g->n) {
        fprintf(stderr, "label %llu of vertex %llu outside [0, n]\n",
                (unsigned long long)label[i], (unsigned long long)i);
        return 1;
      }
      flag[label[i]]++;
    }
    uint64_t nscc = 0, big = 0;
    for (uint32_t f : flag) {
      nscc += f != 0;
      if (f > big) big = f;
    }
    printf("n_scc = %llu\n", (unsigned long long)nscc);
    printf("Largest StronglyConnectedComponents has %llu\n", (unsigned long long)big);
  }
  if (const char *out = P.value("-out")) {
    FILE *f = fopen(out, "wb");
    if (!f || fwrite(label.data(), 8, label.size(), f) != label.size()) {
      fprintf(stderr, "cannot write %s\n", out);
      return 1;
    }
    fclose(f);
  }
  psc_destroy(h);
  psc_dev_graph_free(dg);
  return 0;
}
