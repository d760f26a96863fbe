// Allocation-free SCC: psc_run_noalloc.
//
// For callers that cannot allocate device memory while they run (code under
// HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
// Every working buffer of the run -- bitmaps, frontiers, heavy lists, pivot
// order, label tables, logs, the remainder's compact copy, scan scratch --
// is carved from `scratch`, one fixed array per device, by a bump allocator
// that is emptied at the start of each call.  The labels are computed in
// place in the caller's output array.  The engine behind it is one per
// device, set up by the device's first call (hipGetSymbolAddress of the
// array, launch shapes); it runs on the caller's stream, stages its host
// reads in its own memory (no pinned allocation) and creates no event or
// stream.  Calls on one device share the array and the engine: they must
// not overlap (they race); each call rewrites everything it reads.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "engine.hpp"

namespace {

constexpr size_t SCRATCH_BYTES = PSC_NOALLOC_SCRATCH_BYTES;
static_assert(SCRATCH_BYTES % 256 == 0, "scratch is a whole number of 256-byte units");

static __device__ __align__(256) unsigned char scratch[SCRATCH_BYTES];

constexpr int MAX_DEVICES = 64;

struct Device {
  std::once_flag once;
  int rc = PSC_OK;
  std::string err;
  void *ptr = nullptr;          // device address of this device's `scratch`
  psc::Arena arena;
  psc::Engine *eng = nullptr;   // lives until the process ends
};

Device g_dev[MAX_DEVICES];

// Set-up of device `dev` (current): the array's address on this device and
// the engine's launch shapes.
void setup_device(Device &d) {
  void *p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p) {
    d.rc = PSC_ERR_HIP;
    d.err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
    return;
  }
  d.ptr = p;
  d.arena.base = static_cast<char *>(p);
  d.arena.cap = SCRATCH_BYTES;
  d.eng = new psc::Engine;
  d.eng->arena = &d.arena;
  d.eng->host_static = true;
  d.rc = d.eng->setup();
  if (d.rc) d.err = psc::g_last_error;
}

}  // namespace

extern "C" {

int psc_run_noalloc(const psc_graph *dev_g, const psc_params *p, uint64_t *d_labels_out,
                    void *stream, psc_stats *stats) {
  if (!dev_g) {
    psc::set_error("psc_run_noalloc: no graph");
    return PSC_ERR_INVALID;
  }
  if (dev_g->n > PSC_NOALLOC_MAX_N || dev_g->m > PSC_NOALLOC_MAX_M) {
    psc::set_error("psc_run_noalloc: n = " + std::to_string(dev_g->n) + ", m = " +
                   std::to_string(dev_g->m) + " exceed the limits of the allocation-free "
                   "entry point (n <= " + std::to_string(PSC_NOALLOC_MAX_N) +
                   ", m <= " + std::to_string(PSC_NOALLOC_MAX_M) + ")");
    return PSC_ERR_INVALID;
  }
  if ((dev_g->n && (!dev_g->offset || !dev_g->in_offset || !d_labels_out)) ||
      (dev_g->m && (!dev_g->edges || !dev_g->in_edges))) {
    psc::set_error("psc_run_noalloc: missing graph or label array");
    return PSC_ERR_INVALID;
  }
  psc_params q;
  psc_params_default(&q);
  if (p) q = *p;
  if (q.beta < 1.0) {
    psc::set_error("beta must be >= 1");
    return PSC_ERR_INVALID;
  }
  q.timing = 0;    // timing events and per-launch profiles would allocate
  q.capture = 0;   // so would instantiating captured graphs
  int dev = 0;
  hipError_t he = hipGetDevice(&dev);
  if (he != hipSuccess || dev < 0 || dev >= MAX_DEVICES) {
    psc::set_error(std::string("psc_run_noalloc: hipGetDevice: ") + hipGetErrorString(he));
    return PSC_ERR_HIP;
  }
  Device &d = g_dev[dev];
  std::call_once(d.once, setup_device, std::ref(d));
  if (d.rc) {
    psc::set_error(d.err);
    return d.rc;
  }
  psc::Engine &e = *d.eng;
  e.stream = static_cast<hipStream_t>(stream);
  e.labels_ext = reinterpret_cast<unsigned long long *>(d_labels_out);
  e.arena_reset();
  int rc = e.init(dev_g, false);
  psc_stats tmp;
  if (!rc) rc = e.run(q, e.labels_ext, stats ? stats : &tmp);
  // a failed run may leave work in flight that reads the scratch: it ends
  // before the call returns, so the next call can reuse the array
  if (rc) (void)hipStreamSynchronize(e.stream);
  return rc;
}

uint64_t psc_noalloc_scratch_bytes(void) { return SCRATCH_BYTES; }

uint64_t psc_noalloc_scratch_peak(void) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= MAX_DEVICES) return 0;
  return g_dev[dev].arena.peak;
}

}  // extern "C"
