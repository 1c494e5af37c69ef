// Device scratch of the allocation-free entry points (monte_mc_project_noalloc).
//
// Some callers cannot allocate device memory during a call (code under graph
// capture, tight serving loops, harnesses that forbid hipMalloc).  For them the
// transport launch keeps its descriptors, tables, work queue, counters and
// flight boxes in this file-scope device array instead of a context workspace.
// Each GPU the library runs on holds its own copy until the process ends; its
// address differs per device, so it is looked up once per device with that
// device current (noalloc_setup, called by monte_create) and only that address
// is used from then on.  Every call on a GPU shares that GPU's array: calls
// that overlap on one GPU race, and each call rewrites or clears everything it
// reads (descriptors are copied in, the queue and counters are reset by a
// kernel, the flight boxes are copied in before their reduction).
#include <mutex>

#include "internal.h"

namespace monte {

namespace {

static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_ptr[kMaxDevices] = {};

}  // namespace

int noalloc_setup(int device) {
  if (device < 0 || device >= kMaxDevices) return (int)hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_mu);
  if (g_ptr[device]) return 0;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return (int)e;
  if (!p) return (int)hipErrorInvalidSymbol;
  g_ptr[device] = p;
  return 0;
}

void* noalloc_scratch(int device) {
  if (device < 0 || device >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lock(g_mu);
  return g_ptr[device];
}

}  // namespace monte
