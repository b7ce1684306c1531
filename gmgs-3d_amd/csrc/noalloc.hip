// gmgs_atx_device_noalloc: the level operator AtX on device-resident vectors
// for callers that cannot allocate device memory (code under HIP graph
// capture, tight serving loops, harnesses that forbid hipMalloc).  It runs
// the kernels of gmgs_atx; the two padded level vectors that gmgs_atx
// allocates per call live here in a file-scope device array instead.
#include "../../include/gmgs.h"

#include <mutex>

#include "common.h"

namespace gmgs {
int atx_device_scratch(gmgs_ctx* c, int level, const double* x, double* y,
                       void* (*scratch)(int device, size_t* bytes));
}

namespace {

// Each GPU the library runs on holds its own copy of this array until the
// process ends; every call on that GPU shares it.
constexpr size_t kScratchBytes = size_t(GMGS_NOALLOC_SCRATCH_MIB) << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_addr[kMaxDevices] = {};

// The array's device address on `device` (current when called), looked up
// once per device.  The array's name is a host-side stand-in: only this
// address may be handed to kernels or copies.
void* scratch_on(int device, size_t* bytes) {
  *bytes = 0;
  if (device < 0 || device >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_addr[device]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return nullptr;
    g_addr[device] = p;
  }
  *bytes = kScratchBytes;
  return g_addr[device];
}

}  // namespace

extern "C" int gmgs_atx_device_noalloc(gmgs_ctx* c, int level, const double* x_dev, double* y_dev) {
  return gmgs::atx_device_scratch(c, level, x_dev, y_dev, scratch_on);
}
