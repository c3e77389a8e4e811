// casl_reinit_noalloc: reinitialization for callers that cannot allocate
// device memory (code under HIP graph capture, serving loops, harnesses that
// forbid hipMalloc).  The call's buffers come from one fixed device array per
// GPU instead of the stencil's scratch.
#include <hip/hip_runtime.h>

#include <map>
#include <mutex>

#include "reinit_device.hpp"

// CASL_NOALLOC_SCRATCH_BYTES of device memory, loaded with the code object on
// every GPU the library runs on and held until the process ends.  On the host
// this name is only a stand-in for the symbol, not a device address.
static __device__ __align__(256) unsigned char scratch[CASL_NOALLOC_SCRATCH_BYTES];

namespace {

std::mutex g_mu;
std::map<int, unsigned char *> g_scratch;  // device -> the array's device address there

// The array's address on the current device, looked up once per device with
// that device current (each GPU holds its own copy at its own address).
unsigned char *scratch_here(int device) {
  std::lock_guard<std::mutex> g(g_mu);
  auto it = g_scratch.find(device);
  if (it != g_scratch.end()) return it->second;
  void *ptr = nullptr;
  HIP_CHECK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
  if (!ptr) throw casl::hip_error("hipGetSymbolAddress gave no address for the scratch array");
  g_scratch[device] = static_cast<unsigned char *>(ptr);
  return static_cast<unsigned char *>(ptr);
}

}  // namespace

extern "C" int casl_reinit_noalloc(casl_stencil_t *st, double *phi, int order_space, int order_pseudotime,
                                   int n_iter, double limit_high, double limit_low, int precision, void *stream) {
  return casl::guarded([&]() {
    if (!st || !phi) throw std::invalid_argument("null argument");
    int device = 0;
    HIP_CHECK(hipGetDevice(&device));
    if (device != st->device) throw std::invalid_argument("the stencil was created on another device");
    unsigned char *base = scratch_here(device);
    casl::reinit_in_buffer(st, phi, order_space, order_pseudotime, n_iter, limit_high, limit_low, precision,
                           static_cast<hipStream_t>(stream), base, CASL_NOALLOC_SCRATCH_BYTES);
  });
}
