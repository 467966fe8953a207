// lb_adjoint_noalloc: the discrete adjoint without device allocation.
//
// Its workspace -- the work-queue counter, the recomputed attempts' states
// and stage derivatives, and their coefficient rows -- lives in `scratch`, a
// fixed device array of LB_NOALLOC_SCRATCH_BYTES that every GPU the library
// runs on holds until the process ends.  Host code never passes the array's
// name to a kernel: on the first call on a device, with that device current,
// its device address is looked up once and kept per device.  Calls on one GPU
// share the array: the counter is cleared by a fill kernel before each
// launch, and calls that overlap on one GPU race.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "lb_internal.h"

namespace {

static __device__ __align__(256) double scratch[LB_NOALLOC_SCRATCH_BYTES / sizeof(double)];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_ptr[kMaxDevices] = {};

}  // namespace

extern "C" int lb_adjoint_noalloc(const lb_plan* plan, const lb_solve_args* args, void* stream) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess || dev < 0 || dev >= kMaxDevices) {
    return lb_set_error(LB_ERR_DEVICE, std::string("lb_adjoint_noalloc: no usable current device: ") +
                                           hipGetErrorString(e));
  }
  void* ptr = nullptr;
  {
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_ptr[dev]) {
      // once per device, with that device current: the array's device address
      e = hipGetSymbolAddress(&g_ptr[dev], HIP_SYMBOL(scratch));
      if (e != hipSuccess) {
        g_ptr[dev] = nullptr;
        return lb_set_error(LB_ERR_DEVICE, std::string("lb_adjoint_noalloc: hipGetSymbolAddress: ") +
                                               hipGetErrorString(e));
      }
    }
    ptr = g_ptr[dev];
  }
  return lb_adjoint_launch(plan, args, stream, ptr, LB_NOALLOC_SCRATCH_BYTES);
}
