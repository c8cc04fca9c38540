// fk_forward_euler_noalloc: the forward-Euler loop with its ping-pong state in
// a file-scope device array instead of a caller-supplied workspace, for callers
// that cannot allocate device memory (code under stream capture, tight serving
// loops, harnesses that forbid allocation). The entry point allocates nothing
// and calls no allocating library routine.
//
// The array `scratch` exists once per GPU for the life of the process. In host
// code its name is only a stand-in for the symbol, never a device address: the
// address is obtained per device with hipGetSymbolAddress the first time the
// entry point runs on that device, and only that address reaches a kernel.
#include <hip/hip_runtime.h>

#include <mutex>

#include "cardiax_amd.h"
#include "fk_host.h"

// The ping-pong state: v, w, u planes of up to FK_NOALLOC_MAX_CELLS cells.
static __device__ __align__(256) float scratch[3 * (size_t)FK_NOALLOC_MAX_CELLS];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_scratch[kMaxDevices] = {};  // device address of `scratch`, per device

// Device address of `scratch` on device `dev` (the current device), looked up
// once per device.
int32_t scratch_address(int dev, void** out) {
  if (dev < 0 || dev >= kMaxDevices)
    return fk::set_error(FK_ERR_UNSUPPORTED, "device %d is beyond the %d devices supported", dev,
                         kMaxDevices);
  std::lock_guard<std::mutex> lk(g_mu);
  if (g_scratch[dev] == nullptr) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || p == nullptr)
      return fk::set_error(FK_ERR_HIP, "hipGetSymbolAddress(scratch) on device %d failed: %s",
                           dev, hipGetErrorString(e));
    g_scratch[dev] = p;
  }
  *out = g_scratch[dev];
  return FK_OK;
}

}  // namespace

extern "C" int32_t fk_forward_euler_noalloc(const fk_params* params, const float* D,
                                            const float* v_in, const float* w_in,
                                            const float* u_in, int32_t H, int32_t W,
                                            const fk_stimulus* stimuli, int32_t n_stimuli,
                                            double t, double t_end, float dt, float dx,
                                            float* v_out, float* w_out, float* u_out,
                                            void* stream) {
  fk::ErrorScope es;
  if (!fk::check_grid(H, W)) return fk::last_code();
  if ((int64_t)H * (int64_t)W > (int64_t)FK_NOALLOC_MAX_CELLS)
    return fk::set_error(FK_ERR_ARG,
                         "grid %dx%d has %lld cells, above the %d of the fixed scratch",
                         H, W, (long long)H * W, FK_NOALLOC_MAX_CELLS);
  int dev = -1;
  HIP_CHECK(hipGetDevice(&dev));
  if (stream != nullptr) {
    int sdev = -1;
    HIP_CHECK(hipStreamGetDevice((hipStream_t)stream, &sdev));
    if (sdev != dev)
      return fk::set_error(FK_ERR_ARG, "stream belongs to device %d, the current device is %d",
                           sdev, dev);
  }
  void* work = nullptr;
  const int32_t rc = scratch_address(dev, &work);
  if (rc != FK_OK) return rc;
  // Every launch writes each cell of its target before the next one reads it,
  // so no part of the scratch needs clearing between calls.
  return fk::forward_loop(fk::current_forward_variant(), params, D, v_in, w_in, u_in, H, W,
                          stimuli, n_stimuli, t, t_end, dt, dx, v_out, w_out, u_out, work,
                          3 * (size_t)FK_NOALLOC_MAX_CELLS * sizeof(float), stream, true);
}
