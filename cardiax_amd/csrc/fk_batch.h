// Device helpers shared by the batched kernels (fk_batch.hip, and the batched
// rollout of fk_rollout_batch.hip): stimuli with a per-step activity mask and
// wave-uniform (scalar register) copies of descriptor values.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fk {

struct BatchStim {
  double start, duration, period;
  const float* field;
};

__device__ __forceinline__ bool active_dev(double t, const BatchStim& s) {
  if (!(t >= s.start)) return false;
  double r = fmod(s.start - t + 1.0, s.period);
  if (r != 0.0 && ((r < 0.0) != (s.period < 0.0))) r += s.period;
  return r < s.duration;
}

// Wave-uniform copies (scalar registers) of descriptor values.
__device__ __forceinline__ int uni(int x) { return __builtin_amdgcn_readfirstlane(x); }
__device__ __forceinline__ float uni(float x) {
  return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(x)));
}
template <class P>
__device__ __forceinline__ P* uni(P* p) {
  const uint64_t a = (uint64_t)p;
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)a);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(a >> 32));
  return (P*)(((uint64_t)hi << 32) | lo);
}

// The simulation's stimuli with this step's activity mask.
struct BatchStimView {
  const BatchStim* s;
  int n;
  uint32_t mask;
  __device__ int count() const { return n; }
  __device__ bool on(int m) const { return (mask >> m) & 1u; }
  __device__ const float* field(int m) const { return s[m].field; }
};

}  // namespace fk
