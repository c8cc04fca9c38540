// Helpers shared by the batched kernels (fk_batch.hip, and the batched
// rollout of fk_rollout_batch.hip): stimuli with a per-step activity mask,
// wave-uniform (scalar register) copies of descriptor values, and the XCD
// order of a work list.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <vector>

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

// Work-list order for the chip's placement of workgroup w on XCD w % 8 (each
// XCD has its own L2): `runs` are groups of consecutive workgroups that share
// halo rows (a column of tiles); each run goes whole to the XCD with the least
// work so far, and the XCD lists are interleaved, short ones padded with
// `pad` items that exit at once. `per_wg` items make one workgroup.
constexpr int kXcds = 8;

template <class T>
inline std::vector<T> xcd_order(const std::vector<std::vector<T>>& runs, int per_wg, T pad) {
  std::vector<std::vector<T>> lists(kXcds);
  for (const auto& r : runs) {
    size_t best = 0;
    for (int x = 1; x < kXcds; ++x)
      if (lists[x].size() < lists[best].size()) best = x;
    lists[best].insert(lists[best].end(), r.begin(), r.end());
  }
  size_t m = 0;
  for (auto& l : lists) {
    while (l.size() % per_wg) l.push_back(pad);
    m = std::max(m, l.size() / per_wg);
  }
  std::vector<T> out;
  out.reserve(m * kXcds * per_wg);
  for (size_t j = 0; j < m; ++j)
    for (int x = 0; x < kXcds; ++x)
      for (int k = 0; k < per_wg; ++k)
        out.push_back(j * per_wg + k < lists[x].size() ? lists[x][j * per_wg + k] : pad);
  return out;
}

}  // namespace fk
