// Checkpoint output of the dataset generator: the on-device resize that
// cardiax/io.py:118-124 (imresize) applies to every saved state, D-map and
// stimulus field before it is written (deepx/generate.py:167-207).
//
// imresize is jax.image.resize(a, a.shape[:-2] + (h, w), "bilinear"), whose
// default antialias=True makes it a separable resampling with a triangle
// kernel widened by the downsampling factor (JAX's scale_and_translate,
// translation 0). Along an axis of n_in -> n_out samples, output o sits at
// input coordinate s = (o + 1/2) n_in / n_out - 1/2 and takes the inputs j
// with weights max(0, 1 - |s - j| / k), k = max(n_in / n_out, 1), normalised
// to sum 1 (all weights 0 when their sum is <= 1000 float32 eps). An axis
// whose size does not change reduces to the identity.
//
// Two passes, rows then columns. Weights and sums are in double and the row
// pass keeps its result in double, so the output is the exact resampling
// rounded once to float32.
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>
#include <cstdint>

#include "cardiax_amd.h"
#include "fk_host.h"

namespace fk {

struct Taps {
  int lo, hi;       // inclusive input range with nonzero weight
  double s, inv_k;  // sample position, 1 / kernel width
  double inv_sum;   // 1 / sum of weights, or 0 when the sum is negligible
};

__device__ __forceinline__ Taps taps_of(int o, int n_in, int n_out) {
  Taps t;
  const double scale_inv = (double)n_in / (double)n_out;
  const double k = scale_inv > 1.0 ? scale_inv : 1.0;
  t.s = ((double)o + 0.5) * scale_inv - 0.5;
  t.inv_k = 1.0 / k;
  t.lo = max(0, (int)floor(t.s - k) + 1);
  t.hi = min(n_in - 1, (int)ceil(t.s + k) - 1);
  double sum = 0.0;
  for (int j = t.lo; j <= t.hi; ++j) sum += fmax(0.0, 1.0 - fabs(t.s - (double)j) * t.inv_k);
  t.inv_sum = sum > 1000.0 * (double)FLT_EPSILON ? 1.0 / sum : 0.0;
  return t;
}

__device__ __forceinline__ double weight(const Taps& t, int j) {
  return fmax(0.0, 1.0 - fabs(t.s - (double)j) * t.inv_k) * t.inv_sum;
}

// Row pass: src (n, H, W) float -> tmp (n, h, W) double. One thread per
// output element, W fastest (coalesced row reads).
__global__ void __launch_bounds__(256)
k_resize_rows(const float* __restrict__ src, int H, int W, int h, int64_t total,
              double* __restrict__ tmp) {
  const int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (idx >= total) return;
  const int q = (int)(idx % W);
  const int64_t ni = idx / W;
  const int i = (int)(ni % h);
  const int64_t n = ni / h;
  const Taps t = taps_of(i, H, h);
  const float* col = src + n * (int64_t)H * W + q;
  double acc = 0.0;
  for (int r = t.lo; r <= t.hi; ++r) acc += weight(t, r) * (double)col[(int64_t)r * W];
  tmp[idx] = acc;
}

// Column pass: tmp (n, h, W) double -> dst (n, h, w) float.
__global__ void __launch_bounds__(256)
k_resize_cols(const double* __restrict__ tmp, int W, int w, int64_t total,
              float* __restrict__ dst) {
  const int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (idx >= total) return;
  const int j = (int)(idx % w);
  const int64_t row = idx / w;
  const Taps t = taps_of(j, W, w);
  const double* rp = tmp + row * (int64_t)W;
  double acc = 0.0;
  for (int q = t.lo; q <= t.hi; ++q) acc += weight(t, q) * rp[q];
  dst[idx] = (float)acc;
}

}  // namespace fk

using namespace fk;

extern "C" {

size_t fk_resize_workspace_bytes(int32_t n_images, int32_t H, int32_t W, int32_t h, int32_t w) {
  (void)H;
  (void)w;
  if (n_images <= 0 || W <= 0 || h <= 0) return 0;
  return (size_t)n_images * (size_t)h * (size_t)W * sizeof(double);
}

int32_t fk_resize(const float* src, int32_t n_images, int32_t H, int32_t W, float* dst,
                  int32_t h, int32_t w, void* work, size_t work_bytes, void* stream) {
  ErrorScope es;
  if (n_images < 0 || H < 1 || W < 1 || h < 1 || w < 1)
    return set_error(FK_ERR_ARG, "fk_resize: bad sizes (n %d, %d x %d -> %d x %d)", n_images, H,
                     W, h, w);
  if (n_images == 0) return FK_OK;
  if (!src || !dst || !work) return set_error(FK_ERR_ARG, "fk_resize: null pointer");
  const size_t need = fk_resize_workspace_bytes(n_images, H, W, h, w);
  if (work_bytes < need)
    return set_error(FK_ERR_ARG, "fk_resize: workspace too small: need %zu bytes, got %zu",
                     need, work_bytes);
  const char* s0 = (const char*)src;
  const char* s1 = s0 + (size_t)n_images * H * W * sizeof(float);
  const char* d0 = (const char*)dst;
  const char* d1 = d0 + (size_t)n_images * h * w * sizeof(float);
  if (d0 < s1 && s0 < d1) return set_error(FK_ERR_ARG, "fk_resize: dst overlaps src");
  hipStream_t s = (hipStream_t)stream;
  double* tmp = (double*)work;
  const int64_t t1 = (int64_t)n_images * h * W, t2 = (int64_t)n_images * h * w;
  k_resize_rows<<<(unsigned)((t1 + 255) / 256), 256, 0, s>>>(src, H, W, h, t1, tmp);
  k_resize_cols<<<(unsigned)((t2 + 255) / 256), 256, 0, s>>>(tmp, W, w, t2, dst);
  HIP_CHECK(hipGetLastError());
  return FK_OK;
}

}  // extern "C"
