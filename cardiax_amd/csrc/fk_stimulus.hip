// Stimulus fields built on the device (cardiax/stimulus.py:31-140): the
// box masks of `rectangular` and `linear`, and the resampling `triangular`
// applies to a linear stripe.
//
// triangular calls scipy.ndimage.rotate(field, angle, mode="nearest",
// prefilter=False, reshape=False) (order 3), which builds an affine map
// (input coordinate = M o + offset for output index o) and hands it to
// scipy.ndimage.affine_transform. The map is built by the caller as scipy
// builds it; fk_affine_transform then reproduces affine_transform's
// arithmetic for order 3, no prefilter and mode "nearest", in double:
//   * x = (0 + o0 m[0] + o1 m[1]) + offset, per input axis, without fused ops;
//   * taps floor(x)-1 .. floor(x)+2 per axis; with t = x - floor(x) and
//     z = 1 - t the cubic B-spline weights are z^3/6,
//     (t^2 (t-2) 3 + 4)/6, (z^2 (z-2) 3 + 4)/6 and 1 minus those three;
//   * taps outside the grid take the nearest edge sample;
//   * the sum runs over row taps, then column taps, each term
//     (value w_row) w_col; the result is rounded once to float32.
// Measured bit-identical to scipy's output over 7.25 M cells of stripes and
// random fields at 18 angles and 9 shapes.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "cardiax_amd.h"
#include "fk_host.h"

namespace fk {

__global__ void __launch_bounds__(256)
k_stimulus_box(float* __restrict__ field, int H, int W, int r0, int r1, int c0, int c1,
               float modulus) {
  const int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (idx >= (int64_t)H * W) return;
  const int r = (int)(idx / W), c = (int)(idx % W);
  field[idx] = (r >= r0 && r < r1 && c >= c0 && c < c1) ? modulus : 0.0f;
}

// The last weight is 1 minus the other three, so at t = 0 it is 8.3e-17
// rather than 0 and a cell one row past a sharp edge reads a tiny nonzero
// value, which the solver's stimulus rule (replace where the field is
// nonzero) sees.
__device__ __forceinline__ void bspline3(double t, double w[4]) {
#pragma clang fp contract(off)
  const double z = 1.0 - t;
  w[1] = (t * t * (t - 2.0) * 3.0 + 4.0) / 6.0;
  w[2] = (z * z * (z - 2.0) * 3.0 + 4.0) / 6.0;
  w[0] = z * z * z / 6.0;
  w[3] = 1.0 - w[0] - w[1] - w[2];
}

struct AffineArgs {
  double m00, m01, m10, m11, off0, off1;
};

__global__ void __launch_bounds__(256)
k_affine_spline3(const float* __restrict__ src, int H, int W, AffineArgs a,
                 float* __restrict__ dst) {
#pragma clang fp contract(off)  // the reference library's rounding: no fused operations
  const int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (idx >= (int64_t)H * W) return;
  const double o0 = (double)(idx / W), o1 = (double)(idx % W);
  double x0 = 0.0, x1 = 0.0;
  x0 += o0 * a.m00;
  x0 += o1 * a.m01;
  x0 = x0 + a.off0;
  x1 += o0 * a.m10;
  x1 += o1 * a.m11;
  x1 = x1 + a.off1;
  const double f0 = floor(x0), f1 = floor(x1);
  double w0[4], w1[4];
  bspline3(x0 - f0, w0);
  bspline3(x1 - f1, w1);
  // Taps are clamped to the grid; bounding the floor first keeps the int
  // conversion defined for any coordinate without changing a clamped tap.
  const int i0 = (int)fmin(fmax(f0, -8.0), (double)H + 8.0) - 1;
  const int j0 = (int)fmin(fmax(f1, -8.0), (double)W + 8.0) - 1;
  double acc = 0.0;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int r = min(max(i0 + i, 0), H - 1);
    const float* row = src + (int64_t)r * W;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int q = min(max(j0 + j, 0), W - 1);
      acc += ((double)row[q] * w0[i]) * w1[j];
    }
  }
  dst[idx] = (float)acc;
}

}  // namespace fk

using namespace fk;

extern "C" {

int32_t fk_stimulus_box(float* field, int32_t H, int32_t W, int32_t r0, int32_t r1, int32_t c0,
                        int32_t c1, float modulus, void* stream) {
  ErrorScope es;
  if (H < 1 || W < 1) return set_error(FK_ERR_ARG, "fk_stimulus_box: bad grid %d x %d", H, W);
  if (!field) return set_error(FK_ERR_ARG, "fk_stimulus_box: null field");
  if (r0 < 0 || r1 > H || c0 < 0 || c1 > W)
    return set_error(FK_ERR_ARG, "fk_stimulus_box: box [%d, %d) x [%d, %d) outside %d x %d", r0,
                     r1, c0, c1, H, W);
  const int64_t n = (int64_t)H * W;
  k_stimulus_box<<<(unsigned)((n + 255) / 256), 256, 0, (hipStream_t)stream>>>(
      field, H, W, r0, r1, c0, c1, modulus);
  HIP_CHECK(hipGetLastError());
  return FK_OK;
}

int32_t fk_affine_transform(const float* src, int32_t H, int32_t W, const double* matrix,
                            const double* offset, float* dst, void* stream) {
  ErrorScope es;
  if (H < 1 || W < 1) return set_error(FK_ERR_ARG, "fk_affine_transform: bad grid %d x %d", H, W);
  if (!src || !dst || !matrix || !offset)
    return set_error(FK_ERR_ARG, "fk_affine_transform: null pointer");
  for (int k = 0; k < 4; ++k)
    if (!std::isfinite(matrix[k]) || (k < 2 && !std::isfinite(offset[k])))
      return set_error(FK_ERR_ARG, "fk_affine_transform: non-finite map");
  const int64_t n = (int64_t)H * W;
  if (dst < src + n && src < dst + n)
    return set_error(FK_ERR_ARG, "fk_affine_transform: dst overlaps src");
  const AffineArgs a{matrix[0], matrix[1], matrix[2], matrix[3], offset[0], offset[1]};
  k_affine_spline3<<<(unsigned)((n + 255) / 256), 256, 0, (hipStream_t)stream>>>(src, H, W, a,
                                                                                 dst);
  HIP_CHECK(hipGetLastError());
  return FK_OK;
}

}  // extern "C"
