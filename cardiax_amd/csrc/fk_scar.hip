// The device half of the reference's random scar D-map generator
// (deepx/utils_scars.py:246-426, called by deepx/generate.py:79-83): the
// rasterisation of the closed spline contours into a mask and the Gaussian
// blur that tapers its edges. The contours themselves (a few hundred points)
// are built on the host, as the reference builds them with scipy.
//
// Rasterisation (drawpolyintoemptycanvas, :79-97, via skimage.draw.polygon):
// a pixel (r, c) is inside a polygon when it lies in the polygon's bounding
// box [int(max(0, min r)), ceil(max r)] x [int(max(0, min c)), ceil(max c)]
// and the crossing-number test at its centre (integer coordinates) holds:
// an edge (i, j) counts when (r_i <= r < r_j or r_j <= r < r_i) and
// c < (c_j - c_i)(r - r_i) / (r_j - r_i) + c_i, in double precision. The
// blobs of one scar are OR'ed (MakeAndSumCompositeBlob, :377-381: summed and
// clipped to 1); pixels outside the image are dropped, as JAX's scatter drops
// them.
//
// Blur (blur_scar, :402-406): scipy.signal.convolve2d(scar, kernel, "same"):
// out[i, j] = sum_{p, q} k[p, q] in[i + (kh-1)/2 - p, j + (kw-1)/2 - q]
// (integer halves), zero outside the image; products and sums in double.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "cardiax_amd.h"
#include "fk_host.h"
#include "fk_stream.h"

namespace fk {

struct PolyBox {
  int r0, r1, c0, c1;  // inclusive bounding box of tested pixels
  int begin, end;      // vertex range
};

// Crossing-number test at (x, y) = (column, row) of the closed polygon with
// vertices (cp[k], rp[k]), k in [b, e).
__device__ __forceinline__ bool inside_polygon(const double* __restrict__ rp,
                                               const double* __restrict__ cp, int b, int e,
                                               double y, double x) {
#pragma clang fp contract(off)
  bool in = false;
  int j = e - 1;
  for (int i = b; i < e; ++i) {
    const double yi = rp[i], yj = rp[j], xi = cp[i], xj = cp[j];
    if (((yi <= y && y < yj) || (yj <= y && y < yi)) && (x < (xj - xi) * (y - yi) / (yj - yi) + xi))
      in = !in;
    j = i;
  }
  return in;
}

__global__ void __launch_bounds__(256)
k_raster_polygons(float* __restrict__ mask, int H, int W, const double* __restrict__ rp,
                  const double* __restrict__ cp, const PolyBox* __restrict__ boxes, int n_polys) {
  const int c = blockIdx.x * 64 + (threadIdx.x & 63);
  const int r = blockIdx.y * 4 + (threadIdx.x >> 6);
  if (r >= H || c >= W) return;
  bool on = false;
  for (int p = 0; p < n_polys && !on; ++p) {
    const PolyBox bx = boxes[p];
    if (r < bx.r0 || r > bx.r1 || c < bx.c0 || c > bx.c1) continue;
    on = inside_polygon(rp, cp, bx.begin, bx.end, (double)r, (double)c);
  }
  mask[(size_t)r * W + c] = on ? 1.0f : 0.0f;
}

// R outputs down a column per thread. With d = s - (i0 + oi) the offset of
// input row s from the thread's first output row, output r takes tap row
// p = r - d. The thread walks d over [-(kh-1), R-1] in chunks of R rows and,
// within a chunk, every tap column q: R input values (one load each, reused
// by up to R outputs; the chunk's rows stay in L1 across q) and the chunk's
// 2R - 1 taps of column q, contiguous and wave-uniform in the padded,
// transposed copy of the kernel (k_pad_kernel: zeros outside [0, kh), so no
// test per tap). The inputs come from a zero-padded copy of the image
// (k_pad_image) wide enough for every row and column a thread reaches, so
// the loads need no clamp and no select. Per output the sum runs over
// chunks, then q, then d, in double; rounded once.
constexpr int kConvRows = 8;
constexpr int kConvPadFront = 2 * kConvRows - 2;

__host__ __device__ constexpr int conv_padded_rows(int kh) { return kh + 3 * kConvRows - 3; }

// The zero-padded image: PT rows above and below, PL columns left and PR
// right of the H x W image.
struct ConvPad {
  int PT, PL, Hp, Wp;
};

__host__ __device__ inline ConvPad conv_pad(int H, int W, int kh, int kw) {
  ConvPad p;
  p.PT = kh + 2 * kConvRows;
  p.PL = kw;
  p.Hp = H + 2 * p.PT;
  p.Wp = W + 2 * kw + 64;
  return p;
}

__global__ void __launch_bounds__(256)
k_pad_kernel(const double* __restrict__ k, int kh, int kw, double* __restrict__ kt) {
  const int kp = conv_padded_rows(kh);
  const int idx = blockIdx.x * 256 + threadIdx.x;
  if (idx >= kw * kp) return;
  const int q = idx / kp, p = idx % kp - kConvPadFront;
  kt[idx] = (p >= 0 && p < kh) ? k[(size_t)p * kw + q] : 0.0;
}

__global__ void __launch_bounds__(256)
k_pad_image(const float* __restrict__ src, int H, int W, ConvPad pd, float* __restrict__ pad) {
  const size_t n = (size_t)pd.Hp * pd.Wp;
  for (size_t idx = (size_t)blockIdx.x * 256 + threadIdx.x; idx < n;
       idx += (size_t)gridDim.x * 256) {
    const int r = (int)(idx / pd.Wp) - pd.PT, c = (int)(idx % pd.Wp) - pd.PL;
    pad[idx] = (r >= 0 && r < H && c >= 0 && c < W) ? src[(size_t)r * W + c] : 0.0f;
  }
}

__global__ void __launch_bounds__(256)
k_convolve_same(const float* __restrict__ pad, int H, int W, ConvPad pd,
                const double* __restrict__ kt, int kh, int kw, float* __restrict__ dst) {
  constexpr int R = kConvRows;
  const int j = blockIdx.x * 64 + (threadIdx.x & 63);
  const int i0 = __builtin_amdgcn_readfirstlane((blockIdx.y * 4 + (int)(threadIdx.x >> 6)) * R);
  if (i0 >= H) return;  // wave-uniform
  const int oi = (kh - 1) / 2, oj = (kw - 1) / 2;
  const int kp = conv_padded_rows(kh);
  const int n_chunks = (kh + 2 * R - 2) / R;
  double acc[R];
#pragma unroll
  for (int r = 0; r < R; ++r) acc[r] = 0.0;
  for (int ch = 0; ch < n_chunks; ++ch) {
    const int d0 = -(kh - 1) + ch * R;
    // padded row of input row i0 + oi + d0 + m (wave-uniform)
    const float* row0 = pad + (size_t)(i0 + oi + d0 + pd.PT) * pd.Wp;
    // taps p = -d0 - (R-1) + e of column q sit at kt[q kp + R - 1 - d0 + e]
    const double* tcol = kt + (R - 1 - d0);
    for (int q = 0; q < kw; ++q) {
      const uint32_t cb = (uint32_t)(j + oj - q + pd.PL) * 4u;  // padded column, bytes
      double x[R];
#pragma unroll
      for (int m = 0; m < R; ++m)
        x[m] = (double)ld(row0 + (size_t)m * pd.Wp, cb);
      const double* t = tcol + (size_t)q * kp;
#pragma unroll
      for (int m = 0; m < R; ++m)    // d = d0 + m
#pragma unroll
        for (int r = 0; r < R; ++r)  // p = r - d: tap index r - m + R - 1
          acc[r] = fma(t[r - m + R - 1], x[m], acc[r]);
    }
  }
  if (j < W) {
#pragma unroll
    for (int r = 0; r < R; ++r)
      if (i0 + r < H) dst[(size_t)(i0 + r) * W + j] = (float)acc[r];
  }
}

// The scar-to-diffusivity map of random_diffusivity (deepx/generate.py:79-83,
// deepx/utils_scars.py:426, cardiax/convert.py:23-26): c = 1 - scar, then
// (c - min c) (hi - lo) / (max c - min c) + lo, in double (1 - x is exact
// there for a float x in [0, 1]), rounded once. min and max in two passes
// (per-block partials, then a fixed-order scan of them): deterministic.
constexpr int kMmBlocks = 256;

__global__ void __launch_bounds__(256)
k_minmax_partial(const float* __restrict__ x, int64_t n, double* __restrict__ part) {
  __shared__ double lo[256], hi[256];
  double a = INFINITY, b = -INFINITY;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
    const double c = 1.0 - (double)x[i];
    a = fmin(a, c);
    b = fmax(b, c);
  }
  lo[threadIdx.x] = a;
  hi[threadIdx.x] = b;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) {
      lo[threadIdx.x] = fmin(lo[threadIdx.x], lo[threadIdx.x + o]);
      hi[threadIdx.x] = fmax(hi[threadIdx.x], hi[threadIdx.x + o]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = lo[0];
    part[2 * blockIdx.x + 1] = hi[0];
  }
}

__global__ void __launch_bounds__(256)
k_scar_rescale(const float* __restrict__ x, int64_t n, const double* __restrict__ part,
               int n_part, double lo_d, double hi_d, float* __restrict__ D) {
#pragma clang fp contract(off)
  __shared__ double s_lo, s_hi;
  if (threadIdx.x == 0) {  // every block reduces the partials the same way
    double a = INFINITY, b = -INFINITY;
    for (int k = 0; k < n_part; ++k) {
      a = fmin(a, part[2 * k]);
      b = fmax(b, part[2 * k + 1]);
    }
    s_lo = a;
    s_hi = b;
  }
  __syncthreads();
  // the reference's order of operations: (c - a) * (z - y) / (b - a) + y
  const double a = s_lo, span = s_hi - s_lo, range = hi_d - lo_d;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
    D[i] = (float)((1.0 - (double)x[i] - a) * range / span + lo_d);
}

}  // namespace fk

using namespace fk;

extern "C" {

int32_t fk_rasterize_polygons(float* mask, int32_t H, int32_t W, const double* rows,
                              const double* cols, const int32_t* starts, int32_t n_polys,
                              void* work, size_t work_bytes, void* stream) {
  ErrorScope es;
  if (mask == nullptr || H < 1 || W < 1)
    return set_error(FK_ERR_ARG, "fk_rasterize_polygons: null mask or empty image");
  if (n_polys < 0 || (n_polys > 0 && (!rows || !cols || !starts)))
    return set_error(FK_ERR_ARG, "fk_rasterize_polygons: bad polygon list");
  const int32_t n_verts = n_polys > 0 ? starts[n_polys] : 0;
  for (int p = 0; p < n_polys; ++p)
    if (starts[p + 1] - starts[p] < 3 || starts[p] < 0)
      return set_error(FK_ERR_ARG, "polygon %d has fewer than 3 vertices", p);
  const size_t need = fk_rasterize_workspace_bytes(n_polys, n_verts);
  if (n_polys > 0 && (work == nullptr || work_bytes < need))
    return set_error(FK_ERR_ARG, "workspace too small: need %zu bytes, got %zu", need,
                     work_bytes);
  hipStream_t s = (hipStream_t)stream;
  // vertices and bounding boxes to the device (from host arrays; the call
  // waits for the stream before it returns, so they may be released then)
  char* wk = (char*)work;
  double* d_r = (double*)wk;
  double* d_c = d_r + n_verts;
  PolyBox* d_box = (PolyBox*)(wk + ((2 * (size_t)n_verts * sizeof(double) + 255) / 256) * 256);
  std::vector<PolyBox> boxes((size_t)n_polys);
  for (int p = 0; p < n_polys; ++p) {
    double rmin = rows[starts[p]], rmax = rmin, cmin = cols[starts[p]], cmax = cmin;
    for (int k = starts[p]; k < starts[p + 1]; ++k) {
      if (!std::isfinite(rows[k]) || !std::isfinite(cols[k]))
        return set_error(FK_ERR_ARG, "polygon %d has a non-finite vertex", p);
      rmin = std::fmin(rmin, rows[k]);
      rmax = std::fmax(rmax, rows[k]);
      cmin = std::fmin(cmin, cols[k]);
      cmax = std::fmax(cmax, cols[k]);
    }
    // skimage.draw.polygon's box, limited to the image (pixels beyond it are dropped)
    PolyBox b;
    b.r0 = (int)std::fmax(0.0, rmin);
    b.r1 = (int)std::fmin((double)(H - 1), std::ceil(rmax));
    b.c0 = (int)std::fmax(0.0, cmin);
    b.c1 = (int)std::fmin((double)(W - 1), std::ceil(cmax));
    b.begin = starts[p];
    b.end = starts[p + 1];
    boxes[p] = b;
  }
  if (n_polys > 0) {
    HIP_CHECK(hipMemcpyAsync(d_r, rows, n_verts * sizeof(double), hipMemcpyHostToDevice, s));
    HIP_CHECK(hipMemcpyAsync(d_c, cols, n_verts * sizeof(double), hipMemcpyHostToDevice, s));
    HIP_CHECK(hipMemcpyAsync(d_box, boxes.data(), n_polys * sizeof(PolyBox),
                             hipMemcpyHostToDevice, s));
  }
  const dim3 grid((W + 63) / 64, (H + 3) / 4);
  k_raster_polygons<<<grid, 256, 0, s>>>(mask, H, W, d_r, d_c, d_box, n_polys);
  HIP_CHECK(hipGetLastError());
  // the host-side staging of boxes must outlive the copy
  HIP_CHECK(hipStreamSynchronize(s));
  return FK_OK;
}

int32_t fk_scar_to_diffusivity(const float* scar, int64_t n_cells, double lo, double hi,
                               float* D, void* work, size_t work_bytes, void* stream) {
  ErrorScope es;
  if (!check_ptrs({scar, D, work}) || n_cells < 1)
    return set_error(FK_ERR_ARG, "fk_scar_to_diffusivity: null pointer or empty field");
  if (!std::isfinite(lo) || !std::isfinite(hi))
    return set_error(FK_ERR_ARG, "the diffusivity range must be finite");
  if (work_bytes < 2 * kMmBlocks * sizeof(double))
    return set_error(FK_ERR_ARG, "workspace too small: need %zu bytes, got %zu",
                     2 * kMmBlocks * sizeof(double), work_bytes);
  hipStream_t s = (hipStream_t)stream;
  const int blocks = (int)std::min<int64_t>((n_cells + 255) / 256, kMmBlocks);
  k_minmax_partial<<<blocks, 256, 0, s>>>(scar, n_cells, (double*)work);
  k_scar_rescale<<<blocks, 256, 0, s>>>(scar, n_cells, (const double*)work, blocks, lo, hi, D);
  HIP_CHECK(hipGetLastError());
  return FK_OK;
}

size_t fk_rasterize_workspace_bytes(int32_t n_polys, int32_t n_vertices) {
  if (n_polys <= 0 || n_vertices <= 0) return 0;
  return ((2 * (size_t)n_vertices * sizeof(double) + 255) / 256) * 256 +
         (size_t)n_polys * sizeof(PolyBox);
}

size_t fk_convolve_workspace_bytes(int32_t H, int32_t W, int32_t kh, int32_t kw) {
  if (H < 1 || W < 1 || kh < 1 || kw < 1) return 0;
  const ConvPad pd = conv_pad(H, W, kh, kw);
  const size_t taps = ((size_t)kw * conv_padded_rows(kh) * sizeof(double) + 255) / 256 * 256;
  return taps + (size_t)pd.Hp * pd.Wp * sizeof(float);
}

int32_t fk_convolve_same(const float* src, int32_t H, int32_t W, const double* kernel,
                         int32_t kh, int32_t kw, float* dst, void* work, size_t work_bytes,
                         void* stream) {
  ErrorScope es;
  if (!check_ptrs({src, kernel, dst, work})) return last_code();
  if (H < 1 || W < 1 || kh < 1 || kw < 1)
    return set_error(FK_ERR_ARG, "fk_convolve_same: empty image or kernel");
  if (kh > (1 << 16) || kw > (1 << 16) || H > (1 << 20) || W > (1 << 20))
    return set_error(FK_ERR_ARG, "fk_convolve_same: image %d x %d or kernel %d x %d too large",
                     H, W, kh, kw);
  if (src == dst) return set_error(FK_ERR_ARG, "fk_convolve_same: dst must not be src");
  const size_t need = fk_convolve_workspace_bytes(H, W, kh, kw);
  if (work_bytes < need)
    return set_error(FK_ERR_ARG, "workspace too small: need %zu bytes, got %zu", need,
                     work_bytes);
  hipStream_t s = (hipStream_t)stream;
  double* kt = (double*)work;
  const size_t taps = ((size_t)kw * conv_padded_rows(kh) * sizeof(double) + 255) / 256 * 256;
  float* pad = (float*)((char*)work + taps);
  const ConvPad pd = conv_pad(H, W, kh, kw);
  const int n_pad = kw * conv_padded_rows(kh);
  k_pad_kernel<<<(n_pad + 255) / 256, 256, 0, s>>>(kernel, kh, kw, kt);
  const size_t n_img = (size_t)pd.Hp * pd.Wp;
  k_pad_image<<<(int)std::min<size_t>((n_img + 255) / 256, 4096), 256, 0, s>>>(src, H, W, pd,
                                                                              pad);
  const dim3 grid((W + 63) / 64, (H + 4 * kConvRows - 1) / (4 * kConvRows));
  k_convolve_same<<<grid, 256, 0, s>>>(pad, H, W, pd, kt, kh, kw, dst);
  HIP_CHECK(hipGetLastError());
  return FK_OK;
}

}  // extern "C"
