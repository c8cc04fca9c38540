// Streaming forward-step kernel: one wavefront per (column strip, row band).
//
// Layout. Lane l of a wave holds column cs - 4 + l: 56 output columns plus 4
// halo columns on each side. The wave walks the R rows of its band.
//   * x (rows, axis 0): every lane keeps windows of u (rows rs-4 .. rs+R+3),
//     u_x (rs-2 .. rs+R+1) and D (rs-2 .. rs+R+1) in registers;
//     u_x(q) = G(u)(q), u_xx(i) = G(u_x)(i).
//   * y (columns, axis 1): neighbours +-1..3 columns away come from the lanes
//     next door through DPP wavefront shifts (v_mov_b32 wave_shr / wave_shl),
//     so u_y, u_yy and D_y need neither LDS nor a barrier.
// Row indices are wave-uniform (readfirstlane), so every load and store is a
// scalar row base plus the lane's column offset. All loads of a band are
// issued before its arithmetic.
//
// Boundary rules (cardiax/solve.py:225-254: one-sided differences on the two
// outermost padded entries, edge padding by one cell) only matter within 3
// cells of an edge. Bands and strips are placed so that the first starts at 0
// and the last ends at N (shifted back, overlapping its neighbour with
// identical results), and a band / strip that touches an edge runs the
// general per-position rules through a wave-uniform template switch; all
// others run the central rules only.
#pragma once

#include "fk_device.h"

namespace fk {

constexpr int kStripOut = 56;  // output columns per wave

// Lane l receives the value of lane l-1 (wave_shr:1) / lane l+1 (wave_shl:1).
// The lane without a source reads 0 (bound_ctrl): it is a halo lane whose
// result is unused. With no "old" operand the move needs no initialisation
// and can fold into the consuming instruction as a DPP source.
__device__ __forceinline__ float from_left(float x) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(x), 0x138, 0xf, 0xf, true));
}
__device__ __forceinline__ float from_right(float x) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(x), 0x130, 0xf, 0xf, true));
}

// First difference across columns (times 12, s1_* rules) of a value held one
// per lane. GENERAL: the lane's column `col` selects the rule, as d1_general
// does.
template <bool GENERAL>
__device__ __forceinline__ float d1_lanes(float x, int col, int W) {
#pragma clang fp contract(off)
  const float m1 = from_left(x), p1 = from_right(x);
  const float m2 = from_left(m1), p2 = from_right(p1);
  const float ctr = s1_interior(m2, m1, p1, p2);
  if (!GENERAL) return ctr;
  const float m3 = from_left(m2), p3 = from_right(p2);
  const float fwd = s1_forward(x, p1, p2, p3);
  const float bwd = s1_backward(m3, m2, m1, x);
  return col <= 0 ? fwd : (col >= W - 1 ? bwd : ctr);
}

// Second difference across columns from first differences g held one per
// lane (cardiax/solve.py:51-52 applied to u_y): one-sided at columns 0, W-1.
template <bool GENERAL>
__device__ __forceinline__ float d2_lanes(float g, int col, int W) {
#pragma clang fp contract(off)
  const float m1 = from_left(g), p1 = from_right(g);
  const float m2 = from_left(m1), p2 = from_right(p1);
  const float ctr = s1_interior(m2, m1, p1, p2);
  if (!GENERAL) return ctr;
  const float m3 = from_left(m2), p3 = from_right(p2);
  const float fwd = s1_forward(g, p1, p2, p3);
  const float bwd = s1_backward(m3, m2, m1, g);
  return col == 0 ? fwd : (col == W - 1 ? bwd : ctr);
}

// Work decomposition of one launch, computed on the host.
struct StreamGeom {
  int n_strips, n_bands, n_tiles, blocks;
  int cs_last, rs_last;  // origin of the last strip / band (shifted back to end at N)
};

template <int R>
inline StreamGeom make_stream_geom(int H, int W) {
  StreamGeom g{};
  g.n_strips = (W + kStripOut - 1) / kStripOut;
  g.n_bands = (H + R - 1) / R;
  g.cs_last = W >= kStripOut ? W - kStripOut : 0;
  g.rs_last = H >= R ? H - R : 0;
  g.n_tiles = g.n_strips * g.n_bands;
  g.blocks = (g.n_tiles + 3) / 4;
  return g;
}

// Load one float at byte offset `off` (32-bit, unsigned) from a wave-uniform
// row base: the scalar-base + vector-offset addressing form. Row pointers are
// taken as global-memory pointers (every field lives in device memory), so
// pointers read from a descriptor also compile to global, not flat, access.
typedef const char __attribute__((address_space(1)))* gcbytes;
typedef char __attribute__((address_space(1)))* gbytes;
__device__ __forceinline__ float ld(const float* rowp, uint32_t off) {
  return *(const float __attribute__((address_space(1)))*)((gcbytes)rowp + off);
}
__device__ __forceinline__ void st(float* rowp, uint32_t off, float x) {
  *(float __attribute__((address_space(1)))*)((gbytes)rowp + off) = x;
}
typedef float f32x4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ float4 ld4(const float4* rowp, uint32_t off) {
  const f32x4 x = *(const f32x4 __attribute__((address_space(1)))*)((gcbytes)rowp + off);
  return make_float4(x.x, x.y, x.z, x.w);
}


// Stimulus fields seen by a tile: count(), on(m) (field m active this step)
// and field(m), in list order. This view reads the active list passed as a
// kernel argument.
struct StimListView {
  const StimArgs& sa;
  __device__ int count() const { return sa.n; }
  __device__ bool on(int) const { return true; }
  __device__ const float* field(int m) const { return sa.f[m]; }
};

// The extra inputs and outputs of the two stages of a Heun step
// (cardiax/solve.py:73-85). Stage 1 reads x0 and writes only the slope
// k1 = step(x0). Stage 2 reads x0 (as the state) and k1, rebuilds
// x1 = x0 + dt k1 in registers with the one fmaf stage 1 would have used,
// evaluates k2 = step(x1) and writes x0 + (k1 + k2) dt / 2. x1 never goes to
// memory: 68 bytes per cell-step instead of 92.
struct HeunIO {
  float *kv, *kw, *ku;              // stage 1: k1 out
  const float *k1v, *k1w, *k1u;     // stage 2: k1 in
};

// PC: the static diffusion coefficients come precomputed from `cf`
// (k_stream_coef) instead of being derived from D each step. HEUN: 0 = an
// Euler step (EULER) or the derivatives, 1 / 2 = stage 1 / 2 of a Heun step.
// FIELD_MAJOR: loads field by field with scalar row bases, measured faster
// when the fields stream from DRAM (the batch); otherwise in the order the
// output rows need them, measured faster for grids held in the 256 MB cache.
template <int R, bool EULER, bool XG, bool YG, bool STIM, class SV, bool PC = false,
          int HEUN = 0, bool FIELD_MAJOR = false>
__device__ __forceinline__ void stream_tile(const Coef& c, const float* __restrict__ D,
                                            const float4* __restrict__ cf,
                                            const float* __restrict__ v,
                                            const float* __restrict__ w,
                                            const float* __restrict__ u, int H, int W,
                                            const SV& sv, float dt, float inv_dx2,
                                            float* __restrict__ o_v, float* __restrict__ o_w,
                                            float* __restrict__ o_u, int rs, int cs,
                                            const HeunIO& io = HeunIO{}) {
#pragma clang fp contract(off)  // explicit fmaf only: identical rounding in every kernel
  const int lane = threadIdx.x & 63;
  const int col = cs - 4 + lane;                           // this lane's column
  const uint32_t off = (uint32_t)clampi(col, 0, W - 1) * 4u;  // clamped = edge padding
  const bool out_col = lane >= 4 && lane < 4 + kStripOut && col < W;
  // Row rs - 4 + k of a field. Edge bands clamp the row (edge padding) and
  // address it from a scalar row base. Interior bands (edge_band false) need
  // no clamp. With FIELD_MAJOR they also take scalar row bases (field base
  // plus a row multiple); otherwise one scalar band offset per field base and
  // a 32-bit per-row offset, k * 4W, that every field shares in a vector
  // register.
  const uint32_t rb = (uint32_t)W * 4u;
  const size_t band = XG ? 0 : (size_t)(rs - 4) * (size_t)W * 4u;
  auto ldk = [&](const float* a, int k) {
    if constexpr (XG)
      return ld(a + (size_t)clampi(rs - 4 + k, 0, H - 1) * (size_t)W, off);
    else if constexpr (FIELD_MAJOR)
      return ld(a + (size_t)(rs - 4 + k) * (size_t)W, off);
    else
      return ld((const float*)((const char*)a + band), off + (uint32_t)k * rb);
  };
  auto stk = [&](float* a, int k, float x) {  // rows inside the grid only
    if constexpr (XG || FIELD_MAJOR)
      st(a + (size_t)(rs - 4 + k) * (size_t)W, off, x);
    else
      st((float*)((char*)a + band), off + (uint32_t)k * rb, x);
  };

  // Precomputed coefficients of output row k (PC): float4 rows of pitch 16W
  // bytes, addressed like the fields (clamped scalar row base on edge bands;
  // one band base and a shared 32-bit row offset inside).
  auto ldc = [&](int k) {
    if constexpr (XG || FIELD_MAJOR)
      return ld4(cf + (size_t)clampi(rs + k, 0, H - 1) * (size_t)W, off * 4u);
    else
      return ld4(cf + (size_t)rs * (size_t)W, off * 4u + (uint32_t)k * (rb * 4u));
  };

  float uu[R + 8];           // u rows rs-4 .. rs+R+3
  float dd[PC ? 1 : R + 4];  // D rows rs-2 .. rs+R+1
  float4 cc[PC ? R : 1];     // precomputed coefficients of rows rs .. rs+R-1
  float vv[R], ww[R];
  float h0[HEUN == 2 ? 3 * R : 1], hk[HEUN == 2 ? 3 * R : 1];  // stage 2: x0, k1 of the band
  // Row k of state field f (0 v, 1 w, 2 u) from `a`. In stage 2 of a Heun step
  // that is x1 = fmaf(dt, k1, x0) from x0 (`a`) and k1 (`ka`); the band's own
  // rows also keep x0 and k1 for the final combination.
  auto ldx = [&](const float* a, const float* ka, int k, int f) {
    if constexpr (HEUN == 2) {
      const float x0 = ldk(a, k), k1 = ldk(ka, k);
      if (k >= 4 && k < R + 4) {
        h0[3 * (k - 4) + f] = x0;
        hk[3 * (k - 4) + f] = k1;
      }
      return fmaf(dt, k1, x0);
    } else {
      (void)ka;
      (void)f;
      return ldk(a, k);
    }
  };
  if constexpr (FIELD_MAJOR) {
    // Field by field: every u row of the window, then D, then v and w.
#pragma unroll
    for (int k = 0; k < R + 8; ++k) uu[k] = ldx(u, io.k1u, k, 2);
    if constexpr (!PC) {
#pragma unroll
      for (int k = 0; k < R + 4; ++k) dd[k] = ldk(D, k + 2);
    }
#pragma unroll
    for (int k = 0; k < R; ++k) {
      vv[k] = ldx(v, io.k1v, k + 4, 0);
      ww[k] = ldx(w, io.k1w, k + 4, 1);
    }
    if constexpr (PC) {
#pragma unroll
      for (int k = 0; k < R; ++k) cc[k] = ldc(k);
    }
  } else {
    // In the order the output rows need them: first everything output row 0
    // reads (u rows 0..8 of the window with D rows 2..6 alongside, or its
    // coefficients, then v and w of row 4), then the rows each further output
    // row adds. Loads return in order, so an output row's arithmetic starts
    // once its own rows have arrived while the band's later rows are still in
    // flight.
#pragma unroll
    for (int k = 0; k < 9; ++k) {
      uu[k] = ldx(u, io.k1u, k, 2);
      if constexpr (!PC)
        if (k >= 2 && k < 7) dd[k - 2] = ldk(D, k);
    }
    if constexpr (PC) cc[0] = ldc(0);
    vv[0] = ldx(v, io.k1v, 4, 0);
    ww[0] = ldx(w, io.k1w, 4, 1);
#pragma unroll
    for (int j = 1; j < R; ++j) {
      uu[j + 8] = ldx(u, io.k1u, j + 8, 2);
      if constexpr (!PC) dd[j + 4] = ldk(D, j + 6);
      if constexpr (PC) cc[j] = ldc(j);
      vv[j] = ldx(v, io.k1v, j + 4, 0);
      ww[j] = ldx(w, io.k1w, j + 4, 1);
    }
  }

  // 12 u_x at rows q = rs-2 .. rs+R+1 (every difference below is scaled by
  // 12; the diffusion term divides by 144 once)
  const float k2 = inv_dx2 * (1.0f / 144.0f);
  float gx[R + 4];
#pragma unroll
  for (int k = 0; k < R + 4; ++k) {
    const int q = rs - 2 + k;
    if (XG && k <= 2 && q <= 0)  // forward rule reads u(q..q+3); q <= 0 forces k <= 2
      gx[k] = s1_forward(uu[k + 2], uu[k + 3], uu[k + 4], uu[k + 5]);
    else if (XG && k >= 3 && q >= H - 1)  // backward rule reads u(q-3..q)
      gx[k] = s1_backward(uu[k - 1], uu[k], uu[k + 1], uu[k + 2]);
    else
      gx[k] = s1_interior(uu[k], uu[k + 1], uu[k + 3], uu[k + 4]);
  }

#pragma unroll
  for (int k = 0; k < R; ++k) {
    const int i = rs + k;
    const float uc = uu[k + 4], gxc = gx[k + 2];
    float uxx;
    if (XG && k == 0 && i == 0)  // i = rs + k = 0 forces k = 0
      uxx = s1_forward(gx[k + 2], gx[k + 3], gx[k + 4], gx[k + 5]);
    else if (XG && k >= 1 && i == H - 1)
      uxx = s1_backward(gx[k - 1], gx[k], gx[k + 1], gx[k + 2]);
    else
      uxx = s1_interior(gx[k], gx[k + 1], gx[k + 3], gx[k + 4]);
    const float gyc = d1_lanes<YG>(uc, col, W);
    const float uyy = d2_lanes<YG>(gyc, col, W);
    float del_u;
    if constexpr (PC) {
      del_u = fmaf(cc[k].z, gyc, fmaf(cc[k].y, gxc, cc[k].x * (uxx + uyy)));
    } else {
      const float Dc = dd[k + 2];
      float dDx;
      if (XG && k == 0 && i == 0)
        dDx = s1_forward(dd[k + 2], dd[k + 3], dd[k + 4], dd[k + 5]);
      else if (XG && k >= 1 && i == H - 1)
        dDx = s1_backward(dd[k - 1], dd[k], dd[k + 1], dd[k + 2]);
      else
        dDx = s1_interior(dd[k], dd[k + 1], dd[k + 3], dd[k + 4]);
      const float dDy = d1_lanes<YG>(Dc, col, W);
      del_u = fmaf(dDy, gyc, fmaf(dDx, gxc, Dc * (uxx + uyy))) * k2;
    }
    float dv, dw, jion;
    reaction(c, vv[k], ww[k], uc, dv, dw, jion);
    if (STIM) {
      float sval = 0.0f;
      for (int m = 0; m < sv.count(); ++m) {
        if (!sv.on(m)) continue;
        const float f = ldk(sv.field(m), k + 4);
        sval = (f != 0.0f) ? f : sval;
      }
      jion = (sval != 0.0f) ? sval : jion;
    }
    const float du = del_u + jion;
    if (out_col && (!XG || i < H)) {  // rows past the grid only when H < R
      if constexpr (HEUN == 1) {  // k1 only
        stk(io.kv, k + 4, dv);
        stk(io.kw, k + 4, dw);
        stk(io.ku, k + 4, du);
      } else if constexpr (HEUN == 2) {  // x0 + (k1 + k2) dt / 2
        const float half = dt * 0.5f;
        stk(o_v, k + 4, fmaf(hk[3 * k] + dv, half, h0[3 * k]));
        stk(o_w, k + 4, fmaf(hk[3 * k + 1] + dw, half, h0[3 * k + 1]));
        stk(o_u, k + 4, fmaf(hk[3 * k + 2] + du, half, h0[3 * k + 2]));
      } else if (EULER) {
        stk(o_v, k + 4, fmaf(dt, dv, vv[k]));
        stk(o_w, k + 4, fmaf(dt, dw, ww[k]));
        stk(o_u, k + 4, fmaf(dt, du, uc));
      } else {
        stk(o_v, k + 4, dv);
        stk(o_w, k + 4, dw);
        stk(o_u, k + 4, du);
      }
    }
  }
}

// A band takes the edge path (clamped rows, one-sided rules) when it lies
// within 4 rows of an edge, since the interior path reads rows rs-4 .. rs+R+3
// unclamped, or when R + 8 rows span 4 GiB or more, past the interior path's
// 32-bit row offsets.
__host__ __device__ inline bool edge_band(int rs, int R, int H, int W) {
  return rs < 4 || rs + R + 4 > H || (uint64_t)(R + 8) * (uint64_t)W * 4u > 0xffffffffull;
}

// One wave per (strip, band) tile, band-major within a strip so that
// consecutive tiles share row halos; four waves per block. PC: Euler steps
// from precomputed coefficients (k_stream_c), else derived from D.
template <int R, bool EULER, bool PC, int HEUN = 0>
__device__ __forceinline__ void stream_kernel_body(
    const Coef& c, const float* __restrict__ D, const float4* __restrict__ cf,
    const float* __restrict__ v, const float* __restrict__ w, const float* __restrict__ u,
    int H, int W, const StimArgs& sa, float dt, float inv_dx2, float* __restrict__ o_v,
    float* __restrict__ o_w, float* __restrict__ o_u, const StreamGeom& g,
    const HeunIO& io = HeunIO{}) {
  const int tile = __builtin_amdgcn_readfirstlane(blockIdx.x * 4 + (int)(threadIdx.x >> 6));
  if (tile >= g.n_tiles) return;  // whole wave leaves; the kernel has no barrier
  const int strip = tile / g.n_bands, band = tile - strip * g.n_bands;
  const int cs = min(kStripOut * strip, g.cs_last);
  const int rs = min(R * band, g.rs_last);
  const bool xg = edge_band(rs, R, H, W);
  const bool yg = cs < 3 || cs + kStripOut + 3 > W;  // strip within 3 columns
  const int sel = (xg ? 1 : 0) | (yg ? 2 : 0) | (sa.n ? 4 : 0);
  const StimListView sv{sa};
#define FK_TILE(X, Y, S)                                                                 \
  stream_tile<R, EULER, X, Y, S, StimListView, PC, HEUN>(c, D, cf, v, w, u, H, W, sv, dt,  \
                                                         inv_dx2, o_v, o_w, o_u, rs, cs, io)
  switch (sel) {
    case 0: FK_TILE(false, false, false); break;
    case 1: FK_TILE(true, false, false); break;
    case 2: FK_TILE(false, true, false); break;
    case 3: FK_TILE(true, true, false); break;
    case 4: FK_TILE(false, false, true); break;
    case 5: FK_TILE(true, false, true); break;
    case 6: FK_TILE(false, true, true); break;
    default: FK_TILE(true, true, true); break;
  }
#undef FK_TILE
}

template <int R, bool EULER>
__global__ void __launch_bounds__(256)
k_stream(Coef c, const float* __restrict__ D, const float* __restrict__ v,
         const float* __restrict__ w, const float* __restrict__ u, int H, int W, StimArgs sa,
         float dt, float inv_dx2, float* __restrict__ o_v, float* __restrict__ o_w,
         float* __restrict__ o_u, StreamGeom g) {
  stream_kernel_body<R, EULER, false>(c, D, nullptr, v, w, u, H, W, sa, dt, inv_dx2, o_v, o_w,
                                      o_u, g);
}

// Stage STAGE (1 or 2) of a Heun step on streaming bands of R rows.
template <int R, int STAGE>
__global__ void __launch_bounds__(256)
k_stream_heun(Coef c, const float* __restrict__ D, const float* __restrict__ v,
              const float* __restrict__ w, const float* __restrict__ u, int H, int W,
              StimArgs sa, float dt, float inv_dx2, float* __restrict__ o_v,
              float* __restrict__ o_w, float* __restrict__ o_u, StreamGeom g, HeunIO io) {
  stream_kernel_body<R, true, false, STAGE>(c, D, nullptr, v, w, u, H, W, sa, dt, inv_dx2, o_v,
                                            o_w, o_u, g, io);
}

template <int R>
__global__ void __launch_bounds__(256)
k_stream_c(Coef c, const float4* __restrict__ cf, const float* __restrict__ v,
           const float* __restrict__ w, const float* __restrict__ u, int H, int W, StimArgs sa,
           float dt, float* __restrict__ o_v, float* __restrict__ o_w, float* __restrict__ o_u,
           StreamGeom g) {
  stream_kernel_body<R, true, true>(c, nullptr, cf, v, w, u, H, W, sa, dt, 0.0f, o_v, o_w, o_u,
                                    g);
}

}  // namespace fk
