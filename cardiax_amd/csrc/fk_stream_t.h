// Two Euler steps per launch on streaming bands (temporal blocking in
// registers).
//
// One wavefront per (column strip, row band) tile, as in k_stream, but the
// wave advances its tile two steps: level 1 computes the state at t+1 on the
// R + 8 rows and 56 columns that level 2's stencil reaches, level 2 the state
// at t+2 on the tile's R rows and 48 columns. Nothing of level 1 leaves the
// registers, so a launch reads v, w, u, D and writes v, w, u once for two
// steps, and the fixed cost of a dependent launch (dispatch, drain, the L2
// writeback between kernels) is paid once per two steps.
//
// Layout. Lane l holds column cs - 8 + l: 48 output columns (lanes 8..55)
// and an 8-lane halo on each side, 4 columns per level. Level 1 is valid on
// lanes 4..59, level 2 on lanes 8..55 (DPP shifts zero-fill the outer lanes,
// whose results are never used). The band's rows are a window of registers:
// u at rows rs-8 .. rs+R+7, D at rs-6 .. rs+R+5, v and w at rs-4 .. rs+R+3.
//
// Arithmetic. Every cell of both levels is computed by the same operations,
// in the same order, as in stream_tile (contraction off, explicit fmaf), so
// one launch equals two launches of k_stream bit for bit. Level 2 reuses the
// static D derivatives level 1 derived for its rows, and level 1 derives only
// u on the rows that only level 2's stencil needs.
//
// Boundaries (cardiax/solve.py:225-254: edge padding, one-sided differences).
// Bands are placed from both ends of the axis (the middle two overlap and
// write identical values), so every band is the first (rs = 0), the last
// (rs = H - R) or interior (rs >= 8, rs + R + 8 <= H); likewise strips of 48
// columns. The edge bands know at compile time which window row is the grid's
// first or last: their loads clamp to it (edge padding of the inputs), the
// one-sided rules apply where the reference applies them, and level 1's rows
// beyond the grid are copies of its edge row (edge padding of the state at
// t+1). In the first and last strip the lanes beyond the grid take the edge
// column's level-1 value the same way. Sizes: stream_t_fits.
#pragma once

#include "fk_stream.h"

namespace fk {

constexpr int kTOut = 48;  // output columns per wave

struct StreamTGeom {
  int n_strips, n_bands, n_tiles, blocks;
};

template <int R>
inline StreamTGeom make_stream_t_geom(int H, int W) {
  StreamTGeom g{};
  g.n_strips = (W + kTOut - 1) / kTOut;
  g.n_bands = (H + R - 1) / R;
  g.n_tiles = g.n_strips * g.n_bands;
  g.blocks = (g.n_tiles + 3) / 4;
  return g;
}

// Grids the kernel handles: every band the first, the last or interior (H >=
// 2R + 16), and every strip the first, the last or interior: two strips for
// 56 <= W <= 96 (the last one's left halo inside the grid), three or more
// from W = 104.
template <int R>
__host__ __device__ inline bool stream_t_fits(int H, int W) {
  return H >= 2 * R + 16 && ((W >= 56 && W <= 2 * kTOut) || W >= 2 * kTOut + 8);
}

// Origin of tile k of n along an axis of N cells with tiles of T: the first
// half from the start, the second from the end.
__host__ __device__ inline int stream_t_origin(int k, int n, int T, int N) {
  return k < n / 2 ? T * k : N - T * (n - k);
}

// The reaction of reaction() (cardiax/solve.py:35-42,57-58) by the same
// operations, with both sides of each gate computed and then selected.
__device__ __forceinline__ float fo_sel(const Coef& c, float v, float u, bool p) {
#pragma clang fp contract(off)
  const float a = (u - c.V_c) * (1.0f - u);
  const float exc = fmaf(-(v * a), c.inv_td, c.inv_tr);
  const float rest = u * c.inv_t0;
  return p ? exc : rest;
}

__device__ __forceinline__ void reaction_sel(const Coef& c, float v, float w, float u, float& dv,
                                             float& dw, float& jion) {
#pragma clang fp contract(off)
  const bool p = u >= c.V_c;
  const float fo = fo_sel(c, v, u, p);
  const float s = fmaf(-(w * c.inv_tsi), half_one_plus_tanh(c, u), fo);
  const float inv_tvm = (u >= c.V_v) ? c.inv_tv2m : c.inv_tv1m;
  const float dv_e = -v * c.inv_tvp, dv_r = (1.0f - v) * inv_tvm;
  const float dw_e = -w * c.inv_twp, dw_r = (1.0f - w) * c.inv_twm;
  dv = p ? dv_e : dv_r;
  dw = p ? dw_e : dw_r;
  jion = -s * c.inv_cm;
}

// j_ion alone, by the same operations: level 1 needs only u on the rows
// that only level 2's stencil reaches.
__device__ __forceinline__ float jion_only(const Coef& c, float v, float w, float u) {
#pragma clang fp contract(off)
  const bool p = u >= c.V_c;
  const float fo = fo_sel(c, v, u, p);
  const float s = fmaf(-(w * c.inv_tsi), half_one_plus_tanh(c, u), fo);
  return -s * c.inv_cm;
}

// XE: 0 interior band, 1 first band (rs = 0), 2 last band (rs = H - R).
// YG: the strip touches a side (general column rules; lanes beyond the grid
// padded after level 1). STIM: some stimulus is active at t or t+1. s1, s2:
// the stimuli of the steps at t and t+1 (count(), on(m), field(m)).
template <int R, int XE, bool YG, bool STIM, class SV>
__device__ __forceinline__ void stream_t_tile(const Coef& c, const float* __restrict__ D,
                                              const float* __restrict__ v,
                                              const float* __restrict__ w,
                                              const float* __restrict__ u, int H, int W,
                                              const SV& s1, const SV& s2, float dt,
                                              float inv_dx2, float* __restrict__ o_v,
                                              float* __restrict__ o_w, float* __restrict__ o_u,
                                              int rs, int cs) {
#pragma clang fp contract(off)  // explicit fmaf only: the rounding of stream_tile
  constexpr int NU = R + 16;  // u window: rows rs-8 .. rs+R+7 (window index k)
  constexpr int N1 = R + 8;   // level-1 rows rs-4 .. rs+R+3 (index m = k - 4)
  // Window index whose row the load of index k reads: the edge bands clamp at
  // compile time (rows above 0 read row 0, rows below H-1 read row H-1).
  auto kc = [](int k) constexpr { return XE == 1 ? (k < 8 ? 8 : k) : XE == 2 ? (k > R + 7 ? R + 7 : k) : k; };
  const int lane = threadIdx.x & 63;
  const int col = cs - 8 + lane;
  const uint32_t off = (uint32_t)clampi(col, 0, W - 1) * 4u;
  // Row r of a field from a wave-uniform (scalar) row base plus the lane's
  // column offset: no per-row vector registers, whatever the window's height.
  auto ldk = [&](const float* a, int k) {
    return ld(a + (size_t)(rs - 8 + kc(k)) * (size_t)W, off);
  };
  auto stk = [&](float* a, int k, float x) { st(a + (size_t)(rs - 8 + k) * (size_t)W, off, x); };
  // Level-1 rows computed: the rest are copies of the edge row.
  constexpr int m_lo = XE == 1 ? 4 : 0;
  constexpr int m_hi = XE == 2 ? R + 3 : N1 - 1;
  const float k2 = inv_dx2 * (1.0f / 144.0f);
  const bool out_col = lane >= 8 && lane < 8 + kTOut;

  // The band is walked once, top to bottom. Iteration m loads rows LA ahead,
  // computes level 1's row m and level 2's row m - 9; each difference is
  // taken at the iteration its last input arrives (the one-sided rules of the
  // first band reach one row further than the central ones), so every window
  // lives only as long as its stencils reach and few registers are live.
  constexpr int LA = 9;  // rows of u loaded ahead of level 1's row
  float uu[NU], dd[R + 14], g0[R + 12];  // u and D by window index, 12 u_x of u
  float u1[N1], g1[R + 4];               // level 1: u at t+1 and its 12 u_x
  float v1[R], w1[R], sDx[R], sDy[R];    // level 1 on the tile's rows

  // 12 u_x of the input at window row j + 2: the reference's rules, one-sided
  // on the two padded rows at an edge (compile-time in the edge bands).
  // need0(j): the last window row it reads.
  auto fwd0 = [](int j) constexpr { return XE == 1 && j + 2 - 8 <= 0 && j + 2 - 8 >= -1; };
  auto bwd0 = [](int j) constexpr { return XE == 2 && j + 2 - (R + 7) >= 0 && j + 2 - (R + 7) <= 1; };
  auto need0 = [&](int j) constexpr { return fwd0(j) ? j + 5 : bwd0(j) ? j + 2 : j + 4; };
  auto grad0 = [&](int j) {
    const int k = j + 2;
    if (fwd0(j))
      g0[j] = s1_forward(uu[k], uu[k + 1], uu[k + 2], uu[k + 3]);
    else if (bwd0(j))
      g0[j] = s1_backward(uu[k - 3], uu[k - 2], uu[k - 1], uu[k]);
    else
      g0[j] = s1_interior(uu[k - 2], uu[k - 1], uu[k + 1], uu[k + 2]);
  };
  // 12 u_x of the level-1 state at level-1 row j + 2, and the iteration whose
  // level-1 row completes its inputs (the first band's copies above row 0
  // exist from level-1 row 4 on)
  auto fwd1 = [](int j) constexpr { return XE == 1 && j + 2 - 4 <= 0 && j + 2 - 4 >= -1; };
  auto bwd1 = [](int j) constexpr { return XE == 2 && j + 2 - (R + 3) >= 0 && j + 2 - (R + 3) <= 1; };
  auto avail1 = [&](int j) constexpr {
    const int need = fwd1(j) ? j + 5 : bwd1(j) ? j + 2 : j + 4;
    return need < m_lo ? m_lo : need;
  };
  auto grad1 = [&](int j) {
    const int m = j + 2;
    if (fwd1(j))
      g1[j] = s1_forward(u1[m], u1[m + 1], u1[m + 2], u1[m + 3]);
    else if (bwd1(j))
      g1[j] = s1_backward(u1[m - 3], u1[m - 2], u1[m - 1], u1[m]);
    else
      g1[j] = s1_interior(u1[m - 2], u1[m - 1], u1[m + 1], u1[m + 2]);
  };

  // ---- level 2: t+1 -> t+2 at row rs+n -------------------------------------
  auto level2 = [&](int n) {
    const float uc = u1[n + 4], gxc = g1[n + 2];
    float uxx;
    if (XE == 1 && n == 0)
      uxx = s1_forward(g1[n + 2], g1[n + 3], g1[n + 4], g1[n + 5]);
    else if (XE == 2 && n == R - 1)
      uxx = s1_backward(g1[n - 1], g1[n], g1[n + 1], g1[n + 2]);
    else
      uxx = s1_interior(g1[n], g1[n + 1], g1[n + 3], g1[n + 4]);
    const float gyc = d1_lanes<YG>(uc, col, W);
    const float uyy = d2_lanes<YG>(gyc, col, W);
    const float Dc = dd[n + 8];
    const float del_u = fmaf(sDy[n], gyc, fmaf(sDx[n], gxc, Dc * (uxx + uyy))) * k2;
    float dv, dw, jion;
    reaction_sel(c, v1[n], w1[n], uc, dv, dw, jion);
    if (STIM) {
      float sval = 0.0f;
      for (int q = 0; q < s2.count(); ++q) {
        if (!s2.on(q)) continue;
        const float f = ldk(s2.field(q), n + 8);
        sval = (f != 0.0f) ? f : sval;
      }
      jion = (sval != 0.0f) ? sval : jion;
    }
    const float du = del_u + jion;
    if (out_col) {
      stk(o_v, n + 8, fmaf(dt, dv, v1[n]));
      stk(o_w, n + 8, fmaf(dt, dw, w1[n]));
      stk(o_u, n + 8, fmaf(dt, du, uc));
    }
  };

#pragma unroll
  for (int k = 0; k < LA; ++k) uu[k] = ldk(u, k);
#pragma unroll
  for (int k = 2; k < LA - 2; ++k) dd[k] = ldk(D, k);
#pragma unroll
  for (int j = 0; j < R + 12; ++j)
    if (need0(j) < LA) grad0(j);

#pragma unroll
  for (int m = 0; m <= N1; ++m) {
    const int k = m + 4;  // window index of level-1 row m
    if (m + LA < NU) uu[m + LA] = ldk(u, m + LA);
    if (m + LA - 2 < R + 14) dd[m + LA - 2] = ldk(D, m + LA - 2);
#pragma unroll
    for (int j = 0; j < R + 12; ++j)
      if (need0(j) == m + LA) grad0(j);
    if (m >= m_lo && m <= m_hi) {
      // ---- level 1: t -> t+1 at row rs-4+m --------------------------------
      const float vv = ldk(v, k), ww = ldk(w, k);
      const float uc = uu[k], gxc = g0[m + 2];
      float uxx;
      if (XE == 1 && k == 8)  // row 0
        uxx = s1_forward(g0[m + 2], g0[m + 3], g0[m + 4], g0[m + 5]);
      else if (XE == 2 && k == R + 7)  // row H-1
        uxx = s1_backward(g0[m - 1], g0[m], g0[m + 1], g0[m + 2]);
      else
        uxx = s1_interior(g0[m], g0[m + 1], g0[m + 3], g0[m + 4]);
      const float gyc = d1_lanes<YG>(uc, col, W);
      const float uyy = d2_lanes<YG>(gyc, col, W);
      const float Dc = dd[k];
      float dDx;
      if (XE == 1 && k == 8)
        dDx = s1_forward(dd[k], dd[k + 1], dd[k + 2], dd[k + 3]);
      else if (XE == 2 && k == R + 7)
        dDx = s1_backward(dd[k - 3], dd[k - 2], dd[k - 1], dd[k]);
      else
        dDx = s1_interior(dd[k - 2], dd[k - 1], dd[k + 1], dd[k + 2]);
      const float dDy = d1_lanes<YG>(Dc, col, W);
      const float del_u = fmaf(dDy, gyc, fmaf(dDx, gxc, Dc * (uxx + uyy))) * k2;
      const bool own = m >= 4 && m < R + 4;  // a row of the tile: level 2 needs v, w
      float dv = 0.0f, dw = 0.0f, jion;
      if (own)
        reaction_sel(c, vv, ww, uc, dv, dw, jion);
      else
        jion = jion_only(c, vv, ww, uc);
      if (STIM) {
        float sval = 0.0f;
        for (int q = 0; q < s1.count(); ++q) {
          if (!s1.on(q)) continue;
          const float f = ldk(s1.field(q), k);
          sval = (f != 0.0f) ? f : sval;
        }
        jion = (sval != 0.0f) ? sval : jion;
      }
      const float du = del_u + jion;
      u1[m] = fmaf(dt, du, uc);
      if (own) {
        v1[m - 4] = fmaf(dt, dv, vv);
        w1[m - 4] = fmaf(dt, dw, ww);
        sDx[m - 4] = dDx;
        sDy[m - 4] = dDy;
      }
      if (YG) {  // lanes beyond the grid take the edge column's value
        if (cs == 0) {
          const int e = __builtin_amdgcn_readlane(__float_as_int(u1[m]), 8);
          u1[m] = lane < 8 ? __int_as_float(e) : u1[m];
        } else if (cs == W - kTOut) {
          const int e = __builtin_amdgcn_readlane(__float_as_int(u1[m]), 55);
          u1[m] = lane > 55 ? __int_as_float(e) : u1[m];
        }
      }
      if (m == m_lo)  // edge padding of the state at t+1 above row 0
        for (int r = 0; r < m_lo; ++r) u1[r] = u1[m_lo];
      // Materialise this row's level-1 values here: left to itself the
      // compiler sinks each level-1 row to level 2's first use of it, nine rows
      // later, and every input row stays live meanwhile. (An empty asm that
      // takes the values as inputs; it emits no instruction.)
      if (own)
        asm volatile("" ::"v"(u1[m]), "v"(v1[m - 4]), "v"(w1[m - 4]), "v"(sDx[m - 4]),
                     "v"(sDy[m - 4]));
      else
        asm volatile("" ::"v"(u1[m]));
    } else if (m > m_hi && m < N1) {
      u1[m] = u1[m_hi];  // edge padding below row H-1
    }
#pragma unroll
    for (int j = 0; j < R + 4; ++j)
      if (avail1(j) == m) grad1(j);
    if (m >= 9) level2(m - 9);
    // keep the scheduler from hoisting later iterations' loads above this
    // one: they would hold registers for the whole band
    __builtin_amdgcn_sched_barrier(0);
  }
}

template <int R>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2)))
k_stream_t(Coef c, const float* __restrict__ D, const float* __restrict__ v,
           const float* __restrict__ w, const float* __restrict__ u, int H, int W, StimArgs s1,
           StimArgs s2, float dt, float inv_dx2, float* __restrict__ o_v,
           float* __restrict__ o_w, float* __restrict__ o_u, StreamTGeom g) {
  const int tile = __builtin_amdgcn_readfirstlane(blockIdx.x * 4 + (int)(threadIdx.x >> 6));
  if (tile >= g.n_tiles) return;  // whole wave leaves; the kernel has no barrier
  const int strip = tile / g.n_bands, band = tile - strip * g.n_bands;
  const int cs = stream_t_origin(strip, g.n_strips, kTOut, W);
  const int rs = stream_t_origin(band, g.n_bands, R, H);
  const int xe = rs == 0 ? 1 : (rs == H - R ? 2 : 0);
  const bool yg = cs == 0 || cs == W - kTOut;
  const int sel = xe * 4 + (yg ? 2 : 0) + ((s1.n | s2.n) ? 1 : 0);
  const StimListView v1{s1}, v2{s2};
#define FK_T(X, Y, S)                                                                      \
  stream_t_tile<R, X, Y, S, StimListView>(c, D, v, w, u, H, W, v1, v2, dt, inv_dx2, o_v, o_w, \
                                          o_u, rs, cs)
  switch (sel) {
    case 0: FK_T(0, false, false); break;
    case 1: FK_T(0, false, true); break;
    case 2: FK_T(0, true, false); break;
    case 3: FK_T(0, true, true); break;
    case 4: FK_T(1, false, false); break;
    case 5: FK_T(1, false, true); break;
    case 6: FK_T(1, true, false); break;
    case 7: FK_T(1, true, true); break;
    case 8: FK_T(2, false, false); break;
    case 9: FK_T(2, false, true); break;
    case 10: FK_T(2, true, false); break;
    default: FK_T(2, true, true); break;
  }
#undef FK_T
}

}  // namespace fk
