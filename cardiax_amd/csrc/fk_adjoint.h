// Device side of the adjoint of one forward-Euler step (the reverse step, its
// derived fields and the boundary tables of the transposed operators), shared
// by the per-grid kernels (fk_adjoint.hip) and the batched rollout
// (fk_rollout_batch.hip): one code path, so both give identical bits.
// The mathematics is described at the top of fk_adjoint.hip.
#pragma once

#include <hip/hip_runtime.h>

#include "fk_device.h"

namespace fk {

constexpr int kAdjB = 8;       // tabulated rows at each edge
constexpr int kAdjSlots = 16;  // 2 * kAdjB
constexpr int kAdjBand = 5;    // |column - row| <= 5 in A1, A2

// Tables of one axis: row slot -> A1[row][row + d], A2[row][row + d], d in [-5, 5].
struct AdjAxisTab {
  float a1[kAdjSlots][2 * kAdjBand + 1];
  float a2[kAdjSlots][2 * kAdjBand + 1];
};

// Gradient rule of cardiax/solve.py:230-250 as a matrix on n padded entries:
// the coefficient of entry q in row p.
__device__ __forceinline__ double g_coef(int p, int q, int n) {
  const int d = q - p;
  if (p <= 1) {
    const double c[4] = {-11.0 / 6.0, 3.0, -1.5, 1.0 / 3.0};
    return (d >= 0 && d <= 3) ? c[d] : 0.0;
  }
  if (p >= n - 2) {
    const double c[4] = {-1.0 / 3.0, 1.5, -3.0, 11.0 / 6.0};
    return (d >= -3 && d <= 0) ? c[d + 3] : 0.0;
  }
  const double c[5] = {1.0 / 12.0, -2.0 / 3.0, 0.0, 2.0 / 3.0, -1.0 / 12.0};
  return (d >= -2 && d <= 2) ? c[d + 2] : 0.0;
}

// Edge padding as a matrix: padded entry q takes original cell clamp(q-1).
__device__ __forceinline__ double pad_coef(int q, int i, int N) {
  int s = q - 1;
  s = s < 0 ? 0 : (s > N - 1 ? N - 1 : s);
  return s == i ? 1.0 : 0.0;
}

__device__ __forceinline__ int adj_slot(int i, int N) {
  if (N <= kAdjSlots) return i;
  if (i < kAdjB) return i;
  if (i >= N - kAdjB) return i - (N - kAdjSlots);
  return -1;
}

// A1[j][j + d] and A2[j][j + d] (0 outside the band or the grid).
struct AxisOps {
  const AdjAxisTab* T;
  int N;
  __device__ float a1(int j, int d) const {
    const int s = adj_slot(j, N);
    if (s >= 0) return T->a1[s][d + kAdjBand];
    const float c[5] = {1.0f / 12.0f, -2.0f / 3.0f, 0.0f, 2.0f / 3.0f, -1.0f / 12.0f};
    return (d >= -2 && d <= 2) ? c[d + 2] : 0.0f;
  }
  __device__ float a2(int j, int d) const {
    const int s = adj_slot(j, N);
    if (s >= 0) return T->a2[s][d + kAdjBand];
    const float c[9] = {1.0f / 144.0f, -1.0f / 9.0f, 4.0f / 9.0f, 1.0f / 9.0f, -65.0f / 72.0f,
                        1.0f / 9.0f,   4.0f / 9.0f,  -1.0f / 9.0f, 1.0f / 144.0f};
    return (d >= -4 && d <= 4) ? c[d + 4] : 0.0f;
  }
};


constexpr int kAdjThreads = 256;

// Central stencils of the untabulated rows: A1 = G (d = -2..2) and A2 = G G
// (d = -4..4).
#define FK_A1(d) ((d) == -2 ? 1.0f / 12.0f : (d) == -1 ? -2.0f / 3.0f : (d) == 1 ? 2.0f / 3.0f : (d) == 2 ? -1.0f / 12.0f : 0.0f)
#define FK_A2(d)                                                                          \
  ((d) == 0 ? -65.0f / 72.0f                                                              \
            : ((d) == 1 || (d) == -1) ? 1.0f / 9.0f                                       \
            : ((d) == 2 || (d) == -2) ? 4.0f / 9.0f                                       \
            : ((d) == 3 || (d) == -3) ? -1.0f / 9.0f                                      \
            : ((d) == 4 || (d) == -4) ? 1.0f / 144.0f : 0.0f)

// A1[j][j + d] and A2[j][j + d] without branches on the memory path: the
// table entry is always loaded (slot clamped into range) and then selected,
// so unrolled loops issue every load up front. Zero when j is off the axis.
__device__ __forceinline__ void band_coef(const AdjAxisTab& T, int j, int d, int N, float& a1,
                                          float& a2) {
  const bool on = j >= 0 && j < N;
  const int s = adj_slot(on ? j : 0, N);
  const int sc = s < 0 ? 0 : s;
  const float t1 = T.a1[sc][d + kAdjBand], t2 = T.a2[sc][d + kAdjBand];
  a1 = !on ? 0.0f : (s >= 0 ? t1 : FK_A1(d));
  a2 = !on ? 0.0f : (s >= 0 ? t2 : FK_A2(d));
}

// Row j of an axis of N cells uses the central stencils (not a table row).
__device__ __forceinline__ bool central_row(int j, int N) { return adj_slot(j, N) < 0; }
// Every row within the band of j is central, so (A^T q)(j) uses the
// transposed central stencils only.
__device__ __forceinline__ bool central_cone(int j, int N) {
  return N > kAdjSlots && j >= kAdjB + kAdjBand && j < N - kAdjB - kAdjBand;
}

// Fields the reverse step derives from u and the incoming cotangent before
// applying the transposed operators: L = lam_u (+ injected cotangent),
// lap = A2x u + A2y u, Q2 = D L, Q1x = D_x L, Px = L (A1x u), Q1y = D_y L,
// Py = L (A1y u). Fields read together are stored as pairs.
struct AdjPre {
  float2* Llap;  // (L, lap)
  float2* QPx;   // (Q1x, Px)
  float2* QPy;   // (Q1y, Py)
  float* Q2;
};

// Inputs of the derived fields of one cell: u at offsets -5..5 along both
// axes (clamped into the grid; the coefficients there are 0), D, D_x, D_y.
struct PreIn {
  float xs[2 * kAdjBand + 1], ys[2 * kAdjBand + 1];
  float D, Dx, Dy;
};

__device__ __forceinline__ void pre_load(const float* __restrict__ u,
                                         const float* __restrict__ D,
                                         const float* __restrict__ Dx,
                                         const float* __restrict__ Dy, int i, int c, int H,
                                         int W, PreIn& in) {
  const size_t idx = (size_t)i * W + c;
#pragma unroll
  for (int d = -kAdjBand; d <= kAdjBand; ++d) {
    in.xs[d + kAdjBand] = u[(size_t)clampi(i + d, 0, H - 1) * W + c];
    in.ys[d + kAdjBand] = u[(size_t)i * W + clampi(c + d, 0, W - 1)];
  }
  in.D = D[idx];
  in.Dx = Dx[idx];
  in.Dy = Dy[idx];
}

// The derived fields of one cell, given its cotangent L (lam_u plus any
// injected cotangent) and the loaded inputs. lap is the axis-0 sum plus the
// axis-1 sum, the order k_adj_main uses too.
// Per-axis pieces of the derived fields of the previous step: A1 u and A2 u
// at position i of an axis of N cells, from u at offsets -5..5 along the axis
// (clamped into the grid; the coefficients there are 0).
__device__ __forceinline__ void axis_pre(const float* __restrict__ xs, int i, int N,
                                         const AdjAxisTab& T, float& g, float& lap) {
  g = 0.0f;
  lap = 0.0f;
  if (central_row(i, N)) {
#pragma unroll
    for (int d = -4; d <= 4; ++d) {
      lap += FK_A2(d) * xs[d + kAdjBand];
      g += FK_A1(d) * xs[d + kAdjBand];
    }
  } else {
#pragma unroll
    for (int d = -kAdjBand; d <= kAdjBand; ++d) {
      float a1, a2;
      band_coef(T, i, d, N, a1, a2);
      const bool on = i + d >= 0 && i + d < N;
      g += on ? a1 * xs[d + kAdjBand] : 0.0f;
      lap += on ? a2 * xs[d + kAdjBand] : 0.0f;
    }
  }
}


__device__ __forceinline__ void pre_store(const PreIn& in, float L, int i, int c, int H, int W,
                                          const AdjAxisTab* tab, const AdjPre& o) {
  const size_t idx = (size_t)i * W + c;
  float gx, lapx, gy, lapy;
  axis_pre(in.xs, i, H, tab[0], gx, lapx);
  axis_pre(in.ys, c, W, tab[1], gy, lapy);
  o.Llap[idx] = make_float2(L, lapx + lapy);
  o.QPx[idx] = make_float2(in.Dx * L, L * gx);
  o.QPy[idx] = make_float2(in.Dy * L, L * gy);
  o.Q2[idx] = in.D * L;
}

// The boundary tables of both axes staged in LDS by the whole block (one
// barrier) when the block holds a cell within 13 cells of an edge: the edge
// paths then read coefficients from LDS and only their field loads go to
// memory, all independent of each other. Interior blocks (the condition is
// block-uniform) skip the staging and never read the tables.
__device__ __forceinline__ void stage_tables(const AdjAxisTab* __restrict__ tab, int H, int W,
                                             int rows, AdjAxisTab* s_tab, int bx, int by) {
  const int r0 = by * rows, c0 = bx * 64;
  const bool interior = central_cone(r0, H) && central_cone(r0 + rows - 1, H) &&
                        central_cone(c0, W) && central_cone(c0 + 63, W);
  if (interior) return;
  const float* src = (const float*)tab;
  float* dst = (float*)s_tab;
  constexpr int n = (int)(2 * sizeof(AdjAxisTab) / sizeof(float));
  for (int k = threadIdx.x; k < n; k += blockDim.x) dst[k] = src[k];
  __syncthreads();
}

// The derived fields of one cell of the block at (bx, by) of k_adj_pre's grid
// (64 columns x 4 rows): L = lam_u (+ injected cotangent) and the fields of
// pre_store. `tab` is the block's LDS.
__device__ __forceinline__ void adj_pre_tile(const float* __restrict__ u,
                                             const float* __restrict__ D,
                                             const float* __restrict__ Dx,
                                             const float* __restrict__ Dy,
                                             const float* __restrict__ lu,
                                             const float* __restrict__ inj, int H, int W,
                                             const AdjAxisTab* __restrict__ tab_g,
                                             const AdjPre& o, int bx, int by, AdjAxisTab* tab) {
  stage_tables(tab_g, H, W, 4, tab, bx, by);
  const int c = bx * 64 + (threadIdx.x & 63);
  const int i = by * 4 + (threadIdx.x >> 6);
  if (i >= H || c >= W) return;
  const size_t idx = (size_t)i * W + c;
  PreIn in;
  pre_load(u, D, Dx, Dy, i, c, H, W, in);
  pre_store(in, inj ? lu[idx] + inj[idx] : lu[idx], i, c, H, W, tab, o);
}

// What the reverse step of t+k-1 needs from the step of t+k: the state
// x_{t+k-1}'s u and the cotangent injected into x_{t+k} (may be null).
struct AdjNext {
  const float* u_prev;
  const float* inj;
  const float *D, *Dx, *Dy;
  AdjPre pre;
};

// Transposed operators along one axis at position i (stride between
// neighbours `step` elements, base index `idx` of position i):
// s = A2^T(Q2) + A1^T(Q1), e = A1^T(P), with (Q1, P) stored as pairs.
__device__ __forceinline__ void axis_transpose(const float* __restrict__ Q2,
                                               const float2* __restrict__ QP, size_t idx,
                                               ptrdiff_t step, int i, int N, const AdjAxisTab& T,
                                               float& sv, float& ev) {
  sv = 0.0f;
  ev = 0.0f;
  if (central_cone(i, N)) {
#pragma unroll
    for (int d = -4; d <= 4; ++d) {
      const size_t j = idx - (ptrdiff_t)d * step;
      sv += FK_A2(d) * Q2[j];
      if (d >= -2 && d <= 2 && d != 0) {
        const float2 q = QP[j];
        sv += FK_A1(d) * q.x;
        ev += FK_A1(d) * q.y;
      }
    }
  } else {
#pragma unroll
    for (int d = -kAdjBand; d <= kAdjBand; ++d) {
      const int jr = i - d;
      float a1, a2;
      band_coef(T, jr, d, N, a1, a2);  // 0 when jr is off the axis
      const size_t j = idx + (ptrdiff_t)(clampi(jr, 0, N - 1) - i) * step;
      const float2 q = QP[j];
      sv += a2 * Q2[j] + a1 * q.x;
      ev += a1 * q.y;
    }
  }
}

// Index of x wave `xw` (0 or 1) of the block at (bx, by) of k_adj_main's grid
// (gx blocks wide) among all x waves of the grid: one tau accumulator each.
__device__ __forceinline__ size_t adj_wave_index(int bx, int by, int gx, int xw) {
  return ((size_t)by * gx + bx) * 2 + (size_t)xw;
}

// Number of x waves (tau accumulators) of k_adj_main on an H x W grid.
__host__ __device__ inline size_t adj_n_waves(int H, int W) {
  return (size_t)((W + 63) / 64) * (size_t)((H + 1) / 2) * 2;
}

// Stimulus fields of one step as an active list (count(), on(m), field(m)).
struct AdjStimList {
  const StimArgs& sa;
  __device__ int count() const { return sa.n; }
  __device__ bool on(int) const { return true; }
  __device__ const float* field(int m) const { return sa.f[m]; }
};

// One reverse step from the derived fields, two threads per cell in
// different waves of a block of 64 columns x 2 rows: the "x" thread loads and
// applies the transposed operators along axis 0, evaluates the reaction's
// adjoint and writes every output; the "y" thread applies those along axis 1
// (and, with NEXT, derives the axis-1 part of the previous step's fields) and
// hands its 4 results over through LDS (one barrier). Twice the waves and half
// the dependency chain per wave: the step is latency-bound on small grids.
// The tau-gradient contributions of the 64 cells of an x wave are summed
// across the wave (a fixed butterfly) and lane 0 adds the three sums to the
// wave's own double accumulator g_tau[wave]: 24 bytes per 64 cells per step,
// deterministic (each accumulator is touched by one wave, in step order).
// NEXT: also write the derived fields of the previous step's reverse (they
// depend on the new cotangent only at the same cell), so a segment needs one
// launch per step.
// The block at (bx, by) of a grid gx blocks wide; `tab` and `xch` are the
// block's LDS. Every thread of the block calls it (one barrier inside).
template <bool NEXT, class SV>
__device__ __forceinline__ void adj_main_tile(
    const Coef& c, const float* __restrict__ v, const float* __restrict__ w,
    const float* __restrict__ u, int H, int W, const SV& sv, float dt, float inv_dx2,
    const float* __restrict__ lv, const float* __restrict__ lw, const AdjPre& pre,
    float* __restrict__ lv_o, float* __restrict__ lw_o, float* __restrict__ lu_o,
    float* __restrict__ gD, double* __restrict__ g_tau, const AdjAxisTab* __restrict__ tab_g,
    const AdjNext& nx, int bx, int by, int gx, AdjAxisTab* tab, float4* xch) {
  stage_tables(tab_g, H, W, 2, tab, bx, by);
  const bool ythread = threadIdx.x >= kAdjThreads / 2;  // wave-uniform
  const int q = threadIdx.x & (kAdjThreads / 2 - 1);
  const int col = bx * 64 + (q & 63);
  const int i = by * 2 + (q >> 6);
  const bool in = i < H && col < W;
  const size_t idx = in ? (size_t)i * W + col : 0;

  // x-thread state (kept across the barrier)
  float xs[2 * kAdjBand + 1];
  float nD = 0.0f, nDx = 0.0f, nDy = 0.0f, inj = 0.0f;
  float vv = 0.0f, ww = 0.0f, uu = 0.0f, lvv = 0.0f, lww = 0.0f, L = 0.0f, lap = 0.0f;
  float sx = 0.0f, ex = 0.0f, gx_n = 0.0f, lapx_n = 0.0f;
  float gd = 0.0f;
  bool stim = false;
  if (ythread) {
    float4 r = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (in) {
      float ys[2 * kAdjBand + 1];
      if (NEXT) {
        const float* urow = nx.u_prev + (size_t)i * W;
#pragma unroll
        for (int d = -kAdjBand; d <= kAdjBand; ++d)
          ys[d + kAdjBand] = urow[clampi(col + d, 0, W - 1)];
      }
      axis_transpose(pre.Q2, pre.QPy, idx, 1, col, W, tab[1], r.x, r.y);
      if (NEXT) axis_pre(ys, col, W, tab[1], r.z, r.w);
    }
    xch[q] = r;
  } else if (in) {
    // x thread: everything that does not need the y results
    if (NEXT) {  // issued first: independent of everything below
#pragma unroll
      for (int d = -kAdjBand; d <= kAdjBand; ++d)
        xs[d + kAdjBand] = nx.u_prev[(size_t)clampi(i + d, 0, H - 1) * W + col];
      nD = nx.D[idx];
      nDx = nx.Dx[idx];
      nDy = nx.Dy[idx];
      if (nx.inj) inj = nx.inj[idx];
    }
    vv = v[idx];
    ww = w[idx];
    uu = u[idx];
    const float2 Llap = pre.Llap[idx];
    L = Llap.x;
    lap = Llap.y;
    lvv = lv[idx];
    lww = lw[idx];
    gd = gD[idx];
    for (int k = 0; k < sv.count(); ++k) stim = stim || (sv.on(k) && sv.field(k)[idx] != 0.0f);
    axis_transpose(pre.Q2, pre.QPx, idx, W, i, H, tab[0], sx, ex);
    if (NEXT) axis_pre(xs, i, H, tab[0], gx_n, lapx_n);
  }
  __syncthreads();  // every thread of the block passes this one barrier
  if (ythread) return;  // whole waves: the x waves stay complete for the tau sums
  float t_d = 0.0f, t_r = 0.0f, t_si = 0.0f;  // this cell's tau-gradient terms
  if (in) {
    const float4 yr = xch[q];
    const float sy = yr.x, ey = yr.y;

    // --- reaction, gates as fixed masks -----------------------------------
    const bool p = uu >= c.V_c;
    const float inv_tvm = (uu >= c.V_v) ? c.inv_tv2m : c.inv_tv1m;
    const float dfv_dv = p ? -c.inv_tvp : -inv_tvm;
    const float dfw_dw = p ? -c.inv_twp : -c.inv_twm;
    const float S = half_one_plus_tanh(c, uu);  // (1 + tanh(k (u - V_csi))) / 2
    // dS/du = 2k S (1 - S); kk = -2 k log2(e)
    const float dS = -c.kk * 0.69314718055994531f * S * (1.0f - S);
    // j_ion = -(j_fi + j_so + j_si) / Cm
    const float djdv = p ? (uu - c.V_c) * (1.0f - uu) * c.inv_td * c.inv_cm : 0.0f;
    const float djdw = S * c.inv_tsi * c.inv_cm;
    const float djdu =
        -((p ? -vv * (1.0f + c.V_c - 2.0f * uu) * c.inv_td : 0.0f) + (p ? 0.0f : c.inv_t0) -
          ww * dS * c.inv_tsi) *
        c.inv_cm;
    // stimulus replaces j_ion: no cotangent reaches it where active
    const float lus = stim ? 0.0f : L;

    lv_o[idx] = lvv + dt * (dfv_dv * lvv + djdv * lus);
    lw_o[idx] = lww + dt * (dfw_dw * lww + djdw * lus);
    const float lu_new = L + dt * (djdu * lus + (sx + sy) * inv_dx2);
    lu_o[idx] = lu_new;
    gD[idx] = gd + dt * (L * lap + ex + ey) * inv_dx2;
    // tau gradients: d j_ion / d tau
    const float djd_td = p ? -vv * (uu - c.V_c) * (1.0f - uu) * c.inv_td * c.inv_td * c.inv_cm
                           : 0.0f;
    const float djd_tr = p ? c.inv_tr * c.inv_tr * c.inv_cm : 0.0f;
    const float djd_tsi = -ww * S * c.inv_tsi * c.inv_tsi * c.inv_cm;
    t_d = dt * djd_td * lus;
    t_r = dt * djd_tr * lus;
    t_si = dt * djd_tsi * lus;
    if (NEXT) {  // derived fields of the previous step, at this cell
      const float Ln = lu_new + inj;
      nx.pre.Llap[idx] = make_float2(Ln, lapx_n + yr.w);
      nx.pre.QPx[idx] = make_float2(nDx * Ln, Ln * gx_n);
      nx.pre.QPy[idx] = make_float2(nDy * Ln, Ln * yr.z);
      nx.pre.Q2[idx] = nD * Ln;
    }
  }
  // every lane of the x wave takes part; lanes off the grid add 0
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) {
    t_d += __shfl_xor(t_d, o, 64);
    t_r += __shfl_xor(t_r, o, 64);
    t_si += __shfl_xor(t_si, o, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    double* a = g_tau + 3 * adj_wave_index(bx, by, gx, threadIdx.x >> 6);
    a[0] += (double)t_d;
    a[1] += (double)t_r;
    a[2] += (double)t_si;
  }
}

// Blocks of the first pass of the tau reduction (k_adj_reduce).
constexpr int kRedBlocks = 512;

// Workspace layout: [tables (2 axes)][Dx][Dy][tau accumulators: 3 doubles per x wave of
// k_adj_main][cotangent ping-pong: 3 fields][derived fields: 2 sets of 7 (8 floats per
// cell)][tau block sums]
struct AdjLayout {
  size_t tab, dx, dy, part, lam, pre, red, total;
};

inline AdjLayout adj_layout(int H, int W) {
  AdjLayout L;
  const size_t cells = (size_t)H * W;
  size_t o = 0;
  auto take = [&](size_t bytes) {
    const size_t at = o;
    o += (bytes + 255) / 256 * 256;
    return at;
  };
  L.tab = take(2 * sizeof(AdjAxisTab));
  L.dx = take(cells * sizeof(float));
  L.dy = take(cells * sizeof(float));
  L.part = take(3 * adj_n_waves(H, W) * sizeof(double));
  L.lam = take(3 * cells * sizeof(float));
  L.pre = take(2 * 8 * cells * sizeof(float));
  L.red = take(3 * kRedBlocks * sizeof(double));
  L.total = o;
  return L;
}

// Set `set` of the derived fields: three float2 fields, then Q2 (the pairs
// stay 8-byte aligned for any n).
inline AdjPre pre_fields(const AdjLayout& L, char* wk, int set, size_t n) {
  float* pf = (float*)(wk + L.pre) + (size_t)set * 8 * n;
  return AdjPre{(float2*)pf, (float2*)(pf + 2 * n), (float2*)(pf + 4 * n), pf + 6 * n};
}

}  // namespace fk
