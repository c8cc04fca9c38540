// Adjoint (reverse mode) of one forward-Euler step, for gradients of a loss
// with respect to the diffusivity map D and the time constants tau_d, tau_r,
// tau_si.
//
// Forward step (cardiax/solve.py:26-70): x' = x + dt f(x), f = (f_v, f_w, f_u),
//   f_u = [D (A2x u + A2y u) + (A1x D)(A1x u) + (A1y D)(A1y u)] / dx^2 + J,
// where A1 = crop(G(pad(.))) and A2 = crop(G(G(pad(.)))) are the effective
// per-axis operators of the reference's `gradient` with edge padding, and J is
// j_ion, replaced by the stimulus where one is active. Reverse step, with the
// Heaviside gates p = [u >= V_c], q = [u >= V_v] held as fixed masks and the
// stimulus selection routing no cotangent to j_ion (the semantics jax.grad
// gives greater_equal and where):
//   lam_v' = lam_v + dt (df_v/dv lam_v + dJ/dv lam_u*)
//   lam_w' = lam_w + dt (df_w/dw lam_w + dJ/dw lam_u*)
//   lam_u' = lam_u + dt (dJ/du lam_u* + [A2x^T(D lam_u) + A2y^T(D lam_u)
//                        + A1x^T(D_x lam_u) + A1y^T(D_y lam_u)] / dx^2)
//   gD    += dt [lam_u (A2x u + A2y u) + A1x^T(lam_u A1x u) + A1y^T(lam_u A1y u)] / dx^2
//   g_tau += dt sum(dJ/dtau lam_u*)            (tau_d, tau_r, tau_si)
// with lam_u* = lam_u outside active stimulus cells and 0 inside.
//
// The operators' entries away from the edges are the central stencils; the 8
// rows nearest each edge are tabulated per call by k_adj_tables, which
// composes the reference's gradient rules exactly (any N >= 3).
#include <hip/hip_runtime.h>

#include "cardiax_amd.h"
#include "fk_device.h"
#include "fk_host.h"

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

// A1[j][i] and A2[j][i] for the tabulated rows of both axes.
__global__ void k_adj_tables(int H, int W, AdjAxisTab* tab) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int per_axis = kAdjSlots * (2 * kAdjBand + 1);
  if (t >= 2 * per_axis) return;
  const int axis = t / per_axis, r = t % per_axis;
  const int slot = r / (2 * kAdjBand + 1), d = r % (2 * kAdjBand + 1) - kAdjBand;
  const int N = axis == 0 ? H : W, n = N + 2;
  // row j of the slot
  int j;
  if (N <= kAdjSlots) j = slot;
  else j = slot < kAdjB ? slot : N - kAdjSlots + slot;
  double a1 = 0.0, a2 = 0.0;
  const int i = j + d;
  if (j < N && i >= 0 && i < N) {
    const int pj = j + 1;  // padded row of original j
    for (int q = pj - 3; q <= pj + 3; ++q) {
      if (q < 0 || q >= n) continue;
      const double gq = g_coef(pj, q, n);
      if (gq == 0.0) continue;
      a1 += gq * pad_coef(q, i, N);
      // A2 = crop G G pad: row pj of G times column i of G pad
      double gp = 0.0;
      for (int s = q - 3; s <= q + 3; ++s)
        if (s >= 0 && s < n) gp += g_coef(q, s, n) * pad_coef(s, i, N);
      a2 += gq * gp;
    }
  }
  AdjAxisTab& T = tab[axis];
  if (slot < kAdjSlots) {
    T.a1[slot][d + kAdjBand] = (float)a1;
    T.a2[slot][d + kAdjBand] = (float)a2;
  }
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

// Static first derivatives of D: (A1x D, A1y D).
__global__ void k_adj_dgrad(const float* __restrict__ D, int H, int W,
                            const AdjAxisTab* __restrict__ tab, float* __restrict__ Dx,
                            float* __restrict__ Dy) {
  const int c = blockIdx.x * 64 + threadIdx.x, i = blockIdx.y * 4 + threadIdx.y;
  if (i >= H || c >= W) return;
  const AxisOps ox{&tab[0], H}, oy{&tab[1], W};
  float sx = 0.0f, sy = 0.0f;
  for (int d = -kAdjBand; d <= kAdjBand; ++d) {
    if (i + d >= 0 && i + d < H) sx += ox.a1(i, d) * D[(size_t)(i + d) * W + c];
    if (c + d >= 0 && c + d < W) sy += oy.a1(c, d) * D[(size_t)i * W + c + d];
  }
  Dx[(size_t)i * W + c] = sx;
  Dy[(size_t)i * W + c] = sy;
}

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
                                             int rows, AdjAxisTab* s_tab) {
  const int r0 = blockIdx.y * rows, c0 = blockIdx.x * 64;
  const bool interior = central_cone(r0, H) && central_cone(r0 + rows - 1, H) &&
                        central_cone(c0, W) && central_cone(c0 + 63, W);
  if (interior) return;
  const float* src = (const float*)tab;
  float* dst = (float*)s_tab;
  constexpr int n = (int)(2 * sizeof(AdjAxisTab) / sizeof(float));
  for (int k = threadIdx.x; k < n; k += blockDim.x) dst[k] = src[k];
  __syncthreads();
}

__global__ void __launch_bounds__(kAdjThreads)
k_adj_pre(const float* __restrict__ u, const float* __restrict__ D,
          const float* __restrict__ Dx, const float* __restrict__ Dy,
          const float* __restrict__ lu, const float* __restrict__ inj, int H, int W,
          const AdjAxisTab* __restrict__ tab_g, AdjPre o) {
  __shared__ AdjAxisTab tab[2];
  stage_tables(tab_g, H, W, 4, tab);
  const int c = blockIdx.x * 64 + (threadIdx.x & 63);
  const int i = blockIdx.y * 4 + (threadIdx.x >> 6);
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

// Index of x wave `xw` (0 or 1) of the current block of k_adj_main among all
// x waves of its grid: one tau accumulator each.
__device__ __forceinline__ size_t adj_wave_index(int xw) {
  return ((size_t)blockIdx.y * gridDim.x + blockIdx.x) * 2 + (size_t)xw;
}

// Number of x waves (tau accumulators) of k_adj_main on an H x W grid.
__host__ __device__ inline size_t adj_n_waves(int H, int W) {
  return (size_t)((W + 63) / 64) * (size_t)((H + 1) / 2) * 2;
}

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
template <bool NEXT>
__global__ void __launch_bounds__(kAdjThreads)
k_adj_main(Coef c, const float* __restrict__ v, const float* __restrict__ w,
           const float* __restrict__ u, int H, int W, StimArgs sa, float dt, float inv_dx2,
           const float* __restrict__ lv, const float* __restrict__ lw, AdjPre pre,
           float* __restrict__ lv_o, float* __restrict__ lw_o, float* __restrict__ lu_o,
           float* __restrict__ gD, double* __restrict__ g_tau,
           const AdjAxisTab* __restrict__ tab_g, AdjNext nx) {
  __shared__ AdjAxisTab tab[2];
  __shared__ float4 xch[kAdjThreads / 2];  // y thread -> x thread of each cell
  stage_tables(tab_g, H, W, 2, tab);
  const bool ythread = threadIdx.x >= kAdjThreads / 2;  // wave-uniform
  const int q = threadIdx.x & (kAdjThreads / 2 - 1);
  const int col = blockIdx.x * 64 + (q & 63);
  const int i = blockIdx.y * 2 + (q >> 6);
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
    for (int k = 0; k < sa.n; ++k) stim = stim || (sa.f[k][idx] != 0.0f);
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
    double* a = g_tau + 3 * adj_wave_index(threadIdx.x >> 6);
    a[0] += (double)t_d;
    a[1] += (double)t_r;
    a[2] += (double)t_si;
  }
}

__global__ void k_adj_zero(float* a, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    a[i] = 0.0f;
}

constexpr int kRedBlocks = 512;

// Sums of the three per-wave tau accumulators (n triples of doubles): each
// block reduces a grid-stride slice into part[block][0..2] (k_adj_reduce),
// then one block adds the block sums (k_adj_reduce_final). Fixed order:
// deterministic.
__global__ void __launch_bounds__(256)
k_adj_reduce(const double* __restrict__ acc, size_t n, double* __restrict__ part) {
  __shared__ double s[3][256];
  double a0 = 0.0, a1 = 0.0, a2 = 0.0;
  for (size_t b = (size_t)blockIdx.x * blockDim.x + threadIdx.x; b < n;
       b += (size_t)gridDim.x * blockDim.x) {
    a0 += acc[3 * b];
    a1 += acc[3 * b + 1];
    a2 += acc[3 * b + 2];
  }
  s[0][threadIdx.x] = a0;
  s[1][threadIdx.x] = a1;
  s[2][threadIdx.x] = a2;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o)
      for (int k = 0; k < 3; ++k) s[k][threadIdx.x] += s[k][threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int k = 0; k < 3; ++k) part[3 * (size_t)blockIdx.x + k] = s[k][0];
}

__global__ void __launch_bounds__(256)
k_adj_reduce_final(const double* __restrict__ part, int n_part, float* g_tau) {
  __shared__ double s[3][256];
  for (int k = 0; k < 3; ++k) {
    double a = 0.0;
    for (int b = threadIdx.x; b < n_part; b += blockDim.x) a += part[3 * b + k];
    s[k][threadIdx.x] = a;
  }
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o)
      for (int k = 0; k < 3; ++k) s[k][threadIdx.x] += s[k][threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int k = 0; k < 3; ++k) g_tau[k] = (float)s[k][0];
}

// Workspace layout: [tables (2 axes)][Dx][Dy][tau accumulators: 3 doubles per x wave of
// k_adj_main][cotangent ping-pong: 3 fields][derived fields: 2 sets of 7 (8 floats per
// cell)][tau block sums]
struct AdjLayout {
  size_t tab, dx, dy, part, lam, pre, red, total;
};

static AdjLayout adj_layout(int H, int W) {
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
static AdjPre pre_fields(const AdjLayout& L, char* wk, int set, size_t n) {
  float* pf = (float*)(wk + L.pre) + (size_t)set * 8 * n;
  return AdjPre{(float2*)pf, (float2*)(pf + 2 * n), (float2*)(pf + 4 * n), pf + 6 * n};
}

static hipError_t launch_pre(const AdjLayout& L, char* wk, const float* D, const float* u,
                             int H, int W, const float* lu, const float* inj, int set,
                             hipStream_t s) {
  const size_t n = (size_t)H * W;
  const dim3 grid((W + 63) / 64, (H + 3) / 4);
  k_adj_pre<<<grid, kAdjThreads, 0, s>>>(u, D, (const float*)(wk + L.dx),
                                         (const float*)(wk + L.dy), lu, inj, H, W,
                                         (const AdjAxisTab*)(wk + L.tab),
                                         pre_fields(L, wk, set, n));
  return hipGetLastError();
}

// Reverse step from derived-field set `set`: cotangents lam_in (v, w; u comes
// from the fields) -> lam_out. With u_prev, also writes set 1 - set for the
// step before (u_prev = that step's state u, inj = its injected cotangent).
static hipError_t launch_main(const AdjLayout& L, char* wk, const Coef& c, const float* D,
                              const float* v, const float* w, const float* u, int H, int W,
                              const StimArgs& sa, float dt, float inv_dx2,
                              const float* const lam_in[3], float* const lam_out[3], float* gD,
                              int set, const float* u_prev, const float* inj, hipStream_t s) {
  const size_t n = (size_t)H * W;
  const AdjAxisTab* tab = (const AdjAxisTab*)(wk + L.tab);
  const dim3 grid((W + 63) / 64, (H + 1) / 2);  // 64 x 2 cells, two threads each
  const AdjNext nx{u_prev, inj, D, (const float*)(wk + L.dx), (const float*)(wk + L.dy),
                   pre_fields(L, wk, 1 - set, n)};
  if (u_prev)
    k_adj_main<true><<<grid, kAdjThreads, 0, s>>>(
        c, v, w, u, H, W, sa, dt, inv_dx2, lam_in[0], lam_in[1], pre_fields(L, wk, set, n),
        lam_out[0], lam_out[1], lam_out[2], gD, (double*)(wk + L.part), tab, nx);
  else
    k_adj_main<false><<<grid, kAdjThreads, 0, s>>>(
        c, v, w, u, H, W, sa, dt, inv_dx2, lam_in[0], lam_in[1], pre_fields(L, wk, set, n),
        lam_out[0], lam_out[1], lam_out[2], gD, (double*)(wk + L.part), tab, nx);
  return hipGetLastError();
}

}  // namespace fk

using namespace fk;

extern "C" {

size_t fk_adjoint_workspace_bytes(int32_t H, int32_t W) {
  if (H <= 0 || W <= 0) return 0;
  return adj_layout(H, W).total;
}

int32_t fk_adjoint_prepare(const float* D, int32_t H, int32_t W, void* work, size_t work_bytes,
                           void* stream) {
  ErrorScope es;
  if (!check_grid(H, W) || !check_ptrs({D, work})) return last_code();
  const AdjLayout L = adj_layout(H, W);
  if (work_bytes < L.total)
    return set_error(FK_ERR_ARG, "adjoint workspace too small: need %zu bytes, got %zu",
                     L.total, work_bytes);
  hipStream_t s = (hipStream_t)stream;
  char* wk = (char*)work;
  AdjAxisTab* tab = (AdjAxisTab*)(wk + L.tab);
  const int n_tab = 2 * kAdjSlots * (2 * kAdjBand + 1);
  k_adj_tables<<<(n_tab + 255) / 256, 256, 0, s>>>(H, W, tab);
  HIP_CHECK(hipGetLastError());
  const dim3 block(64, 4), grid((W + 63) / 64, (H + 3) / 4);
  k_adj_dgrad<<<grid, block, 0, s>>>(D, H, W, tab, (float*)(wk + L.dx), (float*)(wk + L.dy));
  HIP_CHECK(hipGetLastError());
  k_adj_zero<<<256, 256, 0, s>>>((float*)(wk + L.part), 6 * adj_n_waves(H, W));
  HIP_CHECK(hipGetLastError());
  return FK_OK;
}

int32_t fk_adjoint_step(const fk_params* params, const float* D, const float* v,
                        const float* w, const float* u, int32_t H, int32_t W,
                        const fk_stimulus* stimuli, int32_t n_stimuli, double t, float dt,
                        float dx, const float* lam_v, const float* lam_w, const float* lam_u,
                        float* lam_v_out, float* lam_w_out, float* lam_u_out, float* grad_D,
                        void* work, size_t work_bytes, void* stream) {
  ErrorScope es;
  if (!check_params(params) || !check_grid(H, W) || !check_stimuli(stimuli, n_stimuli) ||
      !check_ptrs({D, v, w, u, lam_v, lam_w, lam_u, lam_v_out, lam_w_out, lam_u_out, grad_D,
                   work}) ||
      !check_scalar("t", t) || !check_scalar("dt", dt) || !check_dx(dx))
    return last_code();
  const AdjLayout L = adj_layout(H, W);
  if (work_bytes < L.total)
    return set_error(FK_ERR_ARG, "adjoint workspace too small: need %zu bytes, got %zu",
                     L.total, work_bytes);
  StimSchedule sched(stimuli, n_stimuli);
  const Coef c = make_coef(*params);
  const float inv_dx2 = (float)(1.0 / ((double)dx * (double)dx));
  const float* lin[3] = {lam_v, lam_w, lam_u};
  float* lout[3] = {lam_v_out, lam_w_out, lam_u_out};
  hipStream_t s = (hipStream_t)stream;
  HIP_CHECK(launch_pre(L, (char*)work, D, u, H, W, lam_u, nullptr, 0, s));
  HIP_CHECK(launch_main(L, (char*)work, c, D, v, w, u, H, W, sched.at(t), dt, inv_dx2, lin, lout,
                        grad_D, 0, nullptr, nullptr, s));
  return FK_OK;
}

int32_t fk_adjoint_segment(const fk_params* params, const float* D, int32_t H, int32_t W,
                           const fk_stimulus* stimuli, int32_t n_stimuli, double t,
                           int32_t n_steps, float dt, float dx, const float* states,
                           const float* const* inject, float* lam_v, float* lam_w,
                           float* lam_u, float* grad_D, void* work, size_t work_bytes,
                           void* stream) {
  ErrorScope es;
  if (!check_params(params) || !check_grid(H, W) || !check_stimuli(stimuli, n_stimuli) ||
      !check_ptrs({D, lam_v, lam_w, lam_u, grad_D, work}) || !check_scalar("t", t) ||
      !check_scalar("dt", dt) || !check_dx(dx))
    return last_code();
  if (n_steps < 0) return set_error(FK_ERR_ARG, "n_steps must be >= 0, got %d", n_steps);
  if (n_steps > 0 && states == nullptr) return set_error(FK_ERR_ARG, "null states pointer");
  if (t + (double)n_steps > 16777216.0)
    return set_error(FK_ERR_ARG, "time index must stay below 2^24");
  const AdjLayout L = adj_layout(H, W);
  if (work_bytes < L.total)
    return set_error(FK_ERR_ARG, "adjoint workspace too small: need %zu bytes, got %zu",
                     L.total, work_bytes);
  hipStream_t s = (hipStream_t)stream;
  char* wk = (char*)work;
  const size_t n = (size_t)H * W;
  float* alt = (float*)(wk + L.lam);
  float* bufs[2][3] = {{lam_v, lam_w, lam_u}, {alt, alt + n, alt + 2 * n}};
  StimSchedule sched(stimuli, n_stimuli);
  const Coef c = make_coef(*params);
  const float inv_dx2 = (float)(1.0 / ((double)dx * (double)dx));
  int cur = 0, set = 0;
  if (n_steps > 0) {
    const float* x = states + (size_t)(n_steps - 1) * 3 * n;
    HIP_CHECK(launch_pre(L, wk, D, x + 2 * n, H, W, lam_u,
                         inject ? inject[n_steps - 1] : nullptr, set, s));
  }
  for (int32_t k = n_steps - 1; k >= 0; --k) {
    const float* x = states + (size_t)k * 3 * n;
    const float* xp = k > 0 ? states + (size_t)(k - 1) * 3 * n : nullptr;
    const float* lin[3] = {bufs[cur][0], bufs[cur][1], bufs[cur][2]};
    HIP_CHECK(launch_main(L, wk, c, D, x, x + n, x + 2 * n, H, W, sched.at(t + (double)k), dt,
                          inv_dx2, lin, bufs[cur ^ 1], grad_D, set, xp ? xp + 2 * n : nullptr,
                          (xp && inject) ? inject[k - 1] : nullptr, s));
    cur ^= 1;
    set ^= 1;
  }
  if (cur == 1)
    for (int f = 0; f < 3; ++f)
      HIP_CHECK(hipMemcpyAsync(bufs[0][f], bufs[1][f], n * sizeof(float),
                               hipMemcpyDeviceToDevice, s));
  return FK_OK;
}

int32_t fk_adjoint_tau_grads(void* work, int32_t H, int32_t W, float* g_tau, void* stream) {
  ErrorScope es;
  if (!check_grid(H, W) || !check_ptrs({work, g_tau})) return last_code();
  const AdjLayout L = adj_layout(H, W);
  const size_t n = adj_n_waves(H, W);
  const int blocks = (int)std::min<size_t>((n + 255) / 256, (size_t)kRedBlocks);
  double* part = (double*)((char*)work + L.red);
  k_adj_reduce<<<blocks, 256, 0, (hipStream_t)stream>>>(
      (const double*)((char*)work + L.part), n, part);
  k_adj_reduce_final<<<1, 256, 0, (hipStream_t)stream>>>(part, blocks, g_tau);
  HIP_CHECK(hipGetLastError());
  return FK_OK;
}

}  // extern "C"
