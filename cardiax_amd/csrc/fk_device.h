// Device-side building blocks of the Fenton-Karma step shared by every kernel:
// the reaction (gates + ionic currents) and the reference's finite-difference
// rules along one axis.
//
// Reference: cardiax/solve.py:26-65 (step), :225-254 (gradient).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "cardiax_amd.h"

namespace fk {

// Per-launch constants derived from fk_params on the host (double math, then
// rounded once to float). Reciprocals replace the reference's divisions.
struct Coef {
  float V_c, V_v, V_csi;
  float kk;      // -2 k log2(e): exp(-2k(u - V_csi)) = exp2(u * kk + kk_off)
  float kk_off;  // -V_csi * kk
  float inv_tvp, inv_tv1m, inv_tv2m, inv_twp, inv_twm;
  float inv_td, inv_t0, inv_tr, inv_tsi, inv_cm;
};

// Fields of the stimuli active at one step, in list order (later entries
// override earlier ones where their field is nonzero).
struct StimArgs {
  int32_t n;
  const float* f[FK_MAX_STIMULI];
};

__device__ __forceinline__ int clampi(int x, int lo, int hi) {
  return x < lo ? lo : (x > hi ? hi : x);
}

// (1 + tanh(z)) / 2 = 1 / (1 + exp(-2z)), with z = k (u - V_csi). This form has
// no cancellation for any z and saturates to exactly 0 / 1 (exp2 -> inf / 0),
// which the MAXFLOAT sentinel sets (cardiax/params.py:43-58,78-93) rely on.
__device__ __forceinline__ float half_one_plus_tanh(const Coef& c, float u) {
#pragma clang fp contract(off)
  const float e = __builtin_amdgcn_exp2f(fmaf(u, c.kk, c.kk_off));
  return __builtin_amdgcn_rcpf(1.0f + e);
}

// Gates and ionic currents, cardiax/solve.py:35-42,57-58.
// p = [u >= V_c], q = [u >= V_v]; outputs dv, dw and j_ion (before stimulus).
// The gates are written as selects; the compiler turns some of them into
// branches, so a wave whose cells are all at rest skips the excited side.
// Contraction is off in the shared step code: every kernel that inlines it
// rounds identically, so batched and single-grid runs agree bit for bit.
__device__ __forceinline__ void reaction(const Coef& c, float v, float w, float u, float& dv,
                                         float& dw, float& jion) {
#pragma clang fp contract(off)
  const bool p = u >= c.V_c;
  // j_fi + j_so: p ? -v (u - V_c)(1 - u) / tau_d + 1 / tau_r : u / tau_0
  const float a = (u - c.V_c) * (1.0f - u);
  const float fo = p ? fmaf(-(v * a), c.inv_td, c.inv_tr) : u * c.inv_t0;
  // + j_si = -w S / tau_si, S = (1 + tanh(k (u - V_csi))) / 2
  const float s = fmaf(-(w * c.inv_tsi), half_one_plus_tanh(c, u), fo);
  const float inv_tvm = (u >= c.V_v) ? c.inv_tv2m : c.inv_tv1m;
  dv = p ? -v * c.inv_tvp : (1.0f - v) * inv_tvm;
  dw = p ? -w * c.inv_twp : (1.0f - w) * c.inv_twm;
  jion = -s * c.inv_cm;
}

// Stimulus replacement of cardiax/solve.py:257-271 at one cell.
__device__ __forceinline__ float stimulate_cell(const StimArgs& s, size_t idx, float jion) {
  float st = 0.0f;
  for (int k = 0; k < s.n; ++k) {
    const float f = s.f[k][idx];
    st = (f != 0.0f) ? f : st;
  }
  return (st != 0.0f) ? st : jion;
}

// Gradient coefficients of cardiax/solve.py:230-250.
#define FK_C_E0 (-11.0f / 6.0f)
#define FK_C_E1 (3.0f)
#define FK_C_E2 (-1.5f)
#define FK_C_E3 (1.0f / 3.0f)
#define FK_C_I1 (2.0f / 3.0f)
#define FK_C_I2 (-1.0f / 12.0f)

// 4th-order central difference at position q from values at q-2..q+2.
__device__ __forceinline__ float d1_interior(float m2, float m1, float p1, float p2) {
  return FK_C_I1 * (p1 - m1) + FK_C_I2 * (p2 - m2);
}
// 3rd-order one-sided differences: forward from q..q+3, backward from q-3..q.
__device__ __forceinline__ float d1_forward(float a0, float a1, float a2, float a3) {
  return FK_C_E0 * a0 + FK_C_E1 * a1 + FK_C_E2 * a2 + FK_C_E3 * a3;
}
__device__ __forceinline__ float d1_backward(float a3, float a2, float a1, float a0) {
  // a0 = value at q, a1 at q-1, ...: mirror of the forward rule.
  return -(FK_C_E0 * a0 + FK_C_E1 * a1 + FK_C_E2 * a2 + FK_C_E3 * a3);
}

// The same three rules times 12, with exact integer coefficients: kernels
// that keep derivatives in this scaling (12 per application) fold 1/144 into
// the grid factor and save one multiply per difference.
__device__ __forceinline__ float s1_interior(float m2, float m1, float p1, float p2) {
#pragma clang fp contract(off)
  return fmaf(8.0f, p1 - m1, m2 - p2);
}
__device__ __forceinline__ float s1_forward(float a0, float a1, float a2, float a3) {
#pragma clang fp contract(off)
  return fmaf(4.0f, a3, fmaf(-18.0f, a2, fmaf(36.0f, a1, -22.0f * a0)));
}
__device__ __forceinline__ float s1_backward(float a3, float a2, float a1, float a0) {
#pragma clang fp contract(off)
  return -fmaf(4.0f, a3, fmaf(-18.0f, a2, fmaf(36.0f, a1, -22.0f * a0)));
}

// First derivative (no spacing) at original index q in [-1, N] of an axis of
// N cells that the reference edge-pads by one: padded index q+1, one-sided
// rules on padded indices {0,1} and {N, N+1}. `f(r)` returns the padded value
// at original index r (callers clamp r into [0, N-1], which reproduces the
// edge padding for every index these rules read).
template <class F>
__device__ __forceinline__ float d1_general(F f, int q, int N) {
  if (q <= 0) return d1_forward(f(q), f(q + 1), f(q + 2), f(q + 3));
  if (q >= N - 1) return d1_backward(f(q - 3), f(q - 2), f(q - 1), f(q));
  return d1_interior(f(q - 2), f(q - 1), f(q + 1), f(q + 2));
}

// Second derivative G(G(pad(a))) at original index i in [0, N-1], from the
// first derivatives g(q), q in [-1, N] (padded indices 0..N+1).
template <class G>
__device__ __forceinline__ float d2_from_d1(G g, int i, int N) {
  if (i == 0) return d1_forward(g(0), g(1), g(2), g(3));
  if (i == N - 1) return d1_backward(g(N - 4), g(N - 3), g(N - 2), g(N - 1));
  return d1_interior(g(i - 2), g(i - 1), g(i + 1), g(i + 2));
}

// One cell of the step with every rule chosen per position (edge padding,
// one-sided boundary differences), neighbours read from global memory with
// clamped indices. Used where a cell is near the boundary and by the plain
// baseline kernel. EULER: write x + dt * dx/dt, else the derivatives.
template <bool EULER>
__device__ __forceinline__ void cell_general(const Coef& c, const float* __restrict__ D,
                                             const float* __restrict__ v,
                                             const float* __restrict__ w,
                                             const float* __restrict__ u, int H, int W, int i,
                                             int j, const StimArgs& sa, float dt,
                                             float inv_dx2, float* __restrict__ o_v,
                                             float* __restrict__ o_w,
                                             float* __restrict__ o_u) {
  const size_t idx = (size_t)i * W + j;
  auto ux = [&](int r) { return u[(size_t)clampi(r, 0, H - 1) * W + j]; };
  auto uy = [&](int s) { return u[(size_t)i * W + clampi(s, 0, W - 1)]; };
  auto Dx = [&](int r) { return D[(size_t)clampi(r, 0, H - 1) * W + j]; };
  auto Dy = [&](int s) { return D[(size_t)i * W + clampi(s, 0, W - 1)]; };
  auto gx = [&](int q) { return d1_general(ux, q, H); };
  auto gy = [&](int q) { return d1_general(uy, q, W); };
  const float uxx = d2_from_d1(gx, i, H);
  const float uyy = d2_from_d1(gy, j, W);
  const float dDx = d1_general(Dx, i, H);
  const float dDy = d1_general(Dy, j, W);
  const float Dc = D[idx];
  const float del_u = (Dc * (uxx + uyy) + dDx * gx(i) + dDy * gy(j)) * inv_dx2;
  const float vc = v[idx], wc = w[idx], uc = u[idx];
  float dv, dw, jion;
  reaction(c, vc, wc, uc, dv, dw, jion);
  jion = stimulate_cell(sa, idx, jion);
  const float du = del_u + jion;
  if (EULER) {
    o_v[idx] = vc + dt * dv;
    o_w[idx] = wc + dt * dw;
    o_u[idx] = uc + dt * du;
  } else {
    o_v[idx] = dv;
    o_w[idx] = dw;
    o_u[idx] = du;
  }
}

}  // namespace fk
