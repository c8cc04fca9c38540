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
  float kk;  // -2 k log2(e): exp(-2k(u - V_csi)) = exp2((u - V_csi) * kk)
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
  const float e = __builtin_amdgcn_exp2f((u - c.V_csi) * c.kk);
  return __builtin_amdgcn_rcpf(1.0f + e);
}

// Gates and ionic currents, cardiax/solve.py:35-42,57-58.
// p = [u >= V_c], q = [u >= V_v]; outputs dv, dw and j_ion (before stimulus).
__device__ __forceinline__ void reaction(const Coef& c, float v, float w, float u, float& dv,
                                         float& dw, float& jion) {
  const float jsi = -(w * half_one_plus_tanh(c, u)) * c.inv_tsi;
  float jfi, jso;
  if (u >= c.V_c) {
    jfi = -v * (u - c.V_c) * (1.0f - u) * c.inv_td;
    jso = c.inv_tr;
    dv = -v * c.inv_tvp;
    dw = -w * c.inv_twp;
  } else {
    const float inv_tvm = (u >= c.V_v) ? c.inv_tv2m : c.inv_tv1m;
    jfi = 0.0f;
    jso = u * c.inv_t0;
    dv = (1.0f - v) * inv_tvm;
    dw = (1.0f - w) * c.inv_twm;
  }
  jion = -(jfi + jso + jsi) * c.inv_cm;
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

}  // namespace fk
