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
#include "fk_adjoint.h"
#include "fk_device.h"
#include "fk_host.h"

namespace fk {

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


__global__ void __launch_bounds__(kAdjThreads)
k_adj_pre(const float* __restrict__ u, const float* __restrict__ D,
          const float* __restrict__ Dx, const float* __restrict__ Dy,
          const float* __restrict__ lu, const float* __restrict__ inj, int H, int W,
          const AdjAxisTab* __restrict__ tab_g, AdjPre o) {
  __shared__ AdjAxisTab tab[2];
  adj_pre_tile(u, D, Dx, Dy, lu, inj, H, W, tab_g, o, blockIdx.x, blockIdx.y, tab);
}

// One reverse step (adj_main_tile) per block of 64 columns x 2 rows.
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
  const AdjStimList sv{sa};
  adj_main_tile<NEXT>(c, v, w, u, H, W, sv, dt, inv_dx2, lv, lw, pre, lv_o, lw_o, lu_o, gD,
                      g_tau, tab_g, nx, blockIdx.x, blockIdx.y, gridDim.x, tab, xch);
}

__global__ void k_adj_zero(float* a, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    a[i] = 0.0f;
}


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
