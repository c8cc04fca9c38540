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

// One reverse step. Block partial sums of the tau gradients are added to
// g_tau_part[block][0..2] (each block owns its slot: deterministic).
__global__ void __launch_bounds__(kAdjThreads)
k_adjoint_step(Coef c, const float* __restrict__ D, const float* __restrict__ Dx,
               const float* __restrict__ Dy, const float* __restrict__ v,
               const float* __restrict__ w, const float* __restrict__ u, int H, int W,
               StimArgs sa, float dt, float inv_dx2, const float* __restrict__ lv,
               const float* __restrict__ lw, const float* __restrict__ lu,
               float* __restrict__ lv_o, float* __restrict__ lw_o, float* __restrict__ lu_o,
               float* __restrict__ gD, double* __restrict__ g_tau_part,
               const AdjAxisTab* __restrict__ tab) {
  const int c_ = blockIdx.x * 64 + (threadIdx.x & 63);
  const int i = blockIdx.y * 4 + (threadIdx.x >> 6);
  float tdl = 0.0f, trl = 0.0f, tsl = 0.0f;
  if (i < H && c_ < W) {
    const int col = c_;
    const size_t idx = (size_t)i * W + col;
    const AxisOps ox{&tab[0], H}, oy{&tab[1], W};
    auto at = [&](const float* a, int r, int q) { return a[(size_t)r * W + q]; };
    const float vv = v[idx], ww = w[idx], uu = u[idx];
    const float lvv = lv[idx], lww = lw[idx], luu = lu[idx];

    // --- reaction, gates as fixed masks ---------------------------------
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
    bool stim = false;
    for (int k = 0; k < sa.n; ++k) stim = stim || (sa.f[k][idx] != 0.0f);
    const float lus = stim ? 0.0f : luu;

    // --- diffusion transposes --------------------------------------------
    float sx = 0.0f, sy = 0.0f;  // A2^T(D lam) + A1^T(D_? lam)
    float ex = 0.0f, ey = 0.0f;  // A1^T(lam A1 u)
    float lap = 0.0f;            // A2x u + A2y u at the cell
    for (int d = -kAdjBand; d <= kAdjBand; ++d) {
      const int jr = i - d;  // row whose entry [jr][i] is needed: A[jr][jr + d]
      if (jr >= 0 && jr < H) {
        const float a2 = ox.a2(jr, d), a1 = ox.a1(jr, d);
        if (a2 != 0.0f || a1 != 0.0f) {
          const float lam = at(lu, jr, col);
          sx += a2 * at(D, jr, col) * lam + a1 * at(Dx, jr, col) * lam;
          if (a1 != 0.0f) {
            float uxj = 0.0f;
            for (int e = -kAdjBand; e <= kAdjBand; ++e)
              if (jr + e >= 0 && jr + e < H) uxj += ox.a1(jr, e) * at(u, jr + e, col);
            ex += a1 * lam * uxj;
          }
        }
      }
      const int jc = col - d;
      if (jc >= 0 && jc < W) {
        const float a2 = oy.a2(jc, d), a1 = oy.a1(jc, d);
        if (a2 != 0.0f || a1 != 0.0f) {
          const float lam = at(lu, i, jc);
          sy += a2 * at(D, i, jc) * lam + a1 * at(Dy, i, jc) * lam;
          if (a1 != 0.0f) {
            float uyj = 0.0f;
            for (int e = -kAdjBand; e <= kAdjBand; ++e)
              if (jc + e >= 0 && jc + e < W) uyj += oy.a1(jc, e) * at(u, i, jc + e);
            ey += a1 * lam * uyj;
          }
        }
      }
      if (i + d >= 0 && i + d < H) lap += ox.a2(i, d) * at(u, i + d, col);
      if (col + d >= 0 && col + d < W) lap += oy.a2(col, d) * at(u, i, col + d);
    }

    lv_o[idx] = lvv + dt * (dfv_dv * lvv + djdv * lus);
    lw_o[idx] = lww + dt * (dfw_dw * lww + djdw * lus);
    lu_o[idx] = luu + dt * (djdu * lus + (sx + sy) * inv_dx2);
    gD[idx] += dt * (luu * lap + ex + ey) * inv_dx2;
    // tau gradients: d j_ion / d tau
    const float djd_td = p ? -vv * (uu - c.V_c) * (1.0f - uu) * c.inv_td * c.inv_td * c.inv_cm
                           : 0.0f;
    const float djd_tr = p ? c.inv_tr * c.inv_tr * c.inv_cm : 0.0f;
    const float djd_tsi = -ww * S * c.inv_tsi * c.inv_tsi * c.inv_cm;
    tdl = dt * djd_td * lus;
    trl = dt * djd_tr * lus;
    tsl = dt * djd_tsi * lus;
  }
  // block reduction (wave shuffles, then LDS across the 4 waves)
  __shared__ float red[3][kAdjThreads / 64];
  for (int o = 32; o > 0; o >>= 1) {
    tdl += __shfl_xor(tdl, o);
    trl += __shfl_xor(trl, o);
    tsl += __shfl_xor(tsl, o);
  }
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0) {
    red[0][wv] = tdl;
    red[1][wv] = trl;
    red[2][wv] = tsl;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const int b = blockIdx.y * gridDim.x + blockIdx.x;
    for (int k = 0; k < 3; ++k) {
      double s = 0.0;
      for (int q = 0; q < kAdjThreads / 64; ++q) s += (double)red[k][q];
      g_tau_part[3 * (size_t)b + k] += s;
    }
  }
}

__global__ void k_adj_zero(double* a, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    a[i] = 0.0;
}

__global__ void k_adj_reduce(const double* __restrict__ part, size_t n_blocks, float* g_tau) {
  __shared__ double s[3][256];
  for (int k = 0; k < 3; ++k) {
    double acc = 0.0;
    for (size_t b = threadIdx.x; b < n_blocks; b += blockDim.x) acc += part[3 * b + k];
    s[k][threadIdx.x] = acc;
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

// Workspace layout: [tables (2 axes)][Dx][Dy][tau partials]
struct AdjLayout {
  size_t tab, dx, dy, part, total, n_blocks;
};

static AdjLayout adj_layout(int H, int W) {
  AdjLayout L;
  const size_t cells = (size_t)H * W;
  const dim3 grid((W + 63) / 64, (H + 3) / 4);
  L.n_blocks = (size_t)grid.x * grid.y;
  size_t o = 0;
  auto take = [&](size_t bytes) {
    const size_t at = o;
    o += (bytes + 255) / 256 * 256;
    return at;
  };
  L.tab = take(2 * sizeof(AdjAxisTab));
  L.dx = take(cells * sizeof(float));
  L.dy = take(cells * sizeof(float));
  L.part = take(3 * L.n_blocks * sizeof(double));
  L.total = o;
  return L;
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
  k_adj_zero<<<64, 256, 0, s>>>((double*)(wk + L.part), 3 * L.n_blocks);
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
  char* wk = (char*)work;
  StimSchedule sched(stimuli, n_stimuli);
  const Coef c = make_coef(*params);
  const float inv_dx2 = (float)(1.0 / ((double)dx * (double)dx));
  const dim3 grid((W + 63) / 64, (H + 3) / 4);
  k_adjoint_step<<<grid, kAdjThreads, 0, (hipStream_t)stream>>>(
      c, D, (const float*)(wk + L.dx), (const float*)(wk + L.dy), v, w, u, H, W, sched.at(t),
      dt, inv_dx2, lam_v, lam_w, lam_u, lam_v_out, lam_w_out, lam_u_out, grad_D,
      (double*)(wk + L.part), (const AdjAxisTab*)(wk + L.tab));
  HIP_CHECK(hipGetLastError());
  return FK_OK;
}

int32_t fk_adjoint_tau_grads(const void* work, int32_t H, int32_t W, float* g_tau,
                             void* stream) {
  ErrorScope es;
  if (!check_grid(H, W) || !check_ptrs({work, g_tau})) return last_code();
  const AdjLayout L = adj_layout(H, W);
  k_adj_reduce<<<1, 256, 0, (hipStream_t)stream>>>(
      (const double*)((const char*)work + L.part), L.n_blocks, g_tau);
  HIP_CHECK(hipGetLastError());
  return FK_OK;
}

}  // extern "C"
