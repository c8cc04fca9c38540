// Temporally blocked forward-Euler kernel: K steps per launch on an LDS tile.
//
// Each workgroup owns a T x T output tile. It stages u over the tile plus a
// halo of 4K cells on every side (the composite stencil G(G(.)) reaches 4
// cells per step), keeps v, w, D and the static first derivatives D_x, D_y of
// its cells in registers, and advances K Euler steps on chip; only the tile's
// final v, w, u go back to HBM. Per sub-step j the cells whose dependency cone
// is still inside the staged region are updated: a cell m rows/columns away
// from a side where the staged region was cut (not where the grid ends)
// stays exact for floor(m / 4) steps. Where the staged region reaches the
// grid boundary nothing is cut and the reference's boundary rules apply
// (cardiax/solve.py:225-254: edge padding, one-sided differences on the two
// outermost padded entries), evaluated from the staged values with clamped
// indices.
//
// Thread <-> cell map: cell e = tid + NT * m of the E x E staged region
// (E = T + 8K), row-major, so lanes run along columns and every LDS access of
// a wavefront is to consecutive words.
#pragma once

#include "fk_device.h"

namespace fk {

// Stimuli for one launch: all fields, and per sub-step the bit set of the
// active ones (bit m = stimulus m, applied in list order).
struct TbStim {
  uint32_t act[8];
  int32_t n;
  const float* f[FK_MAX_STIMULI];
};

template <int T, int K, int NT>
struct TbShape {
  static constexpr int E = T + 8 * K;
  static constexpr int NC = (E * E + NT - 1) / NT;
};

// Workgroup-uniform tile origin with XCD-aware ordering: consecutive
// workgroups are dispatched round-robin over the 8 XCDs, so workgroup b is
// given tile (b % 8) * (n / 8) + b / 8 when 8 divides n; each XCD then works
// on a contiguous band of tiles and finds its neighbours' halo in its own L2.
__device__ __forceinline__ int tb_tile_index(int b, int n) {
  return (n % 8 == 0) ? (b % 8) * (n / 8) + b / 8 : b;
}

// Body of one tile. BND: the staged region reaches the grid edge somewhere,
// so the per-position boundary rules are needed; interior tiles (BND false)
// contain central differences only.
//
// With BND the stage also holds the ghost cells of the reference's edge
// padding (row -1 / H, column -1 / W): they carry no u of their own (reads of
// u are clamped into the grid, which is the padding) but do hold the ghost
// first derivatives that the second difference of cells 1 and N-2 reads.
template <int T, int K, int NT, bool BND>
__device__ __forceinline__ void tb_body(const Coef& c, const float* __restrict__ D,
                                        const float* __restrict__ v_in,
                                        const float* __restrict__ w_in,
                                        const float* __restrict__ u_in, int H, int W,
                                        const TbStim& st, int nsteps, float dt, float inv_dx2,
                                        float* __restrict__ v_out, float* __restrict__ w_out,
                                        float* __restrict__ u_out, int r0, int c0, float* s_u,
                                        float* s_gx, float* s_gy) {
  using S = TbShape<T, K, NT>;
  constexpr int E = S::E, NC = S::NC, HALO = 4 * K;
  // meta: bits 0-15 margin to the nearest cut side, then flags
  constexpr uint32_t kMarg = 0xffffu, kDom = 1u << 16, kGhR = 1u << 17, kGhC = 1u << 18;
  // A side is open when the stage extends past the grid there (it then holds
  // the ghost row / column); open sides do not shrink the exact region.
  const bool top_open = r0 < 0, bot_open = r0 + E > H;
  const bool lft_open = c0 < 0, rgt_open = c0 + E > W;
  const int tid = threadIdx.x;

  uint32_t meta[NC];
  float rv[NC], rw[NC], ru[NC], rD[NC], rDx[NC], rDy[NC];
#pragma unroll
  for (int m = 0; m < NC; ++m) {
    const int e = tid + NT * m;
    const int a = e / E, b = e % E;
    const int gr = r0 + a, gc = c0 + b;
    const bool in_r = gr >= 0 && gr < H, in_c = gc >= 0 && gc < W;
    const bool dom = e < E * E && in_r && in_c;
    const bool ghr = BND && e < E * E && in_c && (gr == -1 || gr == H);
    const bool ghc = BND && e < E * E && in_r && (gc == -1 || gc == W);
    int mg = 0x7fff;
    if (!top_open) mg = min(mg, a);
    if (!bot_open) mg = min(mg, E - 1 - a);
    if (!lft_open) mg = min(mg, b);
    if (!rgt_open) mg = min(mg, E - 1 - b);
    // cells past the stage (e >= E*E, last iteration) are dead: meta 0
    meta[m] = e < E * E ? ((uint32_t)max(mg, 0) | (dom ? kDom : 0u) | (ghr ? kGhR : 0u) |
                           (ghc ? kGhC : 0u))
                        : 0u;
    float uu = 0.0f, vv = 0.0f, ww = 0.0f, Dd = 0.0f, Dx = 0.0f, Dy = 0.0f;
    if (dom) {
      const size_t idx = (size_t)gr * W + gc;
      uu = u_in[idx];
      vv = v_in[idx];
      ww = w_in[idx];
      Dd = D[idx];
      // D is static: its first derivatives once, from global memory.
      auto fDx = [&](int r) { return D[(size_t)clampi(r, 0, H - 1) * W + gc]; };
      auto fDy = [&](int q) { return D[(size_t)gr * W + clampi(q, 0, W - 1)]; };
      Dx = d1_general(fDx, gr, H);
      Dy = d1_general(fDy, gc, W);
    }
    ru[m] = uu;
    rv[m] = vv;
    rw[m] = ww;
    rD[m] = Dd;
    rDx[m] = Dx;
    rDy[m] = Dy;
    if (e < E * E) s_u[e] = uu;
  }
  __syncthreads();

#pragma unroll
  for (int j = 0; j < K; ++j) {
    if (j >= nsteps) break;
    const int need_g = 4 * j + 2, need_u = 4 * j + 4;
    // First derivatives of u(j) where the stage is still exact within +-2.
#pragma unroll
    for (int m = 0; m < NC; ++m) {
      const uint32_t mt = meta[m];
      if ((int)(mt & kMarg) < need_g) continue;
      const int e = tid + NT * m;
      if (!BND) {
        if (!(mt & kDom)) continue;
        s_gx[e] = d1_interior(s_u[e - 2 * E], s_u[e - E], s_u[e + E], s_u[e + 2 * E]);
        s_gy[e] = d1_interior(s_u[e - 2], s_u[e - 1], s_u[e + 1], s_u[e + 2]);
      } else {
        if (!(mt & (kDom | kGhR | kGhC))) continue;
        const int a = e / E, b = e % E, gr = r0 + a, gc = c0 + b;
        const int grc = clampi(gr, 0, H - 1), gcc = clampi(gc, 0, W - 1);
        // u at (row gr + d, this column) and (this row, column gc + d), clamped
        auto ux = [&](int d) { return s_u[e + (clampi(gr + d, 0, H - 1) - gr) * E + (gcc - gc)]; };
        auto uy = [&](int d) { return s_u[e + (grc - gr) * E + (clampi(gc + d, 0, W - 1) - gc)]; };
        if (mt & (kDom | kGhR)) s_gx[e] = d1_general([&](int q) { return ux(q - gr); }, gr, H);
        if (mt & (kDom | kGhC)) s_gy[e] = d1_general([&](int q) { return uy(q - gc); }, gc, W);
      }
    }
    __syncthreads();

    // Update the cells whose +-4 cone is exact.
    const uint32_t act = st.act[j];
#pragma unroll
    for (int m = 0; m < NC; ++m) {
      const uint32_t mt = meta[m];
      if (!(mt & kDom) || (int)(mt & kMarg) < need_u) continue;
      const int e = tid + NT * m;
      float uxx, uyy;
      if (!BND) {
        uxx = d1_interior(s_gx[e - 2 * E], s_gx[e - E], s_gx[e + E], s_gx[e + 2 * E]);
        uyy = d1_interior(s_gy[e - 2], s_gy[e - 1], s_gy[e + 1], s_gy[e + 2]);
      } else {
        // one-sided at cells 0 and N-1; elsewhere central, reading the
        // staged ghost first derivatives next to cells 1 and N-2
        const int a = e / E, b = e % E, gr = r0 + a, gc = c0 + b;
        auto gxa = [&](int q) { return s_gx[e + (q - gr) * E]; };
        auto gya = [&](int q) { return s_gy[e + (q - gc)]; };
        uxx = d2_from_d1(gxa, gr, H);
        uyy = d2_from_d1(gya, gc, W);
      }
      const float del_u =
          (rD[m] * (uxx + uyy) + rDx[m] * s_gx[e] + rDy[m] * s_gy[e]) * inv_dx2;
      float dv, dw, jion;
      reaction(c, rv[m], rw[m], ru[m], dv, dw, jion);
      if (act) {
        const int a = e / E, b = e % E;
        const size_t idx = (size_t)(r0 + a) * W + (c0 + b);
        float sv = 0.0f;
        for (int k = 0; k < st.n; ++k) {
          if (!((act >> k) & 1u)) continue;
          const float f = st.f[k][idx];
          sv = (f != 0.0f) ? f : sv;
        }
        jion = (sv != 0.0f) ? sv : jion;
      }
      rv[m] = fmaf(dt, dv, rv[m]);
      rw[m] = fmaf(dt, dw, rw[m]);
      ru[m] = fmaf(dt, del_u + jion, ru[m]);
    }
    if (j + 1 < nsteps) {
      __syncthreads();  // every read of s_u, s_gx, s_gy for this sub-step is done
#pragma unroll
      for (int m = 0; m < NC; ++m) {
        const uint32_t mt = meta[m];
        if ((mt & kDom) && (int)(mt & kMarg) >= need_u) s_u[tid + NT * m] = ru[m];
      }
      __syncthreads();
    }
  }

  // The tile's own cells (inside the T x T output tile, margin >= 4K).
#pragma unroll
  for (int m = 0; m < NC; ++m) {
    const int e = tid + NT * m;
    const int a = e / E, b = e % E;
    if (!(meta[m] & kDom) || a < HALO || a >= HALO + T || b < HALO || b >= HALO + T) continue;
    const size_t idx = (size_t)(r0 + a) * W + (c0 + b);
    v_out[idx] = rv[m];
    w_out[idx] = rw[m];
    u_out[idx] = ru[m];
  }
}

template <int T, int K, int NT>
__global__ void __launch_bounds__(NT)
k_tb(Coef c, const float* __restrict__ D, const float* __restrict__ v_in,
     const float* __restrict__ w_in, const float* __restrict__ u_in, int H, int W, TbStim st,
     int nsteps, float dt, float inv_dx2, float* __restrict__ v_out, float* __restrict__ w_out,
     float* __restrict__ u_out, int tiles_x, int n_tiles) {
  using S = TbShape<T, K, NT>;
  constexpr int E = S::E, HALO = 4 * K;
  __shared__ float s_u[E * E];
  __shared__ float s_gx[E * E];
  __shared__ float s_gy[E * E];
  const int tile = tb_tile_index(blockIdx.x, n_tiles);
  const int ty = tile / tiles_x, tx = tile - ty * tiles_x;
  // The last tile of each axis is shifted back to end on the grid edge (its
  // overlap with the previous tile is recomputed to identical values). The
  // one-sided rules at cells 0 / N-1 reach 5 cells inward per step instead of
  // 4; with full tiles, cells within 3 of an edge sit >= T + 4K - 3 cells from
  // the opposite cut, which covers that reach for every K <= T - 3.
  const int oy = (H >= T) ? min(ty * T, H - T) : 0;
  const int ox = (W >= T) ? min(tx * T, W - T) : 0;
  const int r0 = oy - HALO, c0 = ox - HALO;  // global origin of the staged region
  // Boundary rules are needed only when the stage extends past a grid edge;
  // a stage that is cut everywhere updates only cells >= 2 rows / columns
  // inside its cut, where every difference is central.
  const bool bnd = r0 < 0 || c0 < 0 || r0 + E > H || c0 + E > W;
  if (bnd)
    tb_body<T, K, NT, true>(c, D, v_in, w_in, u_in, H, W, st, nsteps, dt, inv_dx2, v_out,
                            w_out, u_out, r0, c0, s_u, s_gx, s_gy);
  else
    tb_body<T, K, NT, false>(c, D, v_in, w_in, u_in, H, W, st, nsteps, dt, inv_dx2, v_out,
                             w_out, u_out, r0, c0, s_u, s_gx, s_gy);
}

}  // namespace fk
