// Streaming forward step with two adjacent columns per lane.
//
// Lane l of a wave holds columns c0 = cs - 4 + 2l and c0 + 1: 128 columns per
// wave, of which lanes 2..61 (columns cs .. cs+119) are outputs and two lanes
// on each side carry the 4-column halo. Compared with one column per lane
// (fk_stream.h) a pair needs 4 lane shifts per difference instead of 8, the
// halo costs 4 of 64 lanes instead of 8, and interior strips move each row
// with one 8-byte load or store per lane. Rows are walked exactly as in
// fk_stream.h: x windows in registers, the reference's edge rules only in
// bands and strips that touch an edge.
//
// Interior strips use 8-byte accesses and need W even (8-byte aligned rows);
// the launcher runs the one-column kernel for odd W.
#pragma once

#include "fk_stream.h"

namespace fk {

constexpr int kStripOut2 = 120;  // output columns per wave

struct StreamGeom2 {
  int n_strips, n_bands, n_tiles, blocks;
  int cs_last, rs_last;  // origin of the last strip / band (shifted back to end at N)
};

template <int R>
inline StreamGeom2 make_stream_geom2(int H, int W) {
  StreamGeom2 g{};
  g.n_strips = (W + kStripOut2 - 1) / kStripOut2;
  g.n_bands = (H + R - 1) / R;
  g.cs_last = W >= kStripOut2 ? W - kStripOut2 : 0;
  g.rs_last = H >= R ? H - R : 0;
  g.n_tiles = g.n_strips * g.n_bands;
  g.blocks = (g.n_tiles + 3) / 4;
  return g;
}

// 12 x first difference across columns of the pair (a at column c0, b at
// c0 + 1), rules as d1_lanes: forward at columns <= 0, backward at >= W-1.
template <bool GENERAL>
__device__ __forceinline__ void d1_pairs(float a, float b, int c0, int W, float& ga,
                                         float& gb) {
#pragma clang fp contract(off)
  const float pa = from_left(a), pb = from_left(b);    // c0-2, c0-1
  const float na = from_right(a), nb = from_right(b);  // c0+2, c0+3
  ga = s1_interior(pa, pb, b, na);
  gb = s1_interior(pb, a, na, nb);
  if (!GENERAL) return;
  const int c1 = c0 + 1;
  const float nna = from_right(na), ppb = from_left(pb);  // c0+4, c0-3
  const float fa = s1_forward(a, b, na, nb), fb = s1_forward(b, na, nb, nna);
  const float ba = s1_backward(ppb, pa, pb, a), bb = s1_backward(pa, pb, a, b);
  ga = c0 <= 0 ? fa : (c0 >= W - 1 ? ba : ga);
  gb = c1 <= 0 ? fb : (c1 >= W - 1 ? bb : gb);
}

// Second difference from first differences held as pairs, rules as d2_lanes:
// one-sided at columns 0 and W-1 only.
template <bool GENERAL>
__device__ __forceinline__ void d2_pairs(float a, float b, int c0, int W, float& ha,
                                         float& hb) {
#pragma clang fp contract(off)
  const float pa = from_left(a), pb = from_left(b);
  const float na = from_right(a), nb = from_right(b);
  ha = s1_interior(pa, pb, b, na);
  hb = s1_interior(pb, a, na, nb);
  if (!GENERAL) return;
  const int c1 = c0 + 1;
  const float nna = from_right(na), ppb = from_left(pb);
  const float fa = s1_forward(a, b, na, nb), fb = s1_forward(b, na, nb, nna);
  const float ba = s1_backward(ppb, pa, pb, a), bb = s1_backward(pa, pb, a, b);
  ha = c0 == 0 ? fa : (c0 == W - 1 ? ba : ha);
  hb = c1 == 0 ? fb : (c1 == W - 1 ? bb : hb);
}

typedef float f32x2 __attribute__((ext_vector_type(2)));

// The pair of a row at columns (c0, c0+1): one 8-byte load in interior strips
// (both columns inside the grid, 8-byte aligned), else two loads with the
// columns clamped into the grid (edge padding).
template <bool EDGE>
__device__ __forceinline__ void ld_pair(const float* rowp, uint32_t off0, uint32_t off1,
                                        float& a, float& b) {
  if (!EDGE) {
    const f32x2 x = *(const f32x2 __attribute__((address_space(1)))*)((gcbytes)rowp + off0);
    a = x.x;
    b = x.y;
  } else {
    a = ld(rowp, off0);
    b = ld(rowp, off1);
  }
}

template <bool EDGE>
__device__ __forceinline__ void st_pair(float* rowp, uint32_t off0, uint32_t off1, bool oa,
                                        bool ob, float a, float b) {
  if (!EDGE) {  // interior strip: both columns are outputs together
    if (oa) {
      f32x2 x;
      x.x = a;
      x.y = b;
      *(f32x2 __attribute__((address_space(1)))*)((gbytes)rowp + off0) = x;
    }
  } else {
    if (oa) st(rowp, off0, a);
    if (ob) st(rowp, off1, b);
  }
}

template <int R, bool EULER, bool XG, bool YG, bool STIM>
__device__ __forceinline__ void stream_tile2(const Coef& c, const float* __restrict__ D,
                                             const float* __restrict__ v,
                                             const float* __restrict__ w,
                                             const float* __restrict__ u, int H, int W,
                                             const StimArgs& sa, float dt, float inv_dx2,
                                             float* __restrict__ o_v, float* __restrict__ o_w,
                                             float* __restrict__ o_u, int rs, int cs) {
#pragma clang fp contract(off)  // explicit fmaf only: identical rounding in every kernel
  const int lane = threadIdx.x & 63;
  const int c0 = cs - 4 + 2 * lane, c1 = c0 + 1;
  const uint32_t off0 = (uint32_t)clampi(c0, 0, W - 1) * 4u;
  const uint32_t off1 = (uint32_t)clampi(c1, 0, W - 1) * 4u;
  const bool out_lane = lane >= 2 && lane < 2 + kStripOut2 / 2;
  const bool oa = out_lane && c0 < W, ob = out_lane && c1 < W;
  auto row = [&](const float* a, int r) {  // wave-uniform row base, clamped
    return a + (size_t)clampi(r, 0, H - 1) * (size_t)W;
  };

  float ua[R + 8], ub[R + 8];  // u rows rs-4 .. rs+R+3
  float da[R + 4], db[R + 4];  // D rows rs-2 .. rs+R+1
  float va[R], vb[R], wa[R], wb[R];
#pragma unroll
  for (int k = 0; k < R + 8; ++k) ld_pair<YG>(row(u, rs - 4 + k), off0, off1, ua[k], ub[k]);
#pragma unroll
  for (int k = 0; k < R + 4; ++k) ld_pair<YG>(row(D, rs - 2 + k), off0, off1, da[k], db[k]);
#pragma unroll
  for (int k = 0; k < R; ++k) {
    ld_pair<YG>(row(v, rs + k), off0, off1, va[k], vb[k]);
    ld_pair<YG>(row(w, rs + k), off0, off1, wa[k], wb[k]);
  }

  const float k2 = inv_dx2 * (1.0f / 144.0f);
  // 12 u_x at rows q = rs-2 .. rs+R+1, both columns
  float gxa[R + 4], gxb[R + 4];
#pragma unroll
  for (int k = 0; k < R + 4; ++k) {
    const int q = rs - 2 + k;
    if (XG && k <= 2 && q <= 0) {
      gxa[k] = s1_forward(ua[k + 2], ua[k + 3], ua[k + 4], ua[k + 5]);
      gxb[k] = s1_forward(ub[k + 2], ub[k + 3], ub[k + 4], ub[k + 5]);
    } else if (XG && k >= 3 && q >= H - 1) {
      gxa[k] = s1_backward(ua[k - 1], ua[k], ua[k + 1], ua[k + 2]);
      gxb[k] = s1_backward(ub[k - 1], ub[k], ub[k + 1], ub[k + 2]);
    } else {
      gxa[k] = s1_interior(ua[k], ua[k + 1], ua[k + 3], ua[k + 4]);
      gxb[k] = s1_interior(ub[k], ub[k + 1], ub[k + 3], ub[k + 4]);
    }
  }

#pragma unroll
  for (int k = 0; k < R; ++k) {
    const int i = rs + k;
    float uxx[2], dDx[2];
    const float* gx2[2] = {gxa, gxb};
    const float* dd2[2] = {da, db};
#pragma unroll
    for (int s = 0; s < 2; ++s) {
      const float* g = gx2[s];
      const float* d = dd2[s];
      if (XG && k == 0 && i == 0) {
        uxx[s] = s1_forward(g[k + 2], g[k + 3], g[k + 4], g[k + 5]);
        dDx[s] = s1_forward(d[k + 2], d[k + 3], d[k + 4], d[k + 5]);
      } else if (XG && k >= 1 && i == H - 1) {
        uxx[s] = s1_backward(g[k - 1], g[k], g[k + 1], g[k + 2]);
        dDx[s] = s1_backward(d[k - 1], d[k], d[k + 1], d[k + 2]);
      } else {
        uxx[s] = s1_interior(g[k], g[k + 1], g[k + 3], g[k + 4]);
        dDx[s] = s1_interior(d[k], d[k + 1], d[k + 3], d[k + 4]);
      }
    }
    float gya, gyb, uyya, uyyb, dDya, dDyb;
    d1_pairs<YG>(ua[k + 4], ub[k + 4], c0, W, gya, gyb);
    d2_pairs<YG>(gya, gyb, c0, W, uyya, uyyb);
    d1_pairs<YG>(da[k + 2], db[k + 2], c0, W, dDya, dDyb);
    const float dela =
        fmaf(dDya, gya, fmaf(dDx[0], gxa[k + 2], da[k + 2] * (uxx[0] + uyya))) * k2;
    const float delb =
        fmaf(dDyb, gyb, fmaf(dDx[1], gxb[k + 2], db[k + 2] * (uxx[1] + uyyb))) * k2;
    float dva, dwa, ja, dvb, dwb, jb;
    reaction(c, va[k], wa[k], ua[k + 4], dva, dwa, ja);
    reaction(c, vb[k], wb[k], ub[k + 4], dvb, dwb, jb);
    if (STIM) {
      float sa_ = 0.0f, sb_ = 0.0f;
      for (int m = 0; m < sa.n; ++m) {
        float fa, fb;
        ld_pair<YG>(row(sa.f[m], i), off0, off1, fa, fb);
        sa_ = (fa != 0.0f) ? fa : sa_;
        sb_ = (fb != 0.0f) ? fb : sb_;
      }
      ja = (sa_ != 0.0f) ? sa_ : ja;
      jb = (sb_ != 0.0f) ? sb_ : jb;
    }
    const float dua = dela + ja, dub = delb + jb;
    if (!XG || i < H) {  // rows past the grid only when H < R
      const size_t ro = (size_t)i * (size_t)W;
      if (EULER) {
        st_pair<YG>(o_v + ro, off0, off1, oa, ob, fmaf(dt, dva, va[k]), fmaf(dt, dvb, vb[k]));
        st_pair<YG>(o_w + ro, off0, off1, oa, ob, fmaf(dt, dwa, wa[k]), fmaf(dt, dwb, wb[k]));
        st_pair<YG>(o_u + ro, off0, off1, oa, ob, fmaf(dt, dua, ua[k + 4]),
                    fmaf(dt, dub, ub[k + 4]));
      } else {
        st_pair<YG>(o_v + ro, off0, off1, oa, ob, dva, dvb);
        st_pair<YG>(o_w + ro, off0, off1, oa, ob, dwa, dwb);
        st_pair<YG>(o_u + ro, off0, off1, oa, ob, dua, dub);
      }
    }
  }
}

// One wave per (128-column strip, R-row band) tile, four waves per block.
template <int R, bool EULER>
__global__ void __launch_bounds__(256)
k_stream2(Coef c, const float* __restrict__ D, const float* __restrict__ v,
          const float* __restrict__ w, const float* __restrict__ u, int H, int W, StimArgs sa,
          float dt, float inv_dx2, float* __restrict__ o_v, float* __restrict__ o_w,
          float* __restrict__ o_u, StreamGeom2 g) {
  const int tile = __builtin_amdgcn_readfirstlane(blockIdx.x * 4 + (int)(threadIdx.x >> 6));
  if (tile >= g.n_tiles) return;  // whole wave leaves; the kernel has no barrier
  const int strip = tile / g.n_bands, band = tile - strip * g.n_bands;
  const int cs = min(kStripOut2 * strip, g.cs_last);
  const int rs = min(R * band, g.rs_last);
  const bool xg = rs < 3 || rs + R + 3 > H;          // band within 3 rows of an edge
  const bool yg = cs < 4 || cs + kStripOut2 + 4 > W;  // the 128 columns leave the grid
  const int sel = (xg ? 1 : 0) | (yg ? 2 : 0) | (sa.n ? 4 : 0);
#define FK_TILE2(X, Y, S)                                                                    \
  stream_tile2<R, EULER, X, Y, S>(c, D, v, w, u, H, W, sa, dt, inv_dx2, o_v, o_w, o_u, rs, cs)
  switch (sel) {
    case 0: FK_TILE2(false, false, false); break;
    case 1: FK_TILE2(true, false, false); break;
    case 2: FK_TILE2(false, true, false); break;
    case 3: FK_TILE2(true, true, false); break;
    case 4: FK_TILE2(false, false, true); break;
    case 5: FK_TILE2(true, false, true); break;
    case 6: FK_TILE2(false, true, true); break;
    default: FK_TILE2(true, true, true); break;
  }
#undef FK_TILE2
}

}  // namespace fk
