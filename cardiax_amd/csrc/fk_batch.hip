// Batched forward Euler: many independent simulations of uneven sizes and
// parameter sets advanced together.
//
// One launch over a work list that holds every (column strip, row band) tile
// of every simulation advances the whole batch by one Euler step (the
// streaming step, stream_tile) or by two (the two-step streaming tile,
// stream_t_tile, which reads and writes the state once for both); each tile
// takes its simulation's constants, D-map and stimuli from a device-resident
// descriptor. Stimulus activity follows the law of cardiax/solve.py:262-267
// (floor modulo in double precision, as on the host path), evaluated on the
// device by one small kernel for up to kGraphSteps steps at a time. Steps are
// recorded once into a hipGraph (that kernel, the step launches and a
// one-thread kernel that advances the device-side step counter) and the graph
// is replayed: the host issues one graph launch per kGraphSteps steps for the
// whole batch.
//
// Reference: the batch has no counterpart in the reference, which runs one
// simulation per call (experiments/generate_fd_data_128.py:25-56 loops over
// seeds); each simulation follows cardiax/solve.py:92-100.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <vector>

#include "cardiax_amd.h"
#include "fk_batch.h"
#include "fk_device.h"
#include "fk_host.h"
#include "fk_stream.h"
#include "fk_stream_t.h"

namespace fk {

// Work-list options, chosen by fk_batch_autotune: one step per launch on bands
// of 2, 4, 8 or 16 rows, or two steps per launch on bands of 8 or 16 rows
// (those only when every simulation fits stream_t_fits).
constexpr int kNumRowOpts = FK_BATCH_N_OPTIONS;
constexpr int kBatchRows[kNumRowOpts] = {2, 4, 8, 16, 8, 16};
constexpr bool kBatchTwo[kNumRowOpts] = {false, false, false, false, true, true};
constexpr int kDefaultRowOpt = 2;  // 8 rows, one step
constexpr int kGraphSteps = 32;   // Euler steps recorded in one graph (even)

// The one-step option with the band height of option o (the odd last step of
// a two-step option).
constexpr int one_step_opt(int o) { return o == 4 ? 2 : o == 5 ? 3 : o; }

struct SimDesc {
  Coef c;
  const float* D;
  float* buf[2][3];  // ping-pong state: [parity][v, w, u]
  int H, W;
  int n_stim;
  float dt, inv_dx2;
  double t0;  // time of the first step
  BatchStim stim[FK_BATCH_MAX_STIMULI];
};

struct TileItem {
  int sim, cs, rs, flags;  // flags: bit0 band near a row edge, bit1 strip near a column edge
};

// Stimulus activity of every simulation at the n_rel steps from the device
// step counter on: bit k of mask[r * n + i] = stimulus k of simulation i is
// active at step r of them.
__global__ void k_batch_masks(const SimDesc* __restrict__ sims, int n,
                              const int64_t* __restrict__ step_base, int n_rel,
                              uint32_t* __restrict__ mask) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n * n_rel) return;
  const int i = idx % n, r = idx / n;
  const SimDesc& S = sims[i];
  const double t = S.t0 + (double)(*step_base + r);
  uint32_t m = 0;
  for (int k = 0; k < S.n_stim; ++k)
    if (active_dev(t, S.stim[k])) m |= 1u << k;
  mask[idx] = m;
}

// One Euler step of every tile (bands of R rows). Reads state[parity],
// writes state[parity ^ 1].
template <int R>
__global__ void __launch_bounds__(256)
k_batch_step(const SimDesc* __restrict__ sims, const TileItem* __restrict__ items, int n_items,
             const uint32_t* __restrict__ masks, int parity) {
  const int it = __builtin_amdgcn_readfirstlane(blockIdx.x * 4 + (int)(threadIdx.x >> 6));
  if (it >= n_items) return;
  // wave-uniform work item (scalar registers)
  TileItem item = items[it];
  item.sim = __builtin_amdgcn_readfirstlane(item.sim);
  item.cs = __builtin_amdgcn_readfirstlane(item.cs);
  item.rs = __builtin_amdgcn_readfirstlane(item.rs);
  item.flags = __builtin_amdgcn_readfirstlane(item.flags);
  const SimDesc& S = sims[item.sim];
  const uint32_t mask = __builtin_amdgcn_readfirstlane(masks[item.sim]);
  const BatchStimView sv{S.stim, uni(S.n_stim), mask};
  Coef c = S.c;
  c.V_c = uni(c.V_c); c.V_v = uni(c.V_v); c.V_csi = uni(c.V_csi); c.kk = uni(c.kk);
  c.kk_off = uni(c.kk_off);
  c.inv_tvp = uni(c.inv_tvp); c.inv_tv1m = uni(c.inv_tv1m); c.inv_tv2m = uni(c.inv_tv2m);
  c.inv_twp = uni(c.inv_twp); c.inv_twm = uni(c.inv_twm); c.inv_td = uni(c.inv_td);
  c.inv_t0 = uni(c.inv_t0); c.inv_tr = uni(c.inv_tr); c.inv_tsi = uni(c.inv_tsi);
  c.inv_cm = uni(c.inv_cm);
  const float* Dp = uni(S.D);
  const float* i0 = uni(S.buf[parity][0]);
  const float* i1 = uni(S.buf[parity][1]);
  const float* i2 = uni(S.buf[parity][2]);
  float* o0 = uni(S.buf[parity ^ 1][0]);
  float* o1 = uni(S.buf[parity ^ 1][1]);
  float* o2 = uni(S.buf[parity ^ 1][2]);
  const int H = uni(S.H), W = uni(S.W);
  const float dt = uni(S.dt), inv_dx2 = uni(S.inv_dx2);
  const int sel = item.flags | (mask ? 4 : 0);
#define FK_BT(X, Y, ST)                                                                    \
  stream_tile<R, true, X, Y, ST, BatchStimView, false, 0, true>(c, Dp, nullptr, i0, i1, i2, H, W,  \
                                                                sv, dt, inv_dx2, o0, o1, o2,      \
                                                                item.rs, item.cs)
  switch (sel) {
    case 0: FK_BT(false, false, false); break;
    case 1: FK_BT(true, false, false); break;
    case 2: FK_BT(false, true, false); break;
    case 3: FK_BT(true, true, false); break;
    case 4: FK_BT(false, false, true); break;
    case 5: FK_BT(true, false, true); break;
    case 6: FK_BT(false, true, true); break;
    default: FK_BT(true, true, true); break;
  }
#undef FK_BT
}

// Two Euler steps of every tile (two-step bands of R rows). Reads
// state[parity], writes state[parity ^ 1]; masks1 / masks2 hold the activity
// of the two steps.
template <int R>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2)))
k_batch_step_t(const SimDesc* __restrict__ sims, const TileItem* __restrict__ items,
               int n_items, const uint32_t* __restrict__ masks1,
               const uint32_t* __restrict__ masks2, int parity) {
  const int it = __builtin_amdgcn_readfirstlane(blockIdx.x * 4 + (int)(threadIdx.x >> 6));
  if (it >= n_items) return;
  TileItem item = items[it];
  item.sim = __builtin_amdgcn_readfirstlane(item.sim);
  item.cs = __builtin_amdgcn_readfirstlane(item.cs);
  item.rs = __builtin_amdgcn_readfirstlane(item.rs);
  item.flags = __builtin_amdgcn_readfirstlane(item.flags);
  const SimDesc& S = sims[item.sim];
  const uint32_t m1 = __builtin_amdgcn_readfirstlane(masks1[item.sim]);
  const uint32_t m2 = __builtin_amdgcn_readfirstlane(masks2[item.sim]);
  const int n_stim = uni(S.n_stim);
  const BatchStimView sv1{S.stim, n_stim, m1}, sv2{S.stim, n_stim, m2};
  Coef c = S.c;
  c.V_c = uni(c.V_c); c.V_v = uni(c.V_v); c.V_csi = uni(c.V_csi); c.kk = uni(c.kk);
  c.kk_off = uni(c.kk_off);
  c.inv_tvp = uni(c.inv_tvp); c.inv_tv1m = uni(c.inv_tv1m); c.inv_tv2m = uni(c.inv_tv2m);
  c.inv_twp = uni(c.inv_twp); c.inv_twm = uni(c.inv_twm); c.inv_td = uni(c.inv_td);
  c.inv_t0 = uni(c.inv_t0); c.inv_tr = uni(c.inv_tr); c.inv_tsi = uni(c.inv_tsi);
  c.inv_cm = uni(c.inv_cm);
  const float* Dp = uni(S.D);
  const float* i0 = uni(S.buf[parity][0]);
  const float* i1 = uni(S.buf[parity][1]);
  const float* i2 = uni(S.buf[parity][2]);
  float* o0 = uni(S.buf[parity ^ 1][0]);
  float* o1 = uni(S.buf[parity ^ 1][1]);
  float* o2 = uni(S.buf[parity ^ 1][2]);
  const int H = uni(S.H), W = uni(S.W);
  const float dt = uni(S.dt), inv_dx2 = uni(S.inv_dx2);
  // flags: bits 0-1 the band's kind (0 interior, 1 first, 2 last), bit 2 a
  // strip at a side
  const int sel = (item.flags & 3) * 4 + ((item.flags & 4) ? 2 : 0) + ((m1 | m2) ? 1 : 0);
#define FK_BT2(X, Y, ST)                                                                  \
  stream_t_tile<R, X, Y, ST, BatchStimView>(c, Dp, i0, i1, i2, H, W, sv1, sv2, dt, inv_dx2, \
                                            o0, o1, o2, item.rs, item.cs)
  switch (sel) {
    case 0: FK_BT2(0, false, false); break;
    case 1: FK_BT2(0, false, true); break;
    case 2: FK_BT2(0, true, false); break;
    case 3: FK_BT2(0, true, true); break;
    case 4: FK_BT2(1, false, false); break;
    case 5: FK_BT2(1, false, true); break;
    case 6: FK_BT2(1, true, false); break;
    case 7: FK_BT2(1, true, true); break;
    case 8: FK_BT2(2, false, false); break;
    case 9: FK_BT2(2, false, true); break;
    case 10: FK_BT2(2, true, false); break;
    default: FK_BT2(2, true, true); break;
  }
#undef FK_BT2
}

__global__ void k_batch_bump(int64_t* step_base, int64_t by) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *step_base += by;
}

__global__ void k_batch_set(int64_t* step_base, int64_t value) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *step_base = value;
}

// After a run: every simulation's next first step is n_steps later.
__global__ void k_batch_advance(SimDesc* sims, int n, int64_t n_steps) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) sims[i].t0 += (double)n_steps;
}

}  // namespace fk

using namespace fk;

// Opaque batch handle.
struct fk_batch {
  int n_sims = 0;
  int opt = kDefaultRowOpt;  // index into kBatchRows
  int n_items_opt[kNumRowOpts] = {};
  std::vector<fk_sim> host_sims;   // caller's descriptors (state pointers)
  std::vector<size_t> offset;      // per-sim cell offset in the pools
  size_t total_cells = 0;
  char* dev = nullptr;             // one allocation: descriptors, items, counter, pools
  SimDesc* d_sims = nullptr;
  TileItem* d_items[kNumRowOpts] = {};
  int64_t* d_step = nullptr;
  uint32_t* d_mask = nullptr;  // stimulus activity: kGraphSteps x n_sims
  float* pool[2] = {nullptr, nullptr};  // ping-pong pools of 3 * total_cells floats
  hipStream_t cap = nullptr;            // private stream used for graph capture
  hipGraphExec_t exec[2] = {nullptr, nullptr};  // per starting parity
};

static int32_t batch_check_sim(const fk_sim& s, int i) {
  if (!check_params(&s.params) || !check_grid(s.H, s.W) ||
      !check_ptrs({s.D, s.v, s.w, s.u}) || !check_time(s.t, s.t) ||
      !check_scalar("dt", s.dt) || !check_dx(s.dx))
    return set_error(last_code(), "simulation %d: %s", i, fk_last_error());
  if (s.n_stimuli < 0 || s.n_stimuli > FK_BATCH_MAX_STIMULI)
    return set_error(FK_ERR_UNSUPPORTED, "simulation %d: %d stimuli, the batch limit is %d", i,
                     s.n_stimuli, FK_BATCH_MAX_STIMULI);
  if (!check_stimuli(s.stimuli, s.n_stimuli))
    return set_error(last_code(), "simulation %d: %s", i, fk_last_error());
  return FK_OK;
}

static void batch_free(fk_batch* b) {
  if (!b) return;
  for (auto& e : b->exec)
    if (e) (void)hipGraphExecDestroy(e);
  if (b->cap) (void)hipStreamDestroy(b->cap);
  if (b->dev) (void)hipFree(b->dev);
  delete b;
}

extern "C" {

int32_t fk_batch_create(const fk_sim* sims, int32_t n, fk_batch** out) {
  ErrorScope es;
  if (out == nullptr) return set_error(FK_ERR_ARG, "out is null");
  *out = nullptr;
  if (n <= 0 || sims == nullptr) return set_error(FK_ERR_ARG, "empty batch");
  for (int i = 0; i < n; ++i) {
    const int32_t rc = batch_check_sim(sims[i], i);
    if (rc != FK_OK) return rc;
  }
  fk_batch* b = new fk_batch();
  b->n_sims = n;
  b->host_sims.assign(sims, sims + n);
  std::vector<TileItem> items[kNumRowOpts];
  b->offset.resize(n);
  for (int i = 0; i < n; ++i) {
    const fk_sim& s = sims[i];
    b->offset[i] = b->total_cells;
    b->total_cells += (size_t)s.H * s.W;
    for (int o = 0; o < kNumRowOpts; ++o) {
      const int R = kBatchRows[o];
      if (kBatchTwo[o]) {  // two-step bands, placed from both ends of each axis
        const StreamTGeom g = o == 4 ? make_stream_t_geom<8>(s.H, s.W)
                                     : make_stream_t_geom<16>(s.H, s.W);
        for (int strip = 0; strip < g.n_strips; ++strip)
          for (int band = 0; band < g.n_bands; ++band) {
            TileItem t;
            t.sim = i;
            t.cs = stream_t_origin(strip, g.n_strips, kTOut, s.W);
            t.rs = stream_t_origin(band, g.n_bands, R, s.H);
            const int xe = t.rs == 0 ? 1 : (t.rs == s.H - R ? 2 : 0);
            const bool yg = t.cs == 0 || t.cs == s.W - kTOut;
            t.flags = xe | (yg ? 4 : 0);
            items[o].push_back(t);
          }
        continue;
      }
      const StreamGeom g = o == 0   ? make_stream_geom<kBatchRows[0]>(s.H, s.W)
                           : o == 1 ? make_stream_geom<kBatchRows[1]>(s.H, s.W)
                           : o == 2 ? make_stream_geom<kBatchRows[2]>(s.H, s.W)
                                    : make_stream_geom<kBatchRows[3]>(s.H, s.W);
      for (int strip = 0; strip < g.n_strips; ++strip)
        for (int band = 0; band < g.n_bands; ++band) {
          TileItem t;
          t.sim = i;
          t.cs = std::min(kStripOut * strip, g.cs_last);
          t.rs = std::min(R * band, g.rs_last);
          const bool xg = edge_band(t.rs, R, s.H, s.W);
          const bool yg = t.cs < 3 || t.cs + kStripOut + 3 > s.W;
          t.flags = (xg ? 1 : 0) | (yg ? 2 : 0);
          items[o].push_back(t);
        }
    }
  }
  // the two-step options exist only when every simulation fits them
  for (int o = 0; o < kNumRowOpts; ++o) {
    if (!kBatchTwo[o]) continue;
    bool all_fit = true;
    for (int i = 0; i < n; ++i)
      all_fit = all_fit && (kBatchRows[o] == 8 ? stream_t_fits<8>(sims[i].H, sims[i].W)
                                               : stream_t_fits<16>(sims[i].H, sims[i].W));
    if (!all_fit) items[o].clear();
  }
  size_t sz_it[kNumRowOpts], sz_items = 0;
  for (int o = 0; o < kNumRowOpts; ++o) {
    b->n_items_opt[o] = (int)items[o].size();
    sz_it[o] = ((sizeof(TileItem) * items[o].size() + 255) / 256) * 256;
    sz_items += sz_it[o];
  }
  const size_t sz_sims = ((sizeof(SimDesc) * n + 255) / 256) * 256;
  const size_t sz_pool = 3 * b->total_cells * sizeof(float);
  const size_t sz_mask = ((sizeof(uint32_t) * n * kGraphSteps + 255) / 256) * 256;
  const size_t total = sz_sims + sz_items + 256 + sz_mask + 2 * sz_pool;
  hipError_t e = hipMalloc((void**)&b->dev, total);
  if (e != hipSuccess) {
    batch_free(b);
    return set_error(FK_ERR_HIP, "hipMalloc(%zu) failed: %s", total, hipGetErrorString(e));
  }
  b->d_sims = (SimDesc*)b->dev;
  {
    size_t at = sz_sims;
    for (int o = 0; o < kNumRowOpts; ++o) {
      b->d_items[o] = (TileItem*)(b->dev + at);
      at += sz_it[o];
    }
  }
  b->d_step = (int64_t*)(b->dev + sz_sims + sz_items);
  b->d_mask = (uint32_t*)(b->dev + sz_sims + sz_items + 256);
  b->pool[0] = (float*)(b->dev + sz_sims + sz_items + 256 + sz_mask);
  b->pool[1] = b->pool[0] + 3 * b->total_cells;
  std::vector<SimDesc> desc(n);
  for (int i = 0; i < n; ++i) {
    const fk_sim& s = sims[i];
    SimDesc& d = desc[i];
    std::memset(&d, 0, sizeof d);
    d.c = make_coef(s.params);
    d.D = s.D;
    const size_t cells = (size_t)s.H * s.W;
    for (int p = 0; p < 2; ++p)
      for (int f = 0; f < 3; ++f) d.buf[p][f] = b->pool[p] + 3 * b->offset[i] + f * cells;
    d.H = s.H;
    d.W = s.W;
    d.n_stim = s.n_stimuli;
    d.dt = s.dt;
    d.inv_dx2 = (float)(1.0 / ((double)s.dx * (double)s.dx));
    d.t0 = s.t;
    for (int k = 0; k < s.n_stimuli; ++k)
      d.stim[k] = {s.stimuli[k].start, s.stimuli[k].duration, s.stimuli[k].period,
                   s.stimuli[k].field};
  }
  e = hipMemcpy(b->d_sims, desc.data(), sizeof(SimDesc) * n, hipMemcpyHostToDevice);
  for (int o = 0; o < kNumRowOpts && e == hipSuccess; ++o)
    if (!items[o].empty())
      e = hipMemcpy(b->d_items[o], items[o].data(), sizeof(TileItem) * items[o].size(),
                    hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&b->cap, hipStreamNonBlocking);
  if (e != hipSuccess) {
    batch_free(b);
    return set_error(FK_ERR_HIP, "batch setup failed: %s", hipGetErrorString(e));
  }
  *out = b;
  return FK_OK;
}

int32_t fk_batch_destroy(fk_batch* b) {
  ErrorScope es;
  if (b == nullptr) return FK_OK;
  HIP_CHECK(hipDeviceSynchronize());
  batch_free(b);
  return FK_OK;
}

int32_t fk_batch_n_items(const fk_batch* b) { return b ? b->n_items_opt[b->opt] : 0; }

// Stimulus masks of the next n_rel (<= kGraphSteps) steps from the counter.
static hipError_t batch_masks(fk_batch* b, int n_rel, hipStream_t s) {
  const int total = b->n_sims * n_rel;
  k_batch_masks<<<(total + 255) / 256, 256, 0, s>>>(b->d_sims, b->n_sims, b->d_step, n_rel,
                                                    b->d_mask);
  return hipGetLastError();
}

// One launch of option o at step rel of the current mask window (its masks
// computed): two steps when o is a two-step option and `two` is set, else
// one. Reads pool[parity], writes pool[parity ^ 1].
static hipError_t batch_launch(fk_batch* b, int o, bool two, int rel, int parity,
                               hipStream_t s) {
  if (!two) o = one_step_opt(o);
  const int n_items = b->n_items_opt[o];
  const int blocks = (n_items + 3) / 4;
  const TileItem* it = b->d_items[o];
  const uint32_t* m1 = b->d_mask + (size_t)rel * b->n_sims;
  switch (o) {
    case 0: k_batch_step<kBatchRows[0]><<<blocks, 256, 0, s>>>(b->d_sims, it, n_items, m1, parity); break;
    case 1: k_batch_step<kBatchRows[1]><<<blocks, 256, 0, s>>>(b->d_sims, it, n_items, m1, parity); break;
    case 2: k_batch_step<kBatchRows[2]><<<blocks, 256, 0, s>>>(b->d_sims, it, n_items, m1, parity); break;
    case 3: k_batch_step<kBatchRows[3]><<<blocks, 256, 0, s>>>(b->d_sims, it, n_items, m1, parity); break;
    case 4: k_batch_step_t<8><<<blocks, 256, 0, s>>>(b->d_sims, it, n_items, m1, m1 + b->n_sims, parity); break;
    default: k_batch_step_t<16><<<blocks, 256, 0, s>>>(b->d_sims, it, n_items, m1, m1 + b->n_sims, parity); break;
  }
  return hipGetLastError();
}

// n (<= kGraphSteps) steps of option o from the counter on: their masks, the
// step launches (two steps each where possible) and the counter bump. Starts
// at pool[*parity] and leaves the pool of the result in *parity.
static hipError_t batch_chunk(fk_batch* b, int o, int n, int* parity, hipStream_t s) {
  hipError_t e = batch_masks(b, n, s);
  for (int r = 0; r < n && e == hipSuccess;) {
    const bool two = kBatchTwo[o] && n - r >= 2;
    e = batch_launch(b, o, two, r, *parity, s);
    *parity ^= 1;
    r += two ? 2 : 1;
  }
  if (e == hipSuccess) {
    k_batch_bump<<<1, 64, 0, s>>>(b->d_step, n);
    e = hipGetLastError();
  }
  return e;
}

// Launches of one chunk of kGraphSteps steps of option o, and so the parity
// change a recorded graph makes.
static int chunk_launches(int o) { return kBatchTwo[o] ? kGraphSteps / 2 : kGraphSteps; }

// Records a chunk of kGraphSteps steps starting at `parity`.
static hipError_t batch_build_graph(fk_batch* b, int parity) {
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamBeginCapture(b->cap, hipStreamCaptureModeThreadLocal);
  if (e != hipSuccess) return e;
  int p = parity;
  e = batch_chunk(b, b->opt, kGraphSteps, &p, b->cap);
  const hipError_t e2 = hipStreamEndCapture(b->cap, &g);
  if (e == hipSuccess) e = e2;
  if (e == hipSuccess) e = hipGraphInstantiate(&b->exec[parity], g, nullptr, nullptr, 0);
  if (g) (void)hipGraphDestroy(g);
  return e;
}

int32_t fk_batch_autotune(fk_batch* b, int32_t n_steps, void* stream, float* times_ms,
                          int32_t* best_rows) {
  ErrorScope es;
  if (b == nullptr) return set_error(FK_ERR_ARG, "batch is null");
  if (n_steps < 1 || n_steps > 10000)
    return set_error(FK_ERR_ARG, "n_steps must be in [1, 10000], got %d", n_steps);
  hipStream_t s = (hipStream_t)stream;
  // Queued runs may still replay this batch's graphs or use its pools.
  HIP_CHECK(hipDeviceSynchronize());
  // The pools are scratch between runs (fk_batch_run copies the state in),
  // so timing may overwrite them; the step counter is reset by every run.
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipError_t e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  float ms[kNumRowOpts];
  for (int o = 0; o < kNumRowOpts; ++o) ms[o] = -1.0f;
  for (int o = 0; o < kNumRowOpts && e == hipSuccess; ++o) {
    if (b->n_items_opt[o] == 0) continue;  // a two-step option some simulation does not fit
    int parity = 0;
    k_batch_set<<<1, 64, 0, s>>>(b->d_step, 0);
    e = batch_chunk(b, o, 2, &parity, s);  // warm-up
    if (e == hipSuccess) e = hipEventRecord(e0, s);
    for (int k = 0; k < n_steps && e == hipSuccess; k += kGraphSteps)
      e = batch_chunk(b, o, std::min(kGraphSteps, n_steps - k), &parity, s);
    if (e == hipSuccess) e = hipEventRecord(e1, s);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    if (e == hipSuccess) e = hipEventElapsedTime(&ms[o], e0, e1);
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  int best = kDefaultRowOpt;
  if (e == hipSuccess)
    for (int o = 0; o < kNumRowOpts; ++o)
      if (ms[o] >= 0.0f && ms[o] < ms[best]) best = o;
  b->opt = best;
  // graphs recorded for another work list are stale
  for (auto& x : b->exec)
    if (x) {
      (void)hipGraphExecDestroy(x);
      x = nullptr;
    }
  if (e != hipSuccess) return set_error(FK_ERR_HIP, "batch autotune failed: %s",
                                        hipGetErrorString(e));
  if (times_ms)
    for (int o = 0; o < kNumRowOpts; ++o) times_ms[o] = ms[o];
  if (best_rows) *best_rows = kBatchRows[b->opt] | (kBatchTwo[b->opt] ? FK_BATCH_TWO_STEP : 0);
  return FK_OK;
}

int32_t fk_batch_set_rows(fk_batch* b, int32_t rows) {
  ErrorScope es;
  if (b == nullptr) return set_error(FK_ERR_ARG, "batch is null");
  int o = -1;
  const bool two = (rows & FK_BATCH_TWO_STEP) != 0;
  for (int k = 0; k < kNumRowOpts; ++k)
    if ((rows & ~FK_BATCH_TWO_STEP) == kBatchRows[k] && two == kBatchTwo[k]) o = k;
  if (o < 0)
    return set_error(FK_ERR_ARG,
                     "band height must be 2, 4, 8 or 16, or 8 or 16 with FK_BATCH_TWO_STEP, "
                     "got %d", rows);
  if (b->n_items_opt[o] == 0)
    return set_error(FK_ERR_UNSUPPORTED,
                     "two steps per launch need every simulation of at least 2R + 16 rows "
                     "and 56..96 or >= 104 columns");
  if (o != b->opt) {
    HIP_CHECK(hipDeviceSynchronize());  // no queued launch may still use the old list
    for (auto& x : b->exec)
      if (x) {
        (void)hipGraphExecDestroy(x);
        x = nullptr;
      }
    b->opt = o;
  }
  return FK_OK;
}

int32_t fk_batch_run(fk_batch* b, int64_t n_steps, int32_t use_graph, void* stream) {
  ErrorScope es;
  if (b == nullptr) return set_error(FK_ERR_ARG, "batch is null");
  if (n_steps < 0) return set_error(FK_ERR_ARG, "negative step count");
  for (int i = 0; i < b->n_sims; ++i)
    if (!check_time(b->host_sims[i].t, b->host_sims[i].t + (double)n_steps)) return last_code();
  hipStream_t s = (hipStream_t)stream;
  // state in: caller buffers -> pool 0
  for (int i = 0; i < b->n_sims; ++i) {
    const fk_sim& sm = b->host_sims[i];
    const size_t cells = (size_t)sm.H * sm.W;
    float* dst = b->pool[0] + 3 * b->offset[i];
    HIP_CHECK(hipMemcpyAsync(dst, sm.v, cells * 4, hipMemcpyDeviceToDevice, s));
    HIP_CHECK(hipMemcpyAsync(dst + cells, sm.w, cells * 4, hipMemcpyDeviceToDevice, s));
    HIP_CHECK(hipMemcpyAsync(dst + 2 * cells, sm.u, cells * 4, hipMemcpyDeviceToDevice, s));
  }
  k_batch_set<<<1, 64, 0, s>>>(b->d_step, 0);
  HIP_CHECK(hipGetLastError());
  int64_t k = 0;
  int parity = 0;
  if (use_graph) {
    while (n_steps - k >= kGraphSteps) {
      if (!b->exec[parity]) HIP_CHECK(batch_build_graph(b, parity));
      HIP_CHECK(hipGraphLaunch(b->exec[parity], s));
      k += kGraphSteps;
      parity ^= chunk_launches(b->opt) & 1;
    }
  }
  // the remainder (or everything without graphs): direct launches
  while (k < n_steps) {
    const int n = (int)std::min<int64_t>(kGraphSteps, n_steps - k);
    HIP_CHECK(batch_chunk(b, b->opt, n, &parity, s));
    k += n;
  }
  // state out: final pool -> caller buffers
  for (int i = 0; i < b->n_sims; ++i) {
    const fk_sim& sm = b->host_sims[i];
    const size_t cells = (size_t)sm.H * sm.W;
    const float* src = b->pool[parity] + 3 * b->offset[i];
    HIP_CHECK(hipMemcpyAsync(sm.v, src, cells * 4, hipMemcpyDeviceToDevice, s));
    HIP_CHECK(hipMemcpyAsync(sm.w, src + cells, cells * 4, hipMemcpyDeviceToDevice, s));
    HIP_CHECK(hipMemcpyAsync(sm.u, src + 2 * cells, cells * 4, hipMemcpyDeviceToDevice, s));
  }
  // the simulations' times advance by n_steps for a later run
  for (auto& sm : b->host_sims) sm.t += (double)n_steps;
  k_batch_advance<<<(b->n_sims + 255) / 256, 256, 0, s>>>(b->d_sims, b->n_sims, n_steps);
  HIP_CHECK(hipGetLastError());
  return FK_OK;
}

}  // extern "C"
