// Batched forward Euler: many independent simulations of uneven sizes and
// parameter sets advanced together.
//
// Each Euler step is ONE launch over a work list that holds every (column
// strip, row band) tile of every simulation; a tile runs the streaming step
// (stream_tile, 8-row bands) with its simulation's constants, D-map and
// stimuli read from a device-resident descriptor. Stimulus activity is
// evaluated on the device with the law of cardiax/solve.py:262-267 (floor
// modulo in double precision, as on the host path). Steps are recorded once
// into a hipGraph of kGraphSteps launches plus a one-thread kernel that advances
// the device-side step counter, and the graph is replayed: the host issues one
// graph launch per kGraphSteps steps for the whole batch.
//
// Reference: the batch has no counterpart in the reference, which runs one
// simulation per call (experiments/generate_fd_data_128.py:25-56 loops over
// seeds); each simulation follows cardiax/solve.py:92-100.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <vector>

#include "cardiax_amd.h"
#include "fk_device.h"
#include "fk_host.h"
#include "fk_stream.h"

namespace fk {

// Rows per wave band: one work list per option, chosen by fk_batch_autotune.
constexpr int kNumRowOpts = 4;
constexpr int kBatchRows[kNumRowOpts] = {2, 4, 8, 16};
constexpr int kDefaultRowOpt = 2;  // 8 rows
constexpr int kGraphSteps = 32;   // Euler steps recorded in one graph (even)

struct BatchStim {
  double start, duration, period;
  const float* field;
};

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

__device__ __forceinline__ bool active_dev(double t, const BatchStim& s) {
  if (!(t >= s.start)) return false;
  double r = fmod(s.start - t + 1.0, s.period);
  if (r != 0.0 && ((r < 0.0) != (s.period < 0.0))) r += s.period;
  return r < s.duration;
}

// Stimulus activity of every simulation at its current step: bit k of
// mask[i] = stimulus k of simulation i is active.
__global__ void k_batch_activity(const SimDesc* __restrict__ sims, int n,
                                 const int64_t* __restrict__ step_base, int rel_step,
                                 uint32_t* __restrict__ mask) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const SimDesc& S = sims[i];
  const double t = S.t0 + (double)(*step_base + rel_step);
  uint32_t m = 0;
  for (int k = 0; k < S.n_stim; ++k)
    if (active_dev(t, S.stim[k])) m |= 1u << k;
  mask[i] = m;
}

// Wave-uniform copies (scalar registers) of descriptor values.
__device__ __forceinline__ int uni(int x) { return __builtin_amdgcn_readfirstlane(x); }
__device__ __forceinline__ float uni(float x) {
  return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(x)));
}
template <class P>
__device__ __forceinline__ P* uni(P* p) {
  const uint64_t a = (uint64_t)p;
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)a);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(a >> 32));
  return (P*)(((uint64_t)hi << 32) | lo);
}

// The simulation's stimuli with this step's activity mask.
struct BatchStimView {
  const BatchStim* s;
  int n;
  uint32_t mask;
  __device__ int count() const { return n; }
  __device__ bool on(int m) const { return (mask >> m) & 1u; }
  __device__ const float* field(int m) const { return s[m].field; }
};

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
  uint32_t* d_mask = nullptr;  // per-simulation stimulus activity of the current step
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
  size_t sz_it[kNumRowOpts], sz_items = 0;
  for (int o = 0; o < kNumRowOpts; ++o) {
    b->n_items_opt[o] = (int)items[o].size();
    sz_it[o] = ((sizeof(TileItem) * items[o].size() + 255) / 256) * 256;
    sz_items += sz_it[o];
  }
  const size_t sz_sims = ((sizeof(SimDesc) * n + 255) / 256) * 256;
  const size_t sz_pool = 3 * b->total_cells * sizeof(float);
  const size_t sz_mask = ((sizeof(uint32_t) * n + 255) / 256) * 256;
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

static hipError_t batch_step(fk_batch* b, int rel, int parity, hipStream_t s) {
  k_batch_activity<<<(b->n_sims + 63) / 64, 64, 0, s>>>(b->d_sims, b->n_sims, b->d_step, rel,
                                                        b->d_mask);
  const int n_items = b->n_items_opt[b->opt];
  const int blocks = (n_items + 3) / 4;
  const TileItem* it = b->d_items[b->opt];
  switch (b->opt) {
    case 0: k_batch_step<kBatchRows[0]><<<blocks, 256, 0, s>>>(b->d_sims, it, n_items, b->d_mask, parity); break;
    case 1: k_batch_step<kBatchRows[1]><<<blocks, 256, 0, s>>>(b->d_sims, it, n_items, b->d_mask, parity); break;
    case 2: k_batch_step<kBatchRows[2]><<<blocks, 256, 0, s>>>(b->d_sims, it, n_items, b->d_mask, parity); break;
    default: k_batch_step<kBatchRows[3]><<<blocks, 256, 0, s>>>(b->d_sims, it, n_items, b->d_mask, parity); break;
  }
  return hipGetLastError();
}

// Records kGraphSteps steps starting at `parity` plus the counter bump.
static hipError_t batch_build_graph(fk_batch* b, int parity) {
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamBeginCapture(b->cap, hipStreamCaptureModeThreadLocal);
  if (e != hipSuccess) return e;
  for (int k = 0; k < kGraphSteps && e == hipSuccess; ++k)
    e = batch_step(b, k, parity ^ (k & 1), b->cap);
  if (e == hipSuccess) {
    k_batch_bump<<<1, 64, 0, b->cap>>>(b->d_step, kGraphSteps);
    e = hipGetLastError();
  }
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
  float ms[kNumRowOpts] = {};
  for (int o = 0; o < kNumRowOpts && e == hipSuccess; ++o) {
    b->opt = o;
    k_batch_set<<<1, 64, 0, s>>>(b->d_step, 0);
    for (int k = 0; k < 2 && e == hipSuccess; ++k) e = batch_step(b, k, k & 1, s);
    if (e == hipSuccess) e = hipEventRecord(e0, s);
    for (int k = 0; k < n_steps && e == hipSuccess; ++k) e = batch_step(b, k, k & 1, s);
    if (e == hipSuccess) e = hipEventRecord(e1, s);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    if (e == hipSuccess) e = hipEventElapsedTime(&ms[o], e0, e1);
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  int best = kDefaultRowOpt;
  if (e == hipSuccess)
    for (int o = 0; o < kNumRowOpts; ++o)
      if (ms[o] < ms[best]) best = o;
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
  if (best_rows) *best_rows = kBatchRows[b->opt];
  return FK_OK;
}

int32_t fk_batch_set_rows(fk_batch* b, int32_t rows) {
  ErrorScope es;
  if (b == nullptr) return set_error(FK_ERR_ARG, "batch is null");
  int o = -1;
  for (int k = 0; k < kNumRowOpts; ++k)
    if (rows == kBatchRows[k]) o = k;
  if (o < 0) return set_error(FK_ERR_ARG, "band height must be 2, 4, 8 or 16, got %d", rows);
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
      k += kGraphSteps;  // kGraphSteps is even: parity unchanged
    }
  }
  // remainder (or everything without graphs): direct launches, counter bumped once
  for (int64_t r = 0; k + r < n_steps; ++r) {
    HIP_CHECK(batch_step(b, (int)r, parity, s));
    parity ^= 1;
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
