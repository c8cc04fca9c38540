// Batched rollout and its adjoint: the frame-producing forward solve of many
// independent problems and the checkpointed reverse pass of their frame loss,
// advanced together, one launch per Euler step (and per reverse step) for all
// of them.
//
// Each problem is the forward-Euler solve of cardiax/solve.py:92-100 from its
// own initial state, D-map, parameters and stimuli, with u sampled every
// `frame_every` steps; the reverse pass is the adjoint of fk_adjoint.hip.
// Per problem the schedule is the one the per-grid path runs (adjoint.py
// with fk_forward_trajectory and fk_adjoint_segment): checkpoints every S
// steps; for each segment, in reverse, the states are recomputed from its
// checkpoint, the derived fields of its last step are formed, and the steps
// are reversed one by one, frame cotangents entering lam_u at the frame
// steps. Every launch covers all problems through a work list (one wave per
// forward tile, one block per reverse tile), and the tiles run the same
// device code as the per-grid kernels: the results are bitwise those of the
// per-grid path.
//
// Reference: there is no counterpart in the reference, whose inverse problem
// is learnt by a network (deepx); the forward step follows
// cardiax/solve.py:26-100 and the loss is the frame MSE of the inverse
// problem built on it.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "cardiax_amd.h"
#include "fk_adjoint.h"
#include "fk_batch.h"
#include "fk_device.h"
#include "fk_host.h"
#include "fk_stream.h"

namespace fk {

constexpr int kRollRows = 4;  // band rows of the forward tiles

struct RollDesc {
  Coef c;
  int H, W, gx;  // gx: blocks of 64 columns of the reverse grid
  float dt, inv_dx2;
  size_t cells;
  const float* D;
  const float* x0[3];  // initial state (the caller's)
  float* ckpt;         // x_S, x_2S, ...: 3 planes each
  float* traj;         // one segment's states after its start: 3 planes each
  float* frames;       // forward output, one plane per frame
  const float* g_frames;  // backward input, one plane per frame
  float* gD;           // backward output
  float* lam;          // cotangent (v, w, u planes) at segment boundaries
  float* lam_alt;      // its ping-pong partner
  const AdjAxisTab* tab;
  const float *Dx, *Dy;
  double* part;        // tau accumulators
  AdjPre pre[2];
  int n_stim;
  BatchStim stim[FK_BATCH_MAX_STIMULI];
};

struct RollTile {
  int sim, cs, rs, flags;  // flags: bit0 band near a row edge, bit1 strip near a column edge
};

struct RollBlock {
  int sim, bx, by;
};

// A pointer field of a descriptor, read as a global-memory pointer. Pointers
// read from memory are generic, and the tile bodies index them directly, so
// without this they compile to flat loads (which also wait on the LDS
// counter). The field's bits are loaded as an address-space-1 pointer and
// then converted, which lets the compiler keep every access global. Every
// field lives in device memory.
template <class T>
__device__ __forceinline__ T* gfield(T* const& f) {
  typedef T __attribute__((address_space(1)))* GT;
  GT g;
  __builtin_memcpy(&g, &f, sizeof g);
  return (T*)g;
}
__device__ __forceinline__ AdjPre gfield(const AdjPre& a) {
  return AdjPre{gfield(a.Llap), gfield(a.QPx), gfield(a.QPy), gfield(a.Q2)};
}

// State x_{c + k} of segment `seg` (c = seg S): the checkpoint at k = 0, else
// slot k - 1 of the segment's trajectory. Plane f: 0 v, 1 w, 2 u.
__device__ __forceinline__ const float* roll_state(const RollDesc& S, int seg, int k, int f) {
  if (k == 0)
    return seg == 0 ? gfield(S.x0[f]) : gfield(S.ckpt) + ((size_t)(seg - 1) * 3 + f) * S.cells;
  return gfield(S.traj) + ((size_t)(k - 1) * 3 + f) * S.cells;
}

// One Euler step of every tile: x_{c+k} -> trajectory slot k. masks: this
// step's stimulus activity, one word per problem.
template <int R>
__global__ void __launch_bounds__(256)
k_roll_step(const RollDesc* __restrict__ sims, const RollTile* __restrict__ items, int n_items,
            const uint32_t* __restrict__ masks, int seg, int k) {
  const int it = __builtin_amdgcn_readfirstlane(blockIdx.x * 4 + (int)(threadIdx.x >> 6));
  if (it >= n_items) return;
  RollTile item = items[it];
  if (uni(item.sim) < 0) return;  // padding of the XCD order
  item.sim = uni(item.sim);
  item.cs = uni(item.cs);
  item.rs = uni(item.rs);
  item.flags = uni(item.flags);
  const RollDesc& S = sims[item.sim];
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
  const float* i0 = uni(roll_state(S, seg, k, 0));
  const float* i1 = uni(roll_state(S, seg, k, 1));
  const float* i2 = uni(roll_state(S, seg, k, 2));
  const size_t cells = S.cells;
  float* o0 = uni(S.traj + ((size_t)k * 3 + 0) * cells);
  float* o1 = uni(S.traj + ((size_t)k * 3 + 1) * cells);
  float* o2 = uni(S.traj + ((size_t)k * 3 + 2) * cells);
  const int H = uni(S.H), W = uni(S.W);
  const float dt = uni(S.dt), inv_dx2 = uni(S.inv_dx2);
  const int sel = item.flags | (mask ? 4 : 0);
#define FK_RS(X, Y, ST)                                                                   \
  stream_tile<R, true, X, Y, ST, BatchStimView, false, 0, true>(c, Dp, nullptr, i0, i1, i2, H, \
                                                                W, sv, dt, inv_dx2, o0, o1,   \
                                                                o2, item.rs, item.cs)
  switch (sel) {
    case 0: FK_RS(false, false, false); break;
    case 1: FK_RS(true, false, false); break;
    case 2: FK_RS(false, true, false); break;
    case 3: FK_RS(true, true, false); break;
    case 4: FK_RS(false, false, true); break;
    case 5: FK_RS(true, false, true); break;
    case 6: FK_RS(false, true, true); break;
    default: FK_RS(true, true, true); break;
  }
#undef FK_RS
}

// After forward segment `seg` (steps c .. c + len - 1): the frames it holds
// (states x_j, j in (c, c + len], j a multiple of fe) and, when `ckpt`, its
// last state as the next checkpoint. blockIdx.y: problem.
__global__ void __launch_bounds__(256)
k_roll_gather(const RollDesc* __restrict__ sims, int seg, int c, int len, int fe, int ckpt) {
  const RollDesc& S = sims[blockIdx.y];
  const size_t cells = S.cells;
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < cells;
       i += (size_t)gridDim.x * 256) {
    for (int j = (c / fe + 1) * fe; j <= c + len; j += fe)
      S.frames[(size_t)(j / fe - 1) * cells + i] = S.traj[((size_t)(j - c - 1) * 3 + 2) * cells + i];
    if (ckpt)
      for (int f = 0; f < 3; ++f)
        S.ckpt[((size_t)seg * 3 + f) * cells + i] = S.traj[((size_t)(len - 1) * 3 + f) * cells + i];
  }
}

// dL/dD and the cotangent start at zero.
__global__ void __launch_bounds__(256) k_roll_zero(const RollDesc* __restrict__ sims) {
  const RollDesc& S = sims[blockIdx.y];
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < S.cells;
       i += (size_t)gridDim.x * 256) {
    S.gD[i] = 0.0f;
    for (int f = 0; f < 3; ++f) S.lam[(size_t)f * S.cells + i] = 0.0f;
  }
}

// The cotangent back from the ping-pong partner (a segment of odd length).
__global__ void __launch_bounds__(256) k_roll_lam_back(const RollDesc* __restrict__ sims) {
  const RollDesc& S = sims[blockIdx.y];
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < 3 * S.cells;
       i += (size_t)gridDim.x * 256)
    S.lam[i] = S.lam_alt[i];
}

// Frame cotangent injected into x_j (null when j is not a frame step).
__device__ __forceinline__ const float* roll_inject(const RollDesc& S, int j, int fe) {
  return (j > 0 && j % fe == 0) ? gfield(S.g_frames) + (size_t)(j / fe - 1) * S.cells : nullptr;
}

// Derived fields (set 0) of the last step k of segment `seg`, whose output
// x_j (j = c + k + 1) takes the frame cotangent of j.
__global__ void __launch_bounds__(kAdjThreads)
k_roll_pre(const RollDesc* __restrict__ sims, const RollBlock* __restrict__ blocks, int seg,
           int k, int j, int fe) {
  __shared__ AdjAxisTab tab[2];
  const RollBlock b = blocks[blockIdx.x];
  if (b.sim < 0) return;  // padding of the XCD order (block-uniform)
  const RollDesc& S = sims[b.sim];
  adj_pre_tile(roll_state(S, seg, k, 2), gfield(S.D), gfield(S.Dx), gfield(S.Dy),
               gfield(S.lam) + 2 * S.cells, roll_inject(S, j, fe), S.H, S.W, gfield(S.tab),
               gfield(S.pre[0]), b.bx, b.by, tab);
}

// Reverse of step c + k of segment `seg`: cotangent set `cur` (0 lam, 1
// lam_alt) -> the other, derived fields set `set`; with NEXT (k > 0) also the
// derived fields of step c + k - 1 into set 1 - set.
template <bool NEXT>
__global__ void __launch_bounds__(kAdjThreads)
k_roll_main(const RollDesc* __restrict__ sims, const RollBlock* __restrict__ blocks,
            const uint32_t* __restrict__ masks, int seg, int k, int fe, int cur, int set,
            int c0) {
  __shared__ AdjAxisTab tab[2];
  __shared__ float4 xch[kAdjThreads / 2];
  const RollBlock b = blocks[blockIdx.x];
  if (b.sim < 0) return;  // padding of the XCD order (block-uniform)
  const RollDesc& S = sims[b.sim];
  const size_t n = S.cells;
  float* lin = cur ? gfield(S.lam_alt) : gfield(S.lam);
  float* lout = cur ? gfield(S.lam) : gfield(S.lam_alt);
  const AdjNext nx{NEXT ? roll_state(S, seg, k - 1, 2) : nullptr,
                   NEXT ? roll_inject(S, c0 + k, fe) : nullptr, gfield(S.D), gfield(S.Dx),
                   gfield(S.Dy), gfield(S.pre[1 - set])};
  const BatchStimView sv{S.stim, S.n_stim, masks[b.sim]};
  adj_main_tile<NEXT>(S.c, roll_state(S, seg, k, 0), roll_state(S, seg, k, 1),
                      roll_state(S, seg, k, 2), S.H, S.W, sv, S.dt, S.inv_dx2, lin, lin + n,
                      gfield(S.pre[set]), lout, lout + n, lout + 2 * n, gfield(S.gD),
                      gfield(S.part), gfield(S.tab), nx, b.bx, b.by, S.gx, tab, xch);
}

}  // namespace fk

using namespace fk;

struct fk_rollout_batch {
  int n = 0, n_steps = 0, fe = 1, S = 1, n_seg = 0;
  std::vector<fk_rollout_spec> specs;
  std::vector<size_t> work_off;  // adjoint workspace of each problem (byte offset)
  std::vector<size_t> work_bytes;
  char* dev = nullptr;
  size_t dev_bytes = 0;
  RollDesc* d_sims = nullptr;
  RollTile* d_tiles = nullptr;
  RollBlock* d_pre = nullptr;
  RollBlock* d_main = nullptr;
  uint32_t* d_masks = nullptr;  // n_steps x n
  int n_tiles = 0, n_pre = 0, n_main = 0;
  size_t max_cells = 0;
  int traj_seg = -1;  // segment whose states the trajectory buffer holds (-1: none)
};

static int32_t roll_check(const fk_rollout_spec& s, int i) {
  if (!check_params(&s.params) || !check_grid(s.H, s.W) ||
      !check_ptrs({s.D, s.v0, s.w0, s.u0, s.frames, s.g_frames, s.grad_D}) ||
      !check_time(s.t0, s.t0) || !check_scalar("dt", s.dt) || !check_dx(s.dx))
    return set_error(last_code(), "problem %d: %s", i, fk_last_error());
  if (s.n_stimuli < 0 || s.n_stimuli > FK_BATCH_MAX_STIMULI)
    return set_error(FK_ERR_UNSUPPORTED, "problem %d: %d stimuli, the limit is %d", i,
                     s.n_stimuli, FK_BATCH_MAX_STIMULI);
  if (!check_stimuli(s.stimuli, s.n_stimuli))
    return set_error(last_code(), "problem %d: %s", i, fk_last_error());
  return FK_OK;
}

static size_t align256(size_t x) { return (x + 255) / 256 * 256; }

static dim3 roll_cell_grid(const fk_rollout_batch* b) {
  const size_t blocks = std::min<size_t>((b->max_cells + 255) / 256, 1024);
  return dim3((unsigned)std::max<size_t>(blocks, 1), (unsigned)b->n);
}

static void roll_free(fk_rollout_batch* b) {
  if (!b) return;
  if (b->dev) (void)hipFree(b->dev);
  delete b;
}

extern "C" {

int32_t fk_rollout_batch_create(const fk_rollout_spec* specs, int32_t n, int32_t n_steps,
                                int32_t frame_every, int32_t checkpoint_every,
                                fk_rollout_batch** out) {
  ErrorScope es;
  if (out == nullptr) return set_error(FK_ERR_ARG, "out is null");
  *out = nullptr;
  if (n <= 0 || specs == nullptr) return set_error(FK_ERR_ARG, "empty batch");
  if (n_steps < 1 || frame_every < 1 || checkpoint_every < 1)
    return set_error(FK_ERR_ARG, "n_steps, frame_every and checkpoint_every must be >= 1 "
                     "(got %d, %d, %d)", n_steps, frame_every, checkpoint_every);
  for (int i = 0; i < n; ++i) {
    const int32_t rc = roll_check(specs[i], i);
    if (rc != FK_OK) return rc;
    if (!check_time(specs[i].t0, specs[i].t0 + (double)n_steps)) return last_code();
  }
  fk_rollout_batch* b = new fk_rollout_batch();
  b->n = n;
  b->n_steps = n_steps;
  b->fe = frame_every;
  b->S = std::min(checkpoint_every, n_steps);
  b->n_seg = (n_steps + b->S - 1) / b->S;
  b->specs.assign(specs, specs + n);
  // work lists: runs of vertically adjacent tiles (which share halo rows),
  // placed whole on one XCD each
  constexpr int kRunTiles = 16, kRunPre = 8, kRunMain = 16;
  std::vector<std::vector<RollTile>> tile_runs;
  std::vector<std::vector<RollBlock>> pre_runs, main_runs;
  for (int i = 0; i < n; ++i) {
    const fk_rollout_spec& s = specs[i];
    const StreamGeom g = make_stream_geom<kRollRows>(s.H, s.W);
    for (int strip = 0; strip < g.n_strips; ++strip)
      for (int band = 0; band < g.n_bands; ++band) {
        RollTile t;
        t.sim = i;
        t.cs = std::min(kStripOut * strip, g.cs_last);
        t.rs = std::min(kRollRows * band, g.rs_last);
        const bool xg = edge_band(t.rs, kRollRows, s.H, s.W);
        const bool yg = t.cs < 3 || t.cs + kStripOut + 3 > s.W;
        t.flags = (xg ? 1 : 0) | (yg ? 2 : 0);
        if (band % kRunTiles == 0) tile_runs.emplace_back();
        tile_runs.back().push_back(t);
      }
    const int gx = (s.W + 63) / 64;
    for (int bx = 0; bx < gx; ++bx)
      for (int by = 0; by < (s.H + 3) / 4; ++by) {
        if (by % kRunPre == 0) pre_runs.emplace_back();
        pre_runs.back().push_back(RollBlock{i, bx, by});
      }
    for (int bx = 0; bx < gx; ++bx)
      for (int by = 0; by < (s.H + 1) / 2; ++by) {
        if (by % kRunMain == 0) main_runs.emplace_back();
        main_runs.back().push_back(RollBlock{i, bx, by});
      }
  }
  const std::vector<RollTile> tiles = xcd_order(tile_runs, 4, RollTile{-1, 0, 0, 0});
  const std::vector<RollBlock> pre = xcd_order(pre_runs, 1, RollBlock{-1, 0, 0});
  const std::vector<RollBlock> main = xcd_order(main_runs, 1, RollBlock{-1, 0, 0});
  b->n_tiles = (int)tiles.size();
  b->n_pre = (int)pre.size();
  b->n_main = (int)main.size();
  // stimulus activity of every step, as the per-grid path evaluates it
  std::vector<uint32_t> masks((size_t)n_steps * n);
  for (int i = 0; i < n; ++i) {
    StimSchedule sched(specs[i].stimuli, specs[i].n_stimuli);
    for (int k = 0; k < n_steps; ++k)
      masks[(size_t)k * n + i] = sched.mask(specs[i].t0 + (double)k);
  }
  // one allocation: descriptors, work lists, masks, then per problem its
  // checkpoints, segment trajectory, cotangent and adjoint workspace
  size_t o = 0;
  auto take = [&](size_t bytes) {
    const size_t at = o;
    o += align256(bytes);
    return at;
  };
  const size_t o_sims = take(sizeof(RollDesc) * n);
  const size_t o_tiles = take(sizeof(RollTile) * tiles.size());
  const size_t o_pre = take(sizeof(RollBlock) * pre.size());
  const size_t o_main = take(sizeof(RollBlock) * main.size());
  const size_t o_masks = take(sizeof(uint32_t) * masks.size());
  std::vector<size_t> o_ckpt(n), o_traj(n), o_lam(n);
  b->work_off.resize(n);
  b->work_bytes.resize(n);
  for (int i = 0; i < n; ++i) {
    const size_t cells = (size_t)specs[i].H * specs[i].W;
    b->max_cells = std::max(b->max_cells, cells);
    o_ckpt[i] = take((size_t)(b->n_seg - 1) * 3 * cells * sizeof(float));
    o_traj[i] = take((size_t)b->S * 3 * cells * sizeof(float));
    o_lam[i] = take(3 * cells * sizeof(float));
    b->work_bytes[i] = adj_layout(specs[i].H, specs[i].W).total;
    b->work_off[i] = take(b->work_bytes[i]);
  }
  b->dev_bytes = o;
  hipError_t e = hipMalloc((void**)&b->dev, o);
  if (e != hipSuccess) {
    roll_free(b);
    return set_error(FK_ERR_HIP, "hipMalloc(%zu) failed: %s", o, hipGetErrorString(e));
  }
  b->d_sims = (RollDesc*)(b->dev + o_sims);
  b->d_tiles = (RollTile*)(b->dev + o_tiles);
  b->d_pre = (RollBlock*)(b->dev + o_pre);
  b->d_main = (RollBlock*)(b->dev + o_main);
  b->d_masks = (uint32_t*)(b->dev + o_masks);
  std::vector<RollDesc> desc(n);
  for (int i = 0; i < n; ++i) {
    const fk_rollout_spec& s = specs[i];
    RollDesc& d = desc[i];
    std::memset(&d, 0, sizeof d);
    const size_t cells = (size_t)s.H * s.W;
    d.c = make_coef(s.params);
    d.H = s.H;
    d.W = s.W;
    d.gx = (s.W + 63) / 64;
    d.dt = s.dt;
    d.inv_dx2 = (float)(1.0 / ((double)s.dx * (double)s.dx));
    d.cells = cells;
    d.D = s.D;
    d.x0[0] = s.v0;
    d.x0[1] = s.w0;
    d.x0[2] = s.u0;
    d.ckpt = (float*)(b->dev + o_ckpt[i]);
    d.traj = (float*)(b->dev + o_traj[i]);
    d.frames = s.frames;
    d.g_frames = s.g_frames;
    d.gD = s.grad_D;
    d.lam = (float*)(b->dev + o_lam[i]);
    char* wk = b->dev + b->work_off[i];
    const AdjLayout L = adj_layout(s.H, s.W);
    d.lam_alt = (float*)(wk + L.lam);
    d.tab = (const AdjAxisTab*)(wk + L.tab);
    d.Dx = (const float*)(wk + L.dx);
    d.Dy = (const float*)(wk + L.dy);
    d.part = (double*)(wk + L.part);
    d.pre[0] = pre_fields(L, wk, 0, cells);
    d.pre[1] = pre_fields(L, wk, 1, cells);
    d.n_stim = s.n_stimuli;
    for (int k = 0; k < s.n_stimuli; ++k)
      d.stim[k] = {s.stimuli[k].start, s.stimuli[k].duration, s.stimuli[k].period,
                   s.stimuli[k].field};
  }
  e = hipMemcpy(b->d_sims, desc.data(), sizeof(RollDesc) * n, hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(b->d_tiles, tiles.data(), sizeof(RollTile) * tiles.size(),
                  hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(b->d_pre, pre.data(), sizeof(RollBlock) * pre.size(), hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(b->d_main, main.data(), sizeof(RollBlock) * main.size(),
                  hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(b->d_masks, masks.data(), sizeof(uint32_t) * masks.size(),
                  hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    roll_free(b);
    return set_error(FK_ERR_HIP, "rollout batch setup failed: %s", hipGetErrorString(e));
  }
  *out = b;
  return FK_OK;
}

size_t fk_rollout_batch_device_bytes(const fk_rollout_batch* b) { return b ? b->dev_bytes : 0; }

int32_t fk_rollout_batch_forward(fk_rollout_batch* b, void* stream) {
  ErrorScope es;
  if (b == nullptr) return set_error(FK_ERR_ARG, "null rollout batch");
  hipStream_t s = (hipStream_t)stream;
  const int step_blocks = (b->n_tiles + 3) / 4;
  const dim3 cg = roll_cell_grid(b);
  for (int seg = 0; seg < b->n_seg; ++seg) {
    const int c = seg * b->S, len = std::min(b->S, b->n_steps - c);
    for (int k = 0; k < len; ++k)
      k_roll_step<kRollRows><<<step_blocks, 256, 0, s>>>(
          b->d_sims, b->d_tiles, b->n_tiles, b->d_masks + (size_t)(c + k) * b->n, seg, k);
    k_roll_gather<<<cg, 256, 0, s>>>(b->d_sims, seg, c, len, b->fe,
                                     seg + 1 < b->n_seg ? 1 : 0);
  }
  HIP_CHECK(hipGetLastError());
  b->traj_seg = b->n_seg - 1;
  return FK_OK;
}

int32_t fk_rollout_batch_backward(fk_rollout_batch* b, float* g_tau, void* stream) {
  ErrorScope es;
  if (b == nullptr) return set_error(FK_ERR_ARG, "null rollout batch");
  hipStream_t s = (hipStream_t)stream;
  for (int i = 0; i < b->n; ++i) {  // boundary tables, D_x, D_y, tau accumulators
    const fk_rollout_spec& sp = b->specs[i];
    const int32_t rc = fk_adjoint_prepare(sp.D, sp.H, sp.W, b->dev + b->work_off[i],
                                          b->work_bytes[i], stream);
    if (rc != FK_OK) return rc;
  }
  const dim3 cg = roll_cell_grid(b);
  k_roll_zero<<<cg, 256, 0, s>>>(b->d_sims);
  const int step_blocks = (b->n_tiles + 3) / 4;
  for (int seg = b->n_seg - 1; seg >= 0; --seg) {
    const int c = seg * b->S, len = std::min(b->S, b->n_steps - c);
    // recompute x_{c+1} .. x_{c+len-1}, unless the forward's last segment is
    // still in place (the same states: same kernels from the same checkpoint)
    if (seg != b->traj_seg)
      for (int k = 0; k + 1 < len; ++k)
        k_roll_step<kRollRows><<<step_blocks, 256, 0, s>>>(
            b->d_sims, b->d_tiles, b->n_tiles, b->d_masks + (size_t)(c + k) * b->n, seg, k);
    b->traj_seg = seg;
    k_roll_pre<<<b->n_pre, kAdjThreads, 0, s>>>(b->d_sims, b->d_pre, seg, len - 1, c + len,
                                                b->fe);
    int cur = 0, set = 0;
    for (int k = len - 1; k >= 0; --k) {
      const uint32_t* m = b->d_masks + (size_t)(c + k) * b->n;
      if (k > 0)
        k_roll_main<true><<<b->n_main, kAdjThreads, 0, s>>>(b->d_sims, b->d_main, m, seg, k,
                                                            b->fe, cur, set, c);
      else
        k_roll_main<false><<<b->n_main, kAdjThreads, 0, s>>>(b->d_sims, b->d_main, m, seg, k,
                                                             b->fe, cur, set, c);
      cur ^= 1;
      set ^= 1;
    }
    if (cur == 1) k_roll_lam_back<<<cg, 256, 0, s>>>(b->d_sims);
  }
  HIP_CHECK(hipGetLastError());
  if (g_tau)
    for (int i = 0; i < b->n; ++i) {
      const fk_rollout_spec& sp = b->specs[i];
      const int32_t rc =
          fk_adjoint_tau_grads(b->dev + b->work_off[i], sp.H, sp.W, g_tau + 3 * i, stream);
      if (rc != FK_OK) return rc;
    }
  return FK_OK;
}

int32_t fk_rollout_batch_destroy(fk_rollout_batch* b) {
  ErrorScope es;
  if (b == nullptr) return FK_OK;
  HIP_CHECK(hipDeviceSynchronize());
  roll_free(b);
  return FK_OK;
}

}  // extern "C"
