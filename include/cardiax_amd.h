/*
 * cardiax_amd -- C-ABI of the MI355X (gfx950) Fenton-Karma forward-Euler solver.
 *
 * The reference (cardiax, pure Python/JAX) has no native interface; its plug
 * point is the integrator callable of cardiax/solve.py:188,209-218,
 *     f(state, t, t_end, params, diffusivity, stimuli, dt, dx) -> State,
 * whose Euler implementation is _forward_euler (cardiax/solve.py:92-100).
 * Each entry point below replaces one public function of cardiax/solve.py and
 * is bound from Python with ctypes (cardiax_amd/_lib.py).
 *
 * Conventions
 *   - Fields are float32, C-order row-major (H, W): axis 0 (rows) is the
 *     reference's "x" (cardiax/solve.py:49), axis 1 (columns) its "y" (:50).
 *   - All array pointers are device pointers; structs and descriptor arrays
 *     are host memory read before the call returns.
 *   - Work is enqueued on `stream` (a hipStream_t; NULL = the null stream) and
 *     is asynchronous. Only the calls documented to wait synchronise:
 *     fk_check_finite (reads its flag back), fk_autotune_forward,
 *     fk_batch_autotune, fk_batch_set_rows and fk_batch_destroy.
 *   - Step indices must lie below 2^24 in magnitude (a loop may end at 2^24),
 *     where the reference's float32 and int32 step counters agree
 *     (FK_ERR_ARG beyond).
 *   - No entry point allocates device memory except fk_batch_create, which
 *     owns the device copy of the batch work list until fk_batch_destroy,
 *     and fk_autotune_forward, which frees its timing scratch before it
 *     returns.
 *   - Return value: FK_OK, or a negative FK_ERR_* code; fk_last_error()
 *     returns a thread-local message for the last failing call.
 */
#ifndef CARDIAX_AMD_H
#define CARDIAX_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  FK_OK = 0,
  FK_ERR_ARG = -1,         /* invalid argument: shape, pointer, count, range */
  FK_ERR_HIP = -2,         /* a HIP runtime call failed */
  FK_ERR_UNSUPPORTED = -3, /* valid request beyond the library's limits */
  FK_ERR_NONFINITE = -4    /* fk_check_finite found NaN or inf in device data */
};

/* Most stimuli one call accepts (the reference takes any list length). */
#define FK_MAX_STIMULI 32

/* cardiax/params.py:4-18 -- the 14 Fenton-Karma constants in Params order. */
typedef struct fk_params {
  float tau_v_plus, tau_v1_minus, tau_v2_minus, tau_w_plus, tau_w_minus;
  float tau_d, tau_0, tau_r, tau_si, k, V_csi, V_c, V_v, Cm;
} fk_params;

/* cardiax/stimulus.py:13-21 -- Protocol(start, duration, period), in steps,
 * and the (H, W) float32 field. Activity at time t follows
 * cardiax/solve.py:262-267: t >= start && mod(start - t + 1, period) < duration
 * with a floor modulo; it is evaluated on the host in double precision, which
 * equals the reference's float32/int32 evaluation for integral times < 2^24. */
typedef struct fk_stimulus {
  double start, duration, period;
  const float* field;
} fk_stimulus;

/* Thread-local description of the last error; "" when none. */
const char* fk_last_error(void);

/* ABI revision of this library (incremented on incompatible change). */
int32_t fk_abi_version(void);

/* Kernel used for the forward step by fk_step, fk_step_euler and
 * fk_forward_euler. No reference counterpart (the reference leaves kernel
 * choice to XLA). Process-wide. Returns the previous setting, or FK_ERR_ARG
 * for an unknown variant.
 *    0      automatic: the autotuned choice for the grid shape, else 15 up to
 *           1.1 M cells, 19 up to 2.5 M cells and 21 above (grids too small
 *           for those: 15, 13 up to 8 M cells, 2 above)
 *    1      plain, one thread per cell
 *    2, 3   LDS tile of 16 / 32 rows
 *    4-6    streaming wavefront bands of 8 / 16 / 32 rows
 *    7-10   temporal blocking: 64x64 tiles with 2 / 4 steps per launch, 32x32
 *           tiles with 2 / 4 steps per launch
 *    11, 12 streaming bands of 8 / 16 rows with the static diffusion
 *           coefficients precomputed once per fk_forward_euler call
 *    13, 15 streaming bands of 4 / 2 rows
 *    14, 16 the same with precomputed coefficients
 *    17, 18 streaming bands of 2 / 4 rows with two adjacent columns per lane
 *           (for odd W: 15 / 13)
 *    19-22  two Euler steps per launch on streaming bands of 8 / 12 / 16 / 24
 *           rows, bit-identical to two steps of 15 (grids below 2R + 16 rows
 *           or 112 columns run 15; an odd last step runs 15)
 * fk_step and fk_step_euler run 4 in place of 7-11, 5 in place of 12, 13 in
 * place of 14 and 15 in place of 16 and 19-22. */
int32_t fk_set_forward_variant(int32_t variant);

/* Autotuner of the forward step kernel. No reference counterpart (the
 * reference leaves kernel choice to XLA). Times every variant except the
 * plain one over n_steps Euler steps (1..100000) of an H x W grid on `stream`,
 * from a state with an excited region and a wavefront (200 steps after an S1
 * pulse over the left 40% of the columns; the branches of the gates make the
 * cost depend on the state), synchronising the stream; scratch memory is
 * allocated and freed here. It records
 * the fastest for that shape: automatic launches (variant 0) on an H x W grid
 * use it from then on, process-wide. times_ms (may be NULL) receives one time
 * per variant index (-1 for variants not timed); best (may be NULL) the
 * chosen variant. */
int32_t fk_autotune_forward(int32_t H, int32_t W, int32_t n_steps, void* stream,
                            float* times_ms, int32_t* best);

/* The autotuner on the caller's own problem. No reference counterpart. Times
 * every variant but the plain one over n_steps Euler steps (1..100000) from
 * the state (v, w, u) at time t with the given parameters, D-map, stimuli, dt
 * and dx, and records the fastest for the H x W shape as fk_autotune_forward
 * does. The inputs are not modified; scratch memory is allocated and freed
 * here, and `stream` is synchronised. Which side of the Heaviside gates the
 * cells take decides how much work the branching kernels skip, so a state
 * from the run to be made (a spiral, say) picks the kernel for that run. */
int32_t fk_autotune_forward_state(const fk_params* params, const float* D, const float* v,
                                  const float* w, const float* u, int32_t H, int32_t W,
                                  const fk_stimulus* stimuli, int32_t n_stimuli, double t,
                                  int32_t n_steps, float dt, float dx, void* stream,
                                  float* times_ms, int32_t* best);

/* The variant tuned for an H x W grid, or 0 when that shape was not tuned. */
int32_t fk_autotune_lookup(int32_t H, int32_t W);

/* Forgets every tuned shape. */
void fk_autotune_reset(void);

/* The variant fk_forward_euler runs on an H x W grid under the current
 * setting: the fk_set_forward_variant choice when it is not 0, else the tuned
 * one, else the default listed there; 17 / 18 on an odd W report the 15 / 13
 * they run. No reference counterpart. FK_ERR_ARG for H or W below 3. */
int32_t fk_forward_variant_for(int32_t H, int32_t W);

/* cardiax/solve.py:18-23 -- v = w = 1, u = 0 over n_cells cells. */
int32_t fk_init(float* v, float* w, float* u, int64_t n_cells, void* stream);

/* The device-side non-finite check: FK_OK when every value of the n_fields
 * device fields (fields: a host array of device pointers, n_cells floats each)
 * is finite, FK_ERR_NONFINITE when any is NaN or infinite. work: at least 4
 * bytes of device memory for the flag. Enqueued on `stream`, which it then
 * synchronises to read the flag back. No reference counterpart: JAX carries
 * NaN through silently; a caller checks a state after a solve with this. */
int32_t fk_check_finite(const float* const* fields, int32_t n_fields, int64_t n_cells,
                        void* work, void* stream);

/* cardiax/solve.py:26-65 -- the time derivatives (d_v, d_w, d_u) of one
 * state at time t: edge-padded fields, gates and ionic currents, stimulus
 * replacement, 4th-order heterogeneous diffusion, crop. H, W >= 3. */
int32_t fk_step(const fk_params* params, const float* D, const float* v, const float* w,
                const float* u, int32_t H, int32_t W, const fk_stimulus* stimuli,
                int32_t n_stimuli, double t, float dx, float* dv, float* dw, float* du,
                void* stream);

/* cardiax/solve.py:68-70 -- one explicit Euler step, out = in + dt * step(in).
 * Outputs must not alias inputs. */
int32_t fk_step_euler(const fk_params* params, const float* D, const float* v,
                      const float* w, const float* u, int32_t H, int32_t W,
                      const fk_stimulus* stimuli, int32_t n_stimuli, double t, float dt,
                      float dx, float* v_out, float* w_out, float* u_out, void* stream);

/* Scratch bytes fk_forward_euler needs for an H x W grid (a ping-pong state
 * and the precomputed diffusion coefficients). */
size_t fk_forward_workspace_bytes(int32_t H, int32_t W);

/* cardiax/solve.py:92-100 -- the Euler loop over i = t, t+1, ... while
 * i < t_end; step i uses stimulus time i. Reads (v_in, w_in, u_in), writes the
 * final state to (v_out, w_out, u_out). Inputs are not modified; outputs must
 * not alias inputs or the workspace. `work` holds at least
 * fk_forward_workspace_bytes(H, W) bytes. */
int32_t fk_forward_euler(const fk_params* params, const float* D, const float* v_in,
                         const float* w_in, const float* u_in, int32_t H, int32_t W,
                         const fk_stimulus* stimuli, int32_t n_stimuli, double t,
                         double t_end, float dt, float dx, float* v_out, float* w_out,
                         float* u_out, void* work, size_t work_bytes, void* stream);

/* Most cells fk_forward_euler_noalloc accepts: 2^20 (a 1024 x 1024 grid). */
#define FK_NOALLOC_MAX_CELLS (1 << 20)

/* No reference counterpart: fk_forward_euler for callers that cannot allocate
 * device memory (code under stream capture, tight serving loops, harnesses
 * that forbid hipMalloc). Same loop, arguments and results, without the
 * workspace: the ping-pong state lives in a fixed device array of
 * 3 * FK_NOALLOC_MAX_CELLS floats (12 MiB) that each GPU the library runs on
 * holds until the process ends. The call allocates nothing. Grids above
 * FK_NOALLOC_MAX_CELLS cells give FK_ERR_ARG. A variant with precomputed
 * diffusion coefficients (11, 12, 14, 16) runs its streaming counterpart
 * (8, 16, 4, 2 rows), which has no room to keep them. `stream` (NULL = the
 * null stream) must belong to the current device. Every call on one GPU
 * shares that GPU's array: calls that overlap in time on one GPU (other
 * streams, other host threads) race and give wrong results. */
int32_t fk_forward_euler_noalloc(const fk_params* params, const float* D, const float* v_in,
                                 const float* w_in, const float* u_in, int32_t H, int32_t W,
                                 const fk_stimulus* stimuli, int32_t n_stimuli, double t,
                                 double t_end, float dt, float dx, float* v_out,
                                 float* w_out, float* u_out, void* stream);

/* Same loop with the plain one-thread-per-cell kernel (global loads, no LDS,
 * one launch per step): the unoptimised fp32 baseline the tuned kernels are
 * measured against. Same contract as fk_forward_euler. */
int32_t fk_forward_euler_plain(const fk_params* params, const float* D, const float* v_in,
                               const float* w_in, const float* u_in, int32_t H, int32_t W,
                               const fk_stimulus* stimuli, int32_t n_stimuli, double t,
                               double t_end, float dt, float dx, float* v_out,
                               float* w_out, float* u_out, void* work, size_t work_bytes,
                               void* stream);

/* Scratch bytes fk_step_heun and fk_forward_heun need for an H x W grid
 * (the stage-1 slope and a ping-pong state: 6 fields). */
size_t fk_heun_workspace_bytes(int32_t H, int32_t W);

/* cardiax/solve.py:73-85 -- one Heun step at time t: k1 = step(x, t),
 * x1 = x + dt k1, k2 = step(x1, t), out = x + (k1 + k2) dt / 2. Two fused
 * launches: the first writes k1 to the workspace, the second rebuilds x1
 * from x and k1 in registers. Outputs must not alias inputs or the
 * workspace. */
int32_t fk_step_heun(const fk_params* params, const float* D, const float* v, const float* w,
                     const float* u, int32_t H, int32_t W, const fk_stimulus* stimuli,
                     int32_t n_stimuli, double t, float dt, float dx, float* v_out,
                     float* w_out, float* u_out, void* work, size_t work_bytes, void* stream);

/* cardiax/solve.py:103-111 -- Heun steps for i = t, t+1, ... while i < t_end,
 * step i at stimulus time i; contract as fk_forward_euler with the workspace
 * of fk_heun_workspace_bytes. */
int32_t fk_forward_heun(const fk_params* params, const float* D, const float* v_in,
                        const float* w_in, const float* u_in, int32_t H, int32_t W,
                        const fk_stimulus* stimuli, int32_t n_stimuli, double t, double t_end,
                        float dt, float dx, float* v_out, float* w_out, float* u_out,
                        void* work, size_t work_bytes, void* stream);

/* The states of cardiax/solve.py:92-100 after each of n_steps Euler steps
 * from (v0, w0, u0) at time t: state k (k = 0 .. n_steps-1) is the state after
 * step t+k, stored at states + 3*H*W*k as its v, w, u planes. Used to
 * recompute a checkpointed segment for the adjoint. The inputs are not
 * modified and must not alias `states`. */
int32_t fk_forward_trajectory(const fk_params* params, const float* D, const float* v0,
                              const float* w0, const float* u0, int32_t H, int32_t W,
                              const fk_stimulus* stimuli, int32_t n_stimuli, double t,
                              int32_t n_steps, float dt, float dx, float* states,
                              void* stream);

/* cardiax/solve.py:225-254 -- first derivative of a (rows, cols) array along
 * `axis` (0 or 1) without grid spacing; needs >= 5 entries along the axis. */
int32_t fk_gradient(const float* a, float* out, int32_t rows, int32_t cols, int32_t axis,
                    void* stream);

/* cardiax/solve.py:257-271 -- out = X with cells replaced by the combined
 * field of the stimuli active at time t (later stimuli override earlier ones
 * where their field is nonzero). Fields have n_cells entries. */
int32_t fk_stimulate(double t, const float* X, float* out, int64_t n_cells,
                     const fk_stimulus* stimuli, int32_t n_stimuli, void* stream);

/* ---------------------------------------------------------------------------
 * Adjoint of the Euler step, for gradients with respect to the D-map and
 * tau_d, tau_r, tau_si. No reference counterpart as a function: it is what
 * jax.grad through cardiax/solve.py:68-70 computes, with the Heaviside gates
 * (greater_equal, :35-36) as fixed masks and no cotangent reaching j_ion
 * where a stimulus replaces it (jnp.where, :271).
 * ------------------------------------------------------------------------- */

/* Scratch bytes of the adjoint for an H x W grid (tables of the boundary
 * operators, D_x, D_y, the per-block tau partial sums, a cotangent
 * ping-pong buffer and the fields derived per reverse step). */
size_t fk_adjoint_workspace_bytes(int32_t H, int32_t W);

/* Fills the workspace for D and zeroes the tau accumulators; call once per
 * backward pass before the fk_adjoint_step calls. */
int32_t fk_adjoint_prepare(const float* D, int32_t H, int32_t W, void* work,
                           size_t work_bytes, void* stream);

/* One reverse Euler step at time t: given the state (v, w, u) at that step
 * and the cotangent lam of the next state, writes the cotangent of this
 * state to lam_*_out (must not alias the inputs), adds this step's
 * contribution to grad_D (H x W, accumulated in place) and to the workspace's
 * tau accumulators. */
int32_t fk_adjoint_step(const fk_params* params, const float* D, const float* v,
                        const float* w, const float* u, int32_t H, int32_t W,
                        const fk_stimulus* stimuli, int32_t n_stimuli, double t, float dt,
                        float dx, const float* lam_v, const float* lam_w, const float* lam_u,
                        float* lam_v_out, float* lam_w_out, float* lam_u_out, float* grad_D,
                        void* work, size_t work_bytes, void* stream);

/* The reverse pass over n_steps Euler steps starting at time t: states holds
 * the n_steps states x_t .. x_{t+n_steps-1} (layout of fk_forward_trajectory).
 * On entry lam_* is the cotangent of x_{t+n_steps}; on return that of x_t.
 * inject, if not NULL, is a host array of n_steps device pointers: a non-NULL
 * inject[k] (H x W) is added to the u cotangent of x_{t+k+1} before step
 * t+k is reversed (loss terms on intermediate u frames). grad_D and the tau
 * accumulators gather every step's contribution, as fk_adjoint_step does. */
int32_t fk_adjoint_segment(const fk_params* params, const float* D, int32_t H, int32_t W,
                           const fk_stimulus* stimuli, int32_t n_stimuli, double t,
                           int32_t n_steps, float dt, float dx, const float* states,
                           const float* const* inject, float* lam_v, float* lam_w,
                           float* lam_u, float* grad_D, void* work, size_t work_bytes,
                           void* stream);

/* Writes the accumulated gradients with respect to (tau_d, tau_r, tau_si) to
 * g_tau (3 floats, device memory). Uses a scratch area of the workspace for
 * block sums; the accumulators themselves are left as they are. */
int32_t fk_adjoint_tau_grads(void* work, int32_t H, int32_t W, float* g_tau, void* stream);

/* ---------------------------------------------------------------------------
 * Batched forward Euler. No reference counterpart: the reference runs one
 * simulation per call (experiments/generate_fd_data_128.py:25-56 loops over
 * seeds). Each simulation follows cardiax/solve.py:92-100 exactly as
 * fk_forward_euler does.
 * ------------------------------------------------------------------------- */

/* Most stimuli one simulation of a batch may carry. */
#define FK_BATCH_MAX_STIMULI 4

/* Work-list options of a batch: bands of 2, 4, 8 or 16 rows advanced one step
 * per launch, and bands of 8 or 16 rows advanced two steps per launch (those
 * two only when every simulation has >= 2R + 16 rows and 56..96 or >= 104
 * columns). A band height or'ed with FK_BATCH_TWO_STEP names a two-step
 * option. */
#define FK_BATCH_N_OPTIONS 6
#define FK_BATCH_TWO_STEP 256

/* One simulation of a batch: its own constants, D-map, state (updated in
 * place by fk_batch_run), stimuli, time of the next step and steps. */
typedef struct fk_sim {
  fk_params params;
  const float* D;
  float* v;
  float* w;
  float* u;
  int32_t H, W;
  int32_t n_stimuli;
  fk_stimulus stimuli[FK_BATCH_MAX_STIMULI];
  double t;
  float dt, dx;
} fk_sim;

typedef struct fk_batch fk_batch; /* opaque */

/* Builds the tile work list of all simulations and allocates the device
 * descriptors and ping-pong state pools (the only allocating entry point;
 * fk_batch_destroy frees them). The fields and state buffers the descriptors
 * point to must outlive the batch. */
int32_t fk_batch_create(const fk_sim* sims, int32_t n, fk_batch** out);

/* Advances every simulation by n_steps Euler steps; results land in each
 * fk_sim's v, w, u and each simulation's time advances by n_steps. Each
 * launch advances the whole batch one step (or two, with a two-step option);
 * the stimulus activity of up to 32 steps comes from one small kernel.
 * use_graph != 0 replays a recorded hipGraph of 32 steps (one graph launch
 * per 32 steps for the whole batch), otherwise direct launches. */
int32_t fk_batch_run(fk_batch* b, int64_t n_steps, int32_t use_graph, void* stream);

/* Autotuner of the batch: times n_steps (1..10000) batched steps with each
 * available work-list option on `stream` (synchronising it) and keeps the
 * fastest for later runs (8 rows, one step per launch, until tuned).
 * Overwrites only the batch's internal ping-pong pools, never a simulation's
 * state. times_ms (FK_BATCH_N_OPTIONS floats in the order 2, 4, 8, 16 rows,
 * then 8, 16 rows with two steps; -1 for an unavailable option; may be NULL)
 * and best_rows (may be NULL) receive the timings and the chosen band height
 * (or'ed with FK_BATCH_TWO_STEP for a two-step option). */
int32_t fk_batch_autotune(fk_batch* b, int32_t n_steps, void* stream, float* times_ms,
                          int32_t* best_rows);

/* Pins the work-list option of later runs, as the autotuner would: a band
 * height of 2, 4, 8 or 16 rows, or 8 or 16 or'ed with FK_BATCH_TWO_STEP
 * (FK_ERR_UNSUPPORTED when some simulation does not fit two steps per
 * launch); waits for the device first. */
int32_t fk_batch_set_rows(fk_batch* b, int32_t rows);

/* Number of (simulation, strip, band) tiles of one batch step. */
int32_t fk_batch_n_items(const fk_batch* b);

/* Waits for the device and releases the batch. */
int32_t fk_batch_destroy(fk_batch* b);

/* ---------------------------------------------------------------------------
 * Batched rollout with its adjoint: many frame-loss problems (the forward
 * solve of cardiax/solve.py:92-100 sampled every frame_every steps, and the
 * reverse pass of fk_adjoint_segment) advanced together, one launch per
 * Euler step and one per reverse step for all of them. No reference
 * counterpart: it batches what the per-grid calls fk_forward_trajectory,
 * fk_adjoint_prepare, fk_adjoint_segment and fk_adjoint_tau_grads do for one
 * problem, and its results are bitwise those of that per-grid sequence with
 * the same checkpoint interval.
 * ------------------------------------------------------------------------- */

typedef struct fk_rollout_spec {
  fk_params params;
  const float* D;   /* H x W, device; read by every forward and backward call */
  const float* v0;  /* initial state, device */
  const float* w0;
  const float* u0;
  int32_t H, W;
  int32_t n_stimuli;
  fk_stimulus stimuli[FK_BATCH_MAX_STIMULI];
  double t0;        /* time index of the first step */
  float dt, dx;
  float* frames;         /* n_steps / frame_every planes of H x W (forward output) */
  const float* g_frames; /* the same shape: dLoss/dframes (backward input) */
  float* grad_D;         /* H x W: dLoss/dD (backward output) */
} fk_rollout_spec;

typedef struct fk_rollout_batch fk_rollout_batch; /* opaque */

/* Plans n problems of n_steps Euler steps each, frames every frame_every
 * steps (frame j-1 = u after j * frame_every steps), checkpoints every
 * checkpoint_every steps. Allocates (once) the checkpoints, one segment of
 * states per problem, the cotangents and the adjoint workspaces; the
 * caller's buffers named in the specs must stay valid until destroy.
 * Synchronous. */
int32_t fk_rollout_batch_create(const fk_rollout_spec* specs, int32_t n, int32_t n_steps,
                                int32_t frame_every, int32_t checkpoint_every,
                                fk_rollout_batch** out);
/* Device memory the plan holds, bytes. */
size_t fk_rollout_batch_device_bytes(const fk_rollout_batch* b);
/* The forward rollout of every problem from its D and initial state: writes
 * the frames and keeps the checkpoints. Stream-ordered. */
int32_t fk_rollout_batch_forward(fk_rollout_batch* b, void* stream);
/* The reverse pass from g_frames after a forward call with the same D:
 * writes grad_D and, when g_tau is not null, the gradients with respect to
 * (tau_d, tau_r, tau_si) of every problem (3 n floats, device). Stream-
 * ordered. */
int32_t fk_rollout_batch_backward(fk_rollout_batch* b, float* g_tau, void* stream);
/* Waits for the device and frees the plan. */
int32_t fk_rollout_batch_destroy(fk_rollout_batch* b);

/* ---------------------------------------------------------------------------
 * Checkpoint output of the dataset generator (deepx/generate.py:135-210).
 * ------------------------------------------------------------------------- */

/* Workspace of fk_resize: n_images * h * W doubles (the row pass's result). */
size_t fk_resize_workspace_bytes(int32_t n_images, int32_t H, int32_t W, int32_t h, int32_t w);

/* cardiax/io.py:118-124 -- imresize(a, (h, w)) = jax.image.resize(a,
 * a.shape[:-2] + (h, w), "bilinear"), which is antialiased by default: n_images
 * contiguous H x W float32 images in `src` resampled to h x w in `dst`, along
 * each axis with a triangle kernel widened by the downsampling factor and
 * weights normalised to 1 (an unchanged axis is copied exactly). Sizes >= 1;
 * dst must not overlap src; work holds fk_resize_workspace_bytes. Stream-
 * ordered. */
int32_t fk_resize(const float* src, int32_t n_images, int32_t H, int32_t W, float* dst,
                  int32_t h, int32_t w, void* work, size_t work_bytes, void* stream);

/* ---------------------------------------------------------------------------
 * Stimulus fields built on the device (cardiax/stimulus.py:31-140).
 * ------------------------------------------------------------------------- */

/* cardiax/stimulus.py:31-60 (rectangular) and :63-106 (linear) -- field =
 * modulus on rows [r0, r1) x columns [c0, c1), 0 elsewhere. The bounds are
 * the reference's slices already normalised by Python slice semantics:
 * 0 <= r0, r1 <= H, 0 <= c0, c1 <= W (an empty range gives an all-zero
 * field). */
int32_t fk_stimulus_box(float* field, int32_t H, int32_t W, int32_t r0, int32_t r1, int32_t c0,
                        int32_t c1, float modulus, void* stream);

/* cardiax/stimulus.py:137-139 (triangular) -- the resampling that
 * scipy.ndimage.rotate(src, angle, mode="nearest", prefilter=False,
 * reshape=False) performs: scipy.ndimage.affine_transform with order 3, no
 * prefilter and mode "nearest". Output cell o = (row, col) of the H x W dst
 * samples the cubic B-spline of the unfiltered src at input coordinate
 * matrix . o + offset (matrix: 4 doubles, row-major; offset: 2 doubles; both
 * host memory read before the call returns); taps outside the grid take the
 * nearest edge value. The caller builds the map as scipy's rotate does:
 * matrix = [[cos a, sin a], [-sin a, cos a]] (scipy.special.cosdg/sindg) and
 * offset = c - matrix @ c, c = ((H-1)/2, (W-1)/2). dst must not overlap src.
 * Stream-ordered. */
int32_t fk_affine_transform(const float* src, int32_t H, int32_t W, const double* matrix,
                            const double* offset, float* dst, void* stream);

/* ---------------------------------------------------------------------------
 * The random scar D-map of the dataset generator (deepx/generate.py:79-83,
 * deepx/utils_scars.py:246-426): its rasterisation and its blur.
 * ------------------------------------------------------------------------- */

/* Device scratch of fk_rasterize_polygons for n_polys polygons with
 * n_vertices vertices in all. */
size_t fk_rasterize_workspace_bytes(int32_t n_polys, int32_t n_vertices);

/* deepx/utils_scars.py:79-97,377-381 (drawpolyintoemptycanvas via
 * skimage.draw.polygon, the blobs summed and clipped to 1): mask (H x W,
 * device) = 1 at every pixel (r, c) inside one of the n_polys closed
 * polygons, 0 elsewhere. Polygon p has the vertices (rows[k], cols[k]) for
 * k in [starts[p], starts[p+1]) (host arrays, >= 3 vertices each, read
 * before the call returns). A pixel is inside when it lies in the polygon's
 * box [int(max(0, min row)), ceil(max row)] x [int(max(0, min col)),
 * ceil(max col)] and the crossing-number test at its centre holds: edge
 * (i, j) counts when (row_i <= r < row_j or row_j <= r < row_i) and
 * c < (col_j - col_i)(r - row_i) / (row_j - row_i) + col_i, in double.
 * Pixels outside the image are dropped. work: fk_rasterize_workspace_bytes
 * of device memory. Waits for `stream` before returning. */
int32_t fk_rasterize_polygons(float* mask, int32_t H, int32_t W, const double* rows,
                              const double* cols, const int32_t* starts, int32_t n_polys,
                              void* work, size_t work_bytes, void* stream);

/* deepx/utils_scars.py:402-406 (blur_scar: jax.scipy.signal.convolve2d(scar,
 * kernel, "same")): dst = the "same"-size 2-D convolution of the H x W src
 * with the kh x kw kernel (device, double, row-major):
 * dst[i, j] = sum_{p, q} kernel[p, q] src[i + (kh-1)/2 - p, j + (kw-1)/2 - q]
 * with integer halves and zero outside the image; products and sums in
 * double, rounded once to float. dst must not be src. work:
 * fk_convolve_workspace_bytes(H, W, kh, kw) of device memory. Stream-ordered. */
int32_t fk_convolve_same(const float* src, int32_t H, int32_t W, const double* kernel,
                         int32_t kh, int32_t kw, float* dst, void* work, size_t work_bytes,
                         void* stream);

/* Device scratch of fk_convolve_same for an H x W image and a kh x kw kernel
 * (the kernel's zero-padded, transposed copy and a zero-padded copy of the
 * image). */
size_t fk_convolve_workspace_bytes(int32_t H, int32_t W, int32_t kh, int32_t kw);

/* deepx/utils_scars.py:426 with deepx/generate.py:79-83 and
 * cardiax/convert.py:23-26: D = diffusivity_rescale(1 - scar, (lo, hi)), i.e.
 * c = 1 - scar, D = (c - min c)(hi - lo) / (max c - min c) + lo, in double,
 * rounded once to float (n_cells cells, device). work: 4096 bytes of device
 * memory (the min and max partials). A scar of one value gives NaN, as the
 * reference's division by zero does. Stream-ordered. */
int32_t fk_scar_to_diffusivity(const float* scar, int64_t n_cells, double lo, double hi,
                               float* D, void* work, size_t work_bytes, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* CARDIAX_AMD_H */
