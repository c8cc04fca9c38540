// Host-side helpers shared by the C-ABI translation units: argument checks,
// the thread-local error message, parameter preprocessing and the stimulus
// schedule (the activity law of cardiax/solve.py:262-267, evaluated per step
// on the host so kernels only see the fields that are active).
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <initializer_list>
#include <string>

#include "cardiax_amd.h"
#include "fk_device.h"

namespace fk {

void set_error_text(int32_t code, const char* msg);
void clear_error();
// Code of the last set_error on this thread (FK_OK after clear_error).
int32_t last_code();

inline int32_t set_error(int32_t code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  set_error_text(code, buf);
  return code;
}

// Clears the thread-local error at the start of every entry point.
struct ErrorScope {
  ErrorScope() { clear_error(); }
};

#define HIP_CHECK(expr)                                                                 \
  do {                                                                                  \
    hipError_t e_ = (expr);                                                             \
    if (e_ != hipSuccess)                                                               \
      return ::fk::set_error(FK_ERR_HIP, "%s failed: %s", #expr, hipGetErrorString(e_)); \
  } while (0)

constexpr int kMaxSide = 1 << 18;

inline bool check_grid(int32_t H, int32_t W) {
  if (H < 3 || W < 3) {
    set_error(FK_ERR_ARG, "grid %dx%d too small: both sides must be >= 3", H, W);
    return false;
  }
  if (H > kMaxSide || W > kMaxSide) {
    set_error(FK_ERR_ARG, "grid %dx%d too large: sides must be <= %d", H, W, kMaxSide);
    return false;
  }
  return true;
}

inline bool check_ptrs(std::initializer_list<const void*> ps) {
  int k = 0;
  for (const void* p : ps) {
    if (p == nullptr) {
      set_error(FK_ERR_ARG, "null array pointer (argument %d of the pointer list)", k);
      return false;
    }
    ++k;
  }
  return true;
}

inline bool check_scalar(const char* name, double x) {
  if (!std::isfinite(x)) {
    set_error(FK_ERR_ARG, "%s must be finite", name);
    return false;
  }
  return true;
}

inline bool check_dx(float dx) {
  if (!std::isfinite(dx) || dx == 0.0f) {
    set_error(FK_ERR_ARG, "dx must be finite and nonzero");
    return false;
  }
  return true;
}

inline bool check_params(const fk_params* p) {
  if (p == nullptr) {
    set_error(FK_ERR_ARG, "params is null");
    return false;
  }
  const float* f = &p->tau_v_plus;
  for (int i = 0; i < 14; ++i)
    if (!std::isfinite(f[i])) {
      set_error(FK_ERR_ARG, "params field %d is not finite", i);
      return false;
    }
  return true;
}

inline bool check_stimuli(const fk_stimulus* s, int32_t n) {
  if (n < 0 || (n > 0 && s == nullptr)) {
    set_error(FK_ERR_ARG, "bad stimulus list (n=%d)", n);
    return false;
  }
  if (n > FK_MAX_STIMULI) {
    set_error(FK_ERR_UNSUPPORTED, "%d stimuli exceed the limit of %d", n, FK_MAX_STIMULI);
    return false;
  }
  for (int i = 0; i < n; ++i) {
    if (s[i].field == nullptr) {
      set_error(FK_ERR_ARG, "stimulus %d has a null field", i);
      return false;
    }
    if (std::isnan(s[i].start) || std::isnan(s[i].duration) || std::isnan(s[i].period)) {
      set_error(FK_ERR_ARG, "stimulus %d has a NaN protocol entry", i);
      return false;
    }
  }
  return true;
}

// Number of iterations of the reference loop i = t, t+1, ... while i < t_end
// (cardiax/solve.py:94-99).
inline int64_t loop_count(double t, double t_end) {
  const double n = std::ceil(t_end - t);
  return n > 0 ? (int64_t)n : 0;
}

// Bound on time indices (step units): 2^24.
constexpr double kMaxTime = 16777216.0;

inline bool check_time(double t, double t_end) {
  if (!std::isfinite(t) || !std::isfinite(t_end)) {
    set_error(FK_ERR_ARG, "t and t_end must be finite");
    return false;
  }
  // The reference's step counter is float32 when reached through forward()
  // and int32 through the dataset generator; the two agree on every step
  // index below 2^24 (a loop may end at 2^24).
  if (std::fabs(t) >= kMaxTime || std::fabs(t_end) > kMaxTime) {
    set_error(FK_ERR_ARG, "time indices must stay below 2^24 (t %.17g, t_end %.17g)", t,
              t_end);
    return false;
  }
  return true;
}

// Floor modulo with the sign of the divisor, as jnp.mod.
inline double floor_mod(double a, double b) {
  double r = std::fmod(a, b);
  if (r != 0.0 && ((r < 0.0) != (b < 0.0))) r += b;
  return r;
}

// cardiax/solve.py:262-267.
inline bool stimulus_active(const fk_stimulus& s, double t) {
  return (t >= s.start) && (floor_mod(s.start - t + 1.0, s.period) < s.duration);
}

// Per-step list of active stimulus fields (order preserved), recomputed only
// when the time changes.
class StimSchedule {
 public:
  StimSchedule(const fk_stimulus* s, int32_t n) : s_(s), n_(n) { args_.n = 0; }
  const StimArgs& at(double t) {
    int k = 0;
    for (int i = 0; i < n_; ++i)
      if (stimulus_active(s_[i], t)) args_.f[k++] = s_[i].field;
    args_.n = k;
    return args_;
  }
  // Bit i set = stimulus i active at time t.
  uint32_t mask(double t) const {
    uint32_t m = 0;
    for (int i = 0; i < n_; ++i)
      if (stimulus_active(s_[i], t)) m |= 1u << i;
    return m;
  }

 private:
  const fk_stimulus* s_;
  int32_t n_;
  StimArgs args_;
};

// fk_forward.hip: the Euler loop behind fk_forward_euler (cardiax/solve.py:92-100)
// with the kernel `variant` (0 = automatic). The workspace holds a ping-pong
// state and, unless `state_only`, the precomputed diffusion coefficients;
// with `state_only` a precomputed-coefficient variant runs its streaming
// counterpart and the workspace needs only 3 * H * W floats.
int32_t forward_loop(int variant, const fk_params* params, const float* D, const float* v_in,
                     const float* w_in, const float* u_in, int32_t H, int32_t W,
                     const fk_stimulus* stimuli, int32_t n_stimuli, double t, double t_end,
                     float dt, float dx, float* v_out, float* w_out, float* u_out, void* work,
                     size_t work_bytes, void* stream, bool state_only);

// The process-wide fk_set_forward_variant setting.
int current_forward_variant();

inline Coef make_coef(const fk_params& p) {
  const double log2e = 1.4426950408889634;
  Coef c;
  c.V_c = p.V_c;
  c.V_v = p.V_v;
  c.V_csi = p.V_csi;
  c.kk = (float)(-2.0 * (double)p.k * log2e);
  c.kk_off = (float)(2.0 * (double)p.k * log2e * (double)p.V_csi);
  c.inv_tvp = (float)(1.0 / (double)p.tau_v_plus);
  c.inv_tv1m = (float)(1.0 / (double)p.tau_v1_minus);
  c.inv_tv2m = (float)(1.0 / (double)p.tau_v2_minus);
  c.inv_twp = (float)(1.0 / (double)p.tau_w_plus);
  c.inv_twm = (float)(1.0 / (double)p.tau_w_minus);
  c.inv_td = (float)(1.0 / (double)p.tau_d);
  c.inv_t0 = (float)(1.0 / (double)p.tau_0);
  c.inv_tr = (float)(1.0 / (double)p.tau_r);
  c.inv_tsi = (float)(1.0 / (double)p.tau_si);
  c.inv_cm = (float)(1.0 / (double)p.Cm);
  return c;
}

}  // namespace fk
