/*
 * C restatement of the cardiax forward-Euler path (oracle, CPU only).
 *
 * TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py, always to check results or as the CPU reference
 * timing, never by the product path.
 *
 * It keeps the reference's operation graph (cardiax/solve.py:26-70): edge-pad
 * v, w, u, D by one cell (:29-32), evaluate gates and currents on the padded
 * grid (:35-42), replace j_ion by the active stimulus fields (:45-46,
 * :257-271), take six `gradient`s on padded arrays (:49-54, :225-254), combine
 * (:55-59), crop (:61-65) and add dt times the derivatives (:68-70), for
 * i = t, t+1, ... while i < t_end (:92-100).
 *
 * Built twice from this file: REAL=double (fk_oracle_*_f64, the FP64 oracle
 * for long horizons) and REAL=float (fk_oracle_*_f32, the fp32 CPU baseline).
 * OpenMP threads over rows; the thread count is the caller's choice.
 */
#include <math.h>
#include <stddef.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

#ifndef REAL
#define REAL double
#define SUFFIX f64
#define TANH tanh
#else
#define TANH tanhf
#endif

#define CAT2(a, b) a##_##b
#define CAT(a, b) CAT2(a, b)
#define FN(name) CAT(name, SUFFIX)

/* Python floor modulo (sign of the divisor), as jnp.mod. */
static double FN(floor_mod)(double a, double b) {
  double r = fmod(a, b);
  if (r != 0.0 && ((r < 0.0) != (b < 0.0))) r += b;
  return r;
}

/* cardiax/solve.py:262-267 */
static int FN(active)(double t, double start, double duration, double period) {
  return (t >= start) && (FN(floor_mod)(start - t + 1.0, period) < duration);
}

/* jnp.pad(a, 1, mode="edge") into a (H+2) x (W+2) buffer. */
static void FN(pad_edge)(const REAL* a, REAL* p, int H, int W) {
  const int PW = W + 2;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < H + 2; ++i) {
    int si = i - 1;
    si = si < 0 ? 0 : (si > H - 1 ? H - 1 : si);
    for (int j = 0; j < W + 2; ++j) {
      int sj = j - 1;
      sj = sj < 0 ? 0 : (sj > W - 1 ? W - 1 : sj);
      p[(size_t)i * PW + j] = a[(size_t)si * W + sj];
    }
  }
}

/* cardiax/solve.py:225-254 on an R x C array along `axis`, scaled by s. */
static void FN(gradient)(const REAL* a, REAL* g, int R, int C, int axis, REAL s) {
  const REAL e0 = (REAL)(-11.0 / 6.0), e1 = (REAL)3.0, e2 = (REAL)(3.0 / 2.0),
             e3 = (REAL)(1.0 / 3.0), c1 = (REAL)(1.0 / 12.0), c2 = (REAL)(2.0 / 3.0);
  const int n = axis == 0 ? R : C;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < R; ++i) {
    for (int j = 0; j < C; ++j) {
      const int p = axis == 0 ? i : j;
      const size_t st = axis == 0 ? (size_t)C : 1;
      const REAL* x = a + (size_t)i * C + j; /* x[k*st] = a at position p + k */
      REAL r;
      if (p <= 1)
        r = e0 * x[0] + e1 * x[st] - e2 * x[2 * st] + e3 * x[3 * st];
      else if (p >= n - 2)
        r = -e3 * x[-3 * (ptrdiff_t)st] + e2 * x[-2 * (ptrdiff_t)st] -
            e1 * x[-(ptrdiff_t)st] + (REAL)(11.0 / 6.0) * x[0];
      else
        r = c1 * x[-2 * (ptrdiff_t)st] - c2 * x[-(ptrdiff_t)st] + c2 * x[st] -
            c1 * x[2 * st];
      g[(size_t)i * C + j] = r / s;
    }
  }
}

/* Bytes of scratch FN(forward) needs. */
size_t FN(fk_oracle_scratch_bytes)(int H, int W) {
  return (size_t)12 * (size_t)(H + 2) * (size_t)(W + 2) * sizeof(REAL);
}

/*
 * In-place Euler loop on (v, w, u). params: the 14 Params fields in order.
 * fields: n_stim packed (H, W) arrays; proto: n_stim x (start, duration,
 * period). scratch: FN(fk_oracle_scratch_bytes)(H, W) bytes. nthreads <= 0
 * keeps the OpenMP default. Returns the number of steps taken.
 *
 * gates / gate_mode: the Heaviside gates p = [u >= V_c], q = [u >= V_v] of
 * every cell and step (one byte per cell-step, bit 0 = p, bit 1 = q, step s
 * at gates + s * H * W). gate_mode 0 ignores `gates`; 1 records them; 2
 * replays them in place of the comparisons. Replay makes the loop the smooth
 * function of D, the time constants and the state that the adjoint
 * differentiates (jax.grad treats greater_equal as constant), so central
 * differences of it converge where those of the true loop jump wherever a
 * gate flips.
 */
int64_t FN(fk_oracle_forward)(const double* params, const REAL* D, REAL* v, REAL* w, REAL* u,
                              int H, int W, int n_stim, const double* proto,
                              const REAL* fields, double t, double t_end, double dt,
                              double dx, void* scratch, int nthreads, uint8_t* gates,
                              int gate_mode) {
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#else
  (void)nthreads;
#endif
  const REAL tvp = (REAL)params[0], tv1m = (REAL)params[1], tv2m = (REAL)params[2],
             twp = (REAL)params[3], twm = (REAL)params[4], td = (REAL)params[5],
             t0 = (REAL)params[6], tr = (REAL)params[7], tsi = (REAL)params[8],
             k = (REAL)params[9], Vcsi = (REAL)params[10], Vc = (REAL)params[11],
             Vv = (REAL)params[12], Cm = (REAL)params[13];
  const int PH = H + 2, PW = W + 2;
  const size_t PN = (size_t)PH * PW;
  REAL* buf = (REAL*)scratch;
  REAL *pv = buf, *pw = buf + PN, *pu = buf + 2 * PN, *pD = buf + 3 * PN;
  REAL *ux = buf + 4 * PN, *uy = buf + 5 * PN, *uxx = buf + 6 * PN, *uyy = buf + 7 * PN;
  REAL *Dx = buf + 8 * PN, *Dy = buf + 9 * PN, *pf = buf + 10 * PN, *st = buf + 11 * PN;
  const REAL sdx = (REAL)dx, sdt = (REAL)dt;

  FN(pad_edge)(D, pD, H, W);
  FN(gradient)(pD, Dx, PH, PW, 0, sdx);
  FN(gradient)(pD, Dy, PH, PW, 1, sdx);

  const double nsteps_d = ceil(t_end - t);
  const int64_t nsteps = nsteps_d > 0 ? (int64_t)nsteps_d : 0;
  for (int64_t s = 0; s < nsteps; ++s) {
    const double ti = t + (double)s;
    FN(pad_edge)(v, pv, H, W);
    FN(pad_edge)(w, pw, H, W);
    FN(pad_edge)(u, pu, H, W);
    FN(gradient)(pu, ux, PH, PW, 0, sdx);
    FN(gradient)(pu, uy, PH, PW, 1, sdx);
    FN(gradient)(ux, uxx, PH, PW, 0, sdx);
    FN(gradient)(uy, uyy, PH, PW, 1, sdx);

    /* combined stimulus on the padded grid (cardiax/solve.py:259-269) */
    memset(st, 0, PN * sizeof(REAL));
    for (int m = 0; m < n_stim; ++m) {
      if (!FN(active)(ti, proto[3 * m], proto[3 * m + 1], proto[3 * m + 2])) continue;
      FN(pad_edge)(fields + (size_t)m * H * W, pf, H, W);
#pragma omp parallel for schedule(static)
      for (size_t q = 0; q < PN; ++q)
        if (pf[q] != 0) st[q] = pf[q];
    }

    uint8_t* gs = gates ? gates + (size_t)s * H * W : NULL;
#pragma omp parallel for schedule(static)
    for (int i = 1; i < PH - 1; ++i) {
      for (int j = 1; j < PW - 1; ++j) {
        const size_t q = (size_t)i * PW + j;
        const size_t o = (size_t)(i - 1) * W + (j - 1);
        const REAL vv = pv[q], ww = pw[q], uu = pu[q];
        REAL p = uu >= Vc ? (REAL)1 : (REAL)0;
        REAL qq = uu >= Vv ? (REAL)1 : (REAL)0;
        if (gate_mode == 1 && gs) gs[o] = (uint8_t)((p != 0) | ((qq != 0) << 1));
        if (gate_mode == 2 && gs) {
          p = (gs[o] & 1) ? (REAL)1 : (REAL)0;
          qq = (gs[o] & 2) ? (REAL)1 : (REAL)0;
        }
        const REAL tvm = ((REAL)1 - qq) * tv1m + qq * tv2m;
        const REAL jfi = -vv * p * (uu - Vc) * ((REAL)1 - uu) / td;
        const REAL jso = (uu * ((REAL)1 - p) / t0) + (p / tr);
        const REAL jsi = -(ww * ((REAL)1 + TANH(k * (uu - Vcsi)))) / ((REAL)2 * tsi);
        REAL jion = -(jfi + jso + jsi) / Cm;
        if (st[q] != 0) jion = st[q];
        const REAL del_u = pD[q] * (uxx[q] + uyy[q]) + Dx[q] * ux[q] + Dy[q] * uy[q];
        const REAL dv = (((REAL)1 - p) * ((REAL)1 - vv) / tvm) - ((p * vv) / tvp);
        const REAL dw = (((REAL)1 - p) * ((REAL)1 - ww) / twm) - ((p * ww) / twp);
        const REAL du = del_u + jion;
        v[o] = vv + dv * sdt;
        w[o] = ww + dw * sdt;
        u[o] = uu + du * sdt;
      }
    }
  }
  return nsteps;
}
