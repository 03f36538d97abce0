/*
 * tdse_oracle.c -- CPU restatement of MP-CITDSE's `tdse` hot path.
 *
 * TEST INFRASTRUCTURE ONLY (the oracle, never the product): loaded by
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg.
 * See tdse_oracle.h for the reference file:line each function follows.
 */
#include "tdse_oracle.h"

#include <dlfcn.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* ------------------------------------------------------------------ */
/* dgemv with reference-BLAS semantics (column-major).                  */
/* ------------------------------------------------------------------ */
enum { COLMAJOR = 102, NOTRANS = 111, TRANS = 112 };

static void own_dgemv(int layout, int trans, int M, int N, double alpha,
                      const double *A, int lda, const double *x, int incx,
                      double beta, double *y, int incy) {
  (void)layout; /* always column-major at the reference's call sites */
  int leny = (trans == NOTRANS) ? M : N;
  if (beta != 1.0) {
    for (int i = 0; i < leny; ++i)
      y[i * incy] = (beta == 0.0) ? 0.0 : beta * y[i * incy];
  }
  if (trans == NOTRANS) {
    /* y += alpha * A * x, A is M x N column-major: axpy per column. */
    for (int j = 0; j < N; ++j) {
      double temp = alpha * x[j * incx];
      const double *col = A + (size_t)j * lda;
      for (int i = 0; i < M; ++i) y[i * incy] += temp * col[i];
    }
  } else {
    /* y += alpha * A^T * x: dot product per column. */
    for (int j = 0; j < N; ++j) {
      const double *col = A + (size_t)j * lda;
      double temp = 0.0;
      for (int i = 0; i < M; ++i) temp += col[i] * x[i * incx];
      y[j * incy] += alpha * temp;
    }
  }
}

static or_dgemv_fn g_dgemv = own_dgemv;

void or_set_dgemv(or_dgemv_fn fn) { g_dgemv = fn ? fn : own_dgemv; }

int or_use_blas(const char *path) {
  void *h = dlopen(path, RTLD_NOW | RTLD_GLOBAL);
  if (!h) return -1;
  void *f = dlsym(h, "cblas_dgemv");
  if (!f) return -2;
  g_dgemv = (or_dgemv_fn)f;
  return 0;
}

/* ------------------------------------------------------------------ */
/* Pulses: src/tdse/pulse.cpp:3-97, au.hpp:30-32                         */
/* ------------------------------------------------------------------ */
void or_to_au(double Io_W, double w_eV, double *Io_au, double *w_au) {
  /* pulse.cpp:3-6 with conv::I_W_cm2_au_ and conv::En_ev_au_ */
  *Io_au = Io_W * ((1. / 3.509338) * 1.e-16);
  *w_au = w_eV * (1. / 27.211396181);
}

double or_period(double w, int cycles) { return cycles * 2.0 * M_PI / w; }

static void zero_pulse(or_pulse *p) { memset(p, 0, sizeof(*p)); }

void or_sine_a_setup(double Io, double w, double cepd, int cycles,
                     or_pulse *p) { /* pulse.cpp:41-47 */
  zero_pulse(p);
  p->kind = OR_SINE2;
  p->Eo = sqrt(Io) / w;
  p->w = w;
  p->cepd = cepd * (2.0 * M_PI / w);
  p->a = w / (2.0 * cycles);
}

void or_sine_e_setup(double Io, double w, double cepd, int cycles,
                     or_pulse *p) { /* pulse.cpp:69-75 */
  zero_pulse(p);
  p->kind = OR_SINE2;
  p->Eo = sqrt(Io);
  p->w = w;
  p->cepd = cepd * (2 * M_PI / w);
  p->a = w / (2.0 * cycles);
}

static void gauss_common(double Eo, double w, double tau, int cycles,
                         double cepd, or_pulse *p) { /* pulse.cpp:10-32 */
  const double width = 0.58870501125773734549;
  zero_pulse(p);
  p->kind = OR_GAUSS;
  p->Eo = Eo;
  p->w = w;
  double To = 2 * M_PI / w;
  int cycles_fwhm = 2 + (int)(cycles / 10);
  p->cepd = cepd;
  p->a = 0.5 * To * cycles_fwhm / width;
  p->b = tau;
}

void or_gauss_a_setup(double Io, double w, double tau, int cycles, double cepd,
                      or_pulse *p) {
  gauss_common(sqrt(Io) / w, w, tau, cycles, cepd, p);
}

void or_gauss_e_setup(double Io, double w, double tau, int cycles, double cepd,
                      or_pulse *p) {
  gauss_common(sqrt(Io), w, tau, cycles, cepd, p);
}

double or_field(const or_pulse *p, double t) {
  switch (p->kind) {
  case OR_SINE2: { /* sineAA / sineEE, pulse.cpp:52-56, 93-97 */
    double sinWt = sin(p->a * t);
    return (double)(t >= 0.0 && t <= M_PI / p->a) * p->Eo * sinWt * sinWt *
           sin(p->w * t + p->cepd);
  }
  case OR_GAUSS: { /* gauss, pulse.cpp:34-39 */
    double t_diff = t - 0.5 * p->b;
    return (double)(t >= 0.0 && t <= p->b) * p->Eo *
           exp(-t_diff * t_diff / (p->a * p->a)) * sin(p->w * t + p->cepd);
  }
  case OR_PWL: { /* extension: piecewise-linear envelope */
    if (!(t >= 0.0 && t <= p->t_end) || p->n_env < 2) return 0.0;
    double dx = p->t_end / (p->n_env - 1);
    double u = t / dx;
    int k = (int)floor(u);
    if (k > p->n_env - 2) k = p->n_env - 2;
    if (k < 0) k = 0;
    double f = u - k;
    double env = (1.0 - f) * p->env[k] + f * p->env[k + 1];
    return p->scale * env * sin(p->w * t + p->cepd);
  }
  }
  return 0.0;
}

/* ------------------------------------------------------------------ */
/* RHS: src/tdse/tdse.cpp:14-73 (velocity) and 85-143 (length)           */
/* ------------------------------------------------------------------ */
int or_state_size(const or_sys *sys) {
  int n = 0;
  for (int L = 0; L < sys->nb; ++L) n += sys->s[L];
  return n;
}

/* The diagonal term -i E x added to y (y = -iEx + bt2*y, tdse.cpp:44-48). */
static void diag_add(const double *E, int n, const double *x, double *y,
                     int overwrite) {
  for (int i = 0; i < n; ++i) {
    int i2 = 2 * i;
    double im = -E[i] * x[i2];
    double re = E[i] * x[i2 + 1];
    if (overwrite) {
      y[i2 + 1] = im;
      y[i2] = re;
    } else {
      y[i2 + 1] = im + 1.0 * y[i2 + 1];
      y[i2] = re + 1.0 * y[i2];
    }
  }
}

static void rhs_impl(const or_sys *sys, double F, const double *x,
                     double *dxdt, int with_diag) {
  const int nb = sys->nb;
  const int *s = sys->s;
  int off[64];
  off[0] = 0;
  for (int L = 0; L < nb; ++L) off[L + 1] = off[L] + s[L];
  const int velocity = (sys->gauge == 'v');
  if (nb == 1) {
    if (with_diag)
      diag_add(sys->E[0], s[0], x, dxdt, 1);
    else
      memset(dxdt, 0, sizeof(double) * 2 * s[0]);
    return;
  }
  for (int L = 0; L < nb; ++L) {
    double *y = dxdt + 2 * off[L];
    const int first = (L == 0);
    if (first) {
      /* block 0 starts from the diagonal term (tdse.cpp:22-26) */
      if (with_diag)
        diag_add(sys->E[0], s[0], x, y, 1);
      else
        memset(y, 0, sizeof(double) * 2 * s[0]);
    } else {
      /* D_{L-1}^T x_{L-1}, beta = 0 overwrite (tdse.cpp:37-42, 108-113) */
      const double *xm = x + 2 * off[L - 1];
      const double *A = sys->D[L - 1];
      if (velocity) {
        g_dgemv(COLMAJOR, NOTRANS, s[L], s[L - 1], F, A, s[L], xm, 2, 0.0, y,
                2);
        g_dgemv(COLMAJOR, NOTRANS, s[L], s[L - 1], F, A, s[L], xm + 1, 2, 0.0,
                y + 1, 2);
      } else {
        g_dgemv(COLMAJOR, NOTRANS, s[L], s[L - 1], -F, A, s[L], xm, 2, 0.0,
                y + 1, 2);
        g_dgemv(COLMAJOR, NOTRANS, s[L], s[L - 1], F, A, s[L], xm + 1, 2, 0.0,
                y, 2);
      }
      if (with_diag) diag_add(sys->E[L], s[L], x + 2 * off[L], y, 0);
    }
    if (L < nb - 1) {
      /* D_L x_{L+1}, beta = 1 accumulate (tdse.cpp:28-31, 50-55, 99-102,
       * 121-126) */
      const double *xp = x + 2 * off[L + 1];
      const double *A = sys->D[L];
      if (velocity) {
        g_dgemv(COLMAJOR, TRANS, s[L + 1], s[L], -F, A, s[L + 1], xp, 2, 1.0,
                y, 2);
        g_dgemv(COLMAJOR, TRANS, s[L + 1], s[L], -F, A, s[L + 1], xp + 1, 2,
                1.0, y + 1, 2);
      } else {
        g_dgemv(COLMAJOR, TRANS, s[L + 1], s[L], -F, A, s[L + 1], xp, 2, 1.0,
                y + 1, 2);
        g_dgemv(COLMAJOR, TRANS, s[L + 1], s[L], F, A, s[L + 1], xp + 1, 2,
                1.0, y, 2);
      }
    }
  }
}

void or_rhs(const or_sys *sys, double F, const double *x, double *dxdt) {
  rhs_impl(sys, F, x, dxdt, 1);
}

void or_coupling(const or_sys *sys, const double *x, double *out) {
  rhs_impl(sys, 1.0, x, out, 0);
}

/* ------------------------------------------------------------------ */
/* Fehlberg 7(8) tableau (E. Fehlberg, NASA TR R-287, 1968), as used by   */
/* boost::odeint runge_kutta_fehlberg78.                                */
/* ------------------------------------------------------------------ */
static double A_[13][12];
static double B_[13];  /* order-8 weights (the propagated solution) */
static double DB_[13]; /* b - bhat (embedded error weights) */
static double C_[13];
static int tab_ready = 0;

static void set_tableau(void) {
  if (tab_ready) return;
  memset(A_, 0, sizeof(A_));
  A_[1][0] = 2.0 / 27.0;
  A_[2][0] = 1.0 / 36.0;
  A_[2][1] = 1.0 / 12.0;
  A_[3][0] = 1.0 / 24.0;
  A_[3][2] = 1.0 / 8.0;
  A_[4][0] = 5.0 / 12.0;
  A_[4][2] = -25.0 / 16.0;
  A_[4][3] = 25.0 / 16.0;
  A_[5][0] = 1.0 / 20.0;
  A_[5][3] = 1.0 / 4.0;
  A_[5][4] = 1.0 / 5.0;
  A_[6][0] = -25.0 / 108.0;
  A_[6][3] = 125.0 / 108.0;
  A_[6][4] = -65.0 / 27.0;
  A_[6][5] = 125.0 / 54.0;
  A_[7][0] = 31.0 / 300.0;
  A_[7][4] = 61.0 / 225.0;
  A_[7][5] = -2.0 / 9.0;
  A_[7][6] = 13.0 / 900.0;
  A_[8][0] = 2.0;
  A_[8][3] = -53.0 / 6.0;
  A_[8][4] = 704.0 / 45.0;
  A_[8][5] = -107.0 / 9.0;
  A_[8][6] = 67.0 / 90.0;
  A_[8][7] = 3.0;
  A_[9][0] = -91.0 / 108.0;
  A_[9][3] = 23.0 / 108.0;
  A_[9][4] = -976.0 / 135.0;
  A_[9][5] = 311.0 / 54.0;
  A_[9][6] = -19.0 / 60.0;
  A_[9][7] = 17.0 / 6.0;
  A_[9][8] = -1.0 / 12.0;
  A_[10][0] = 2383.0 / 4100.0;
  A_[10][3] = -341.0 / 164.0;
  A_[10][4] = 4496.0 / 1025.0;
  A_[10][5] = -301.0 / 82.0;
  A_[10][6] = 2133.0 / 4100.0;
  A_[10][7] = 45.0 / 82.0;
  A_[10][8] = 45.0 / 164.0;
  A_[10][9] = 18.0 / 41.0;
  A_[11][0] = 3.0 / 205.0;
  A_[11][5] = -6.0 / 41.0;
  A_[11][6] = -3.0 / 205.0;
  A_[11][7] = -3.0 / 41.0;
  A_[11][8] = 3.0 / 41.0;
  A_[11][9] = 6.0 / 41.0;
  A_[12][0] = -1777.0 / 4100.0;
  A_[12][3] = -341.0 / 164.0;
  A_[12][4] = 4496.0 / 1025.0;
  A_[12][5] = -289.0 / 82.0;
  A_[12][6] = 2193.0 / 4100.0;
  A_[12][7] = 51.0 / 82.0;
  A_[12][8] = 33.0 / 164.0;
  A_[12][9] = 12.0 / 41.0;
  A_[12][11] = 1.0;
  const double b[13] = {0, 0, 0, 0, 0, 34.0 / 105.0, 9.0 / 35.0, 9.0 / 35.0,
                        9.0 / 280.0, 9.0 / 280.0, 0, 41.0 / 840.0,
                        41.0 / 840.0};
  const double bh[13] = {41.0 / 840.0, 0, 0, 0, 0, 34.0 / 105.0, 9.0 / 35.0,
                         9.0 / 35.0, 9.0 / 280.0, 9.0 / 280.0, 41.0 / 840.0,
                         0, 0};
  const double c[13] = {0, 2.0 / 27.0, 1.0 / 9.0, 1.0 / 6.0, 5.0 / 12.0, 0.5,
                        5.0 / 6.0, 1.0 / 6.0, 2.0 / 3.0, 1.0 / 3.0, 1.0, 0,
                        1.0};
  for (int i = 0; i < 13; ++i) {
    B_[i] = b[i];
    DB_[i] = b[i] - bh[i];
    C_[i] = c[i];
  }
  tab_ready = 1;
}

/* Exposed so tests can read back exactly what the oracle integrates with. */
void or_tableau(double *a13x12, double *b13, double *db13, double *c13) {
  set_tableau();
  memcpy(a13x12, A_, sizeof(A_));
  memcpy(b13, B_, sizeof(B_));
  memcpy(db13, DB_, sizeof(DB_));
  memcpy(c13, C_, sizeof(C_));
}

/* work: k[0..12] then xtmp (14 * 2N doubles). */
void or_rkf78_step_stages(const or_sys *sys, const double *F13, double *x,
                          double dt, double *work, double *err,
                          double *k1_out) {
  set_tableau();
  const int n2 = 2 * or_state_size(sys);
  double *k[13];
  for (int s = 0; s < 13; ++s) k[s] = work + (size_t)s * n2;
  double *xt = work + (size_t)13 * n2;
  or_rhs(sys, F13[0], x, k[0]);
  for (int s = 1; s < 13; ++s) {
    /* x_tmp = 1.0*x + sum_j (dt*a_sj) k_j, left to right (boost scale_sum) */
    for (int i = 0; i < n2; ++i) {
      double v = 1.0 * x[i];
      for (int j = 0; j < s; ++j) v += (dt * A_[s][j]) * k[j][i];
      xt[i] = v;
    }
    or_rhs(sys, F13[s], xt, k[s]);
  }
  if (k1_out) memcpy(k1_out, k[0], sizeof(double) * n2);
  if (err) {
    for (int i = 0; i < n2; ++i) {
      double e = 0.0;
      for (int j = 0; j < 13; ++j) e += (dt * DB_[j]) * k[j][i];
      err[i] = e;
    }
  }
  for (int i = 0; i < n2; ++i) {
    double v = 1.0 * x[i];
    for (int j = 0; j < 13; ++j) v += (dt * B_[j]) * k[j][i];
    x[i] = v;
  }
}

void or_rkf78_step(const or_sys *sys, double F, double *x, double dt,
                   double *work) {
  double F13[13];
  for (int s = 0; s < 13; ++s) F13[s] = F;
  or_rkf78_step_stages(sys, F13, x, dt, work, NULL, NULL);
}

double or_dznrm2(int n_complex, const double *x) {
  double ss = 0.0;
  for (int i = 0; i < 2 * n_complex; ++i) ss += x[i] * x[i];
  return sqrt(ss);
}

static double *alloc_work(const or_sys *sys) {
  size_t n2 = 2 * (size_t)or_state_size(sys);
  return (double *)malloc(sizeof(double) * 16 * n2);
}

/* propV/propL loop body (tdse.cpp:186-204), without the file output. */
double or_prop_fixed(const or_sys *sys, const or_pulse *p, double t0,
                     double dt, int steps, double *x, int track,
                     double *pop_out, double *field_out) {
  const int n = or_state_size(sys);
  double *work = alloc_work(sys);
  double t = t0;
  for (int st = 0; st < steps; ++st) {
    double F = or_field(p, t + dt);
    if (field_out) field_out[st] = F;
    or_rkf78_step(sys, F, x, dt, work);
    t += dt;
    double nrm = or_dznrm2(n, x);
    for (int i = 0; i < 2 * n; ++i) x[i] /= nrm;
    if (pop_out)
      pop_out[st] = x[2 * track] * x[2 * track] +
                    x[2 * track + 1] * x[2 * track + 1];
  }
  free(work);
  return t;
}

/* Stage-time fields F(t + c_s h). */
static void stage_fields(const or_pulse *p, double t, double h, double *F13) {
  set_tableau();
  for (int s = 0; s < 13; ++s) F13[s] = or_field(p, t + C_[s] * h);
}

int or_prop_adaptive(const or_sys *sys, const or_pulse *p, double t0,
                     double t1, double h0, double atol, double rtol,
                     int max_attempts, double *x, or_stats *st,
                     double *h_log, int max_log) {
  const int n = or_state_size(sys);
  const size_t n2 = 2 * (size_t)n;
  double *work = alloc_work(sys);
  double *y = (double *)malloc(sizeof(double) * n2 * 3);
  double *err = y + n2;
  double *k1 = y + 2 * n2;
  double t = t0, h = h0;
  int acc = 0, rej = 0, attempts = 0, rc = 0;
  while (t < t1) {
    if (attempts++ >= max_attempts) {
      rc = -1;
      break;
    }
    int last = (t + h >= t1);
    double hh = last ? (t1 - t) : h;
    double F13[13];
    stage_fields(p, t, hh, F13);
    memcpy(y, x, sizeof(double) * n2);
    or_rkf78_step_stages(sys, F13, y, hh, work, err, k1);
    /* default_error_checker: max_i |err_i| / (atol + rtol (|x_i| + h|k1_i|)) */
    double E = 0.0;
    for (size_t i = 0; i < n2; ++i) {
      double r = fabs(err[i]) / (atol + rtol * (fabs(x[i]) + hh * fabs(k1[i])));
      if (r > E) E = r;
    }
    if (E > 1.0) {
      /* reject: dt *= max(0.9 E^(-1/(q-1)), 1/5) with error order q = 7 */
      double f = 0.9 * pow(E, -1.0 / 6.0);
      h = hh * (f > 0.2 ? f : 0.2);
      ++rej;
      continue;
    }
    double nrm = or_dznrm2(n, y);
    for (size_t i = 0; i < n2; ++i) x[i] = y[i] / nrm;
    if (h_log && acc < max_log) h_log[acc] = hh;
    ++acc;
    if (last) {
      t = t1;
      break;
    }
    t += hh;
    if (E < 0.5) {
      /* increase: dt *= 0.9 max(E, 5^-8)^(-1/8) with stepper order 8 */
      double Em = E > 1.0 / 390625.0 ? E : 1.0 / 390625.0;
      h = hh * (0.9 * pow(Em, -1.0 / 8.0));
    } else {
      h = hh;
    }
  }
  if (st) {
    st->accepted = acc;
    st->rejected = rej;
    st->t = t;
    st->h = h;
  }
  free(y);
  free(work);
  return rc;
}

void or_prop_schedule(const or_sys *sys, const or_pulse *p, double t0, int n,
                      const double *hs, double *x) {
  const int ns = or_state_size(sys);
  double *work = alloc_work(sys);
  double t = t0;
  for (int k = 0; k < n; ++k) {
    double F13[13];
    stage_fields(p, t, hs[k], F13);
    or_rkf78_step_stages(sys, F13, x, hs[k], work, NULL, NULL);
    double nrm = or_dznrm2(ns, x);
    for (int i = 0; i < 2 * ns; ++i) x[i] /= nrm;
    t += hs[k];
  }
  free(work);
}

void or_observables(const or_sys *sys, const double *x, double thr,
                    double *out) {
  /* pes.cpp:178-213, 250-254 */
  double yield = 0.0, exc = 0.0, nrm = 0.0;
  int off = 0;
  for (int L = 0; L < sys->nb; ++L) {
    const int sz = sys->s[L];
    for (int i = 1; i < sz - 1; ++i) {
      const double *c = x + 2 * (off + i);
      double p = c[0] * c[0] + c[1] * c[1];
      if (sys->E[L][i] - thr > 0.0) yield += p;
      if (off + i > 0) exc += p;
    }
    off += sz;
  }
  for (int i = 0; i < off; ++i) nrm += x[2 * i] * x[2 * i] + x[2 * i + 1] * x[2 * i + 1];
  out[0] = x[0] * x[0] + x[1] * x[1];
  out[1] = yield;
  out[2] = exc;
  out[3] = nrm;
}

void or_observables_1e(const or_sys *sys, const double *x, double *out) {
  /* pes.cpp:90-110 (yield), 138-146 (ground population, norm).  PES_En is
   * the l = 1 block of energies; state i >= 1 of block l counts toward the
   * yield through the first bin k in [1, s_0 - 2] that holds its energy, if
   * PES_En[k] > 0. */
  const double *pes_en = sys->E[1];
  double yield = 0.0, nrm = 0.0;
  int off = 0;
  for (int l = 0; l < sys->nb; ++l) {
    for (int i = 1; i < sys->s[l]; ++i) {
      for (int k = 1; k < sys->s[0] - 1; ++k) {
        double ndiff = fabs(pes_en[k] - pes_en[k - 1]);
        double pdiff = fabs(pes_en[k + 1] - pes_en[k]);
        double en = sys->E[l][i];
        if (en > pes_en[k] - 0.5 * ndiff && en < pes_en[k] + 0.5 * pdiff) {
          const double *c = x + 2 * (off + i);
          yield += (pes_en[k] > 0) * (c[0] * c[0] + c[1] * c[1]);
          break;
        }
      }
    }
    off += sys->s[l];
  }
  for (int i = 0; i < off; ++i)
    nrm += x[2 * i] * x[2 * i] + x[2 * i + 1] * x[2 * i + 1];
  out[0] = x[0] * x[0] + x[1] * x[1];
  out[1] = yield;
  out[2] = 0.0;
  out[3] = nrm;
}
