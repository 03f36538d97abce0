/*
 * tdse_oracle.h -- CPU restatement of MP-CITDSE's `tdse` propagation path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the oracle for the HIP
 * implementation: only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it.  The product never links or calls it.
 *
 * What each piece restates (reference = aforembs/MP-CITDSE):
 *   or_rhs             MatVecV::operator() / MatVecL::operator()
 *                      (src/tdse/tdse.cpp:14-73, 85-143), including its
 *                      cblas_dgemv call pattern (ColMajor, Trans/NoTrans,
 *                      stride-2 Re/Im lanes).
 *   or_rkf78_step      boost::odeint runge_kutta_fehlberg78::do_step as
 *                      called at src/tdse/tdse.cpp:190,273 (13 RHS calls,
 *                      order-8 weights, no error estimate); the tableau is
 *                      Fehlberg's 7(8) pair (NASA TR R-287, 1968).
 *   or_prop_fixed      the propV/propL step loop (src/tdse/tdse.cpp:186-215):
 *                      field frozen at t+dt, do_step, t += dt, divide by
 *                      dznrm2, record the tracked population.
 *   or_pulse_*         src/tdse/pulse.cpp:3-97 and pulse.hpp:20-26.
 *   or_prop_adaptive   EXTENSION (no reference counterpart): embedded-error
 *                      controlled RKF78 with the field evaluated at stage
 *                      times, renormalised after each accepted step.  The
 *                      controller follows boost::odeint's documented
 *                      default_error_checker / default_step_adjuster rules
 *                      (boost absent here, version unpinned).
 *   or_prop_schedule   EXTENSION: the adaptive scheme's map replayed over a
 *                      given list of step sizes (used for finite differences).
 *   or_observables     src/pes/pes.cpp:151-262 (2e yield, excitation,
 *                      ground population, norm).
 *   or_observables_1e  src/pes/pes.cpp:48-149 (1e yield binned onto the
 *                      l = 1 energies, ground population, norm).
 *
 * Parity pins (tests/test_oracle_*.py): exact-rational order conditions of
 * the tableau; the RKF78 stability polynomial R(z) on diagonal and 2x2
 * constant systems; dense numpy -iH on tiny random blocks; closed forms of
 * the pulses; hand-computed observables (2e masks and a 1e binning case
 * with energies on and between the bin edges).  boost's exact summation order and
 * BLAS's reduction order are parity unpinned (not needed at 1e-12).
 */
#ifndef MPCI_TDSE_ORACLE_H_
#define MPCI_TDSE_ORACLE_H_

#ifdef __cplusplus
extern "C" {
#endif

/* Block-tridiagonal system: nb = L_max + 1 blocks of sizes s[L];
 * E[L] holds s[L] energies; D[L] (L < nb-1) is s[L] x s[L+1] row-major. */
typedef struct {
  int nb;
  const int *s;
  const double *const *E;
  const double *const *D;
  char gauge; /* 'v' velocity (MatVecV), 'l' length (MatVecL) */
} or_sys;

enum { OR_SINE2 = 0, OR_GAUSS = 1, OR_PWL = 2 };

/* pulse::params (pulse.hpp:20-26) plus the piecewise-linear extension. */
typedef struct {
  int kind;
  double Eo, w, cepd, a, b;
  /* OR_PWL: F(t) = scale * env(t) * sin(w t + cepd) on [0, t_end], where
   * env interpolates n_env samples placed uniformly on [0, t_end]. */
  int n_env;
  const double *env;
  double t_end;
  double scale;
} or_pulse;

/* cblas_dgemv-compatible pointer (CblasColMajor=102, NoTrans=111, Trans=112). */
typedef void (*or_dgemv_fn)(int layout, int trans, int M, int N, double alpha,
                            const double *A, int lda, const double *x,
                            int incx, double beta, double *y, int incy);

/* Swap the dgemv used by or_rhs (NULL restores the built-in loops).
 * Used by the CPU-baseline timing to run the reference's BLAS pattern. */
void or_set_dgemv(or_dgemv_fn fn);
/* dlopen a BLAS that exports cblas_dgemv; returns 0 on success. */
int or_use_blas(const char *path);

/* --- pulses (src/tdse/pulse.cpp) --- */
void or_to_au(double Io_W, double w_eV, double *Io_au, double *w_au);
double or_period(double w, int cycles);
void or_sine_a_setup(double Io, double w, double cepd, int cycles, or_pulse *p);
void or_sine_e_setup(double Io, double w, double cepd, int cycles, or_pulse *p);
void or_gauss_a_setup(double Io, double w, double tau, int cycles, double cepd,
                      or_pulse *p);
void or_gauss_e_setup(double Io, double w, double tau, int cycles, double cepd,
                      or_pulse *p);
double or_field(const or_pulse *p, double t);

/* --- RHS and step (src/tdse/tdse.cpp) --- */
/* dxdt = f(x) at frozen field F; x, dxdt interleaved (Re, Im), 2N doubles. */
void or_rhs(const or_sys *sys, double F, const double *x, double *dxdt);
/* Coupling part only: c_L = D_{L-1}^T x_{L-1} (+/-) D_L x_{L+1} with the
 * gauge's signs and Re/Im routing at unit field and zero energies. */
void or_coupling(const or_sys *sys, const double *x, double *out);
int or_state_size(const or_sys *sys);
/* One boost-semantics RKF78 step with a frozen field; work >= 14*2N. */
void or_rkf78_step(const or_sys *sys, double F, double *x, double dt,
                   double *work);
/* Same step with stage-time fields F[13]; also returns the embedded error
 * (order-8 minus order-7 solution) in err when err != NULL. */
void or_rkf78_step_stages(const or_sys *sys, const double *F13, double *x,
                          double dt, double *work, double *err,
                          double *k1_out);
double or_dznrm2(int n_complex, const double *x);
/* Copy out the tableau the oracle integrates with (a is 13x12 row-major). */
void or_tableau(double *a13x12, double *b13, double *db13, double *c13);

/* Reference-mode loop: steps fixed steps from t0; F(t+dt) frozen per step;
 * renormalise; pop_out[st] = |c_track|^2 after each step (may be NULL);
 * field_out[st] = F used (may be NULL).  Returns final t. */
double or_prop_fixed(const or_sys *sys, const or_pulse *p, double t0,
                     double dt, int steps, double *x, int track,
                     double *pop_out, double *field_out);

typedef struct {
  int accepted;
  int rejected;
  double t;      /* final time */
  double h;      /* next proposed step */
} or_stats;

/* Adaptive mode (extension).  Integrates from t0 to t1 exactly (the last
 * step is clipped).  h_log (capacity max_log) receives the accepted step
 * sizes in order (may be NULL).  Returns 0, or -1 if max_attempts hit. */
int or_prop_adaptive(const or_sys *sys, const or_pulse *p, double t0,
                     double t1, double h0, double atol, double rtol,
                     int max_attempts, double *x, or_stats *st,
                     double *h_log, int max_log);

/* Replay of the adaptive map over a given step list (no controller). */
void or_prop_schedule(const or_sys *sys, const or_pulse *p, double t0,
                      int n, const double *hs, double *x);

/* Observables (src/pes/pes.cpp:151-262).  thr = first 1e l=0 energy.
 * out[0]=ground pop, out[1]=2e yield, out[2]=excitation, out[3]=norm. */
void or_observables(const or_sys *sys, const double *x, double thr,
                    double *out);
/* One-electron observables (src/pes/pes.cpp:48-149): out[0]=ground pop,
 * out[1]=1e yield, out[2]=0 (pes -e 1 prints no excitation), out[3]=norm.
 * Needs nb >= 2 and s[0] <= s[1] (the bins read l = 1 energies up to
 * index s[0] - 1). */
void or_observables_1e(const or_sys *sys, const double *x, double *out);

#ifdef __cplusplus
}
#endif
#endif
