/*
 * mpci_tdse.h -- C ABI of the MI355X (gfx950) tdse propagation engine.
 *
 * Drop-in replacement for the propagation path of MP-CITDSE's `tdse`
 * program.  Plain C types only: host pointers and sizes, no torch or HIP
 * types.  Every function returns MPCI_OK (0) or a negative MPCI_ERR_* code
 * and never throws; mpci_last_error() returns the message of the last
 * failure on the calling thread.  All arithmetic is FP64.
 *
 * Objects:
 *   mpci_system  the device-resident block-tridiagonal operator: energies
 *                E_L and dipole blocks D_L (what tdrd::readEnergies and
 *                tdrd::readDipoles load, src/tdse/td_read.cpp:50-120, and
 *                what MatVecV / MatVecL hold, src/tdse/tdse.cpp:5-12,76-83).
 *   mpci_prop    P independent pulses propagated on one system, each with
 *                its own coefficient vector, field and step control.
 *
 * State vectors are 2N doubles, interleaved (Re c_k, Im c_k), exactly the
 * reference's `ct` (src/tdse/tdse_main.cpp:148-150).  Host buffers stay
 * owned by the caller and are copied; no caller pointer is retained after a
 * call returns.  A mpci_prop owns one HIP stream; it is not thread-safe,
 * several may coexist on one device.  Calls block until their work is done.
 */
#ifndef MPCI_TDSE_H_
#define MPCI_TDSE_H_

#ifdef __cplusplus
extern "C" {
#endif

#define MPCI_OK 0
#define MPCI_ERR_ARG (-1)      /* invalid argument */
#define MPCI_ERR_HIP (-2)      /* HIP runtime error (no GPU, OOM, ...) */
#define MPCI_ERR_STATE (-3)    /* object not in a state for this call */
#define MPCI_ERR_LIMIT (-4)    /* attempt limit reached (adaptive) */
#define MPCI_ERR_NOTSUP (-5)   /* combination not supported */

#define MPCI_PULSE_SINE2 0     /* sin^2 envelope, pulse.cpp:41-56,69-97 */
#define MPCI_PULSE_GAUSS 1     /* gaussian envelope, pulse.cpp:10-39 */
#define MPCI_PULSE_PWL 2       /* extension: piecewise-linear envelope */

/* pulse::params (src/tdse/include/pulse.hpp:20-26) plus, for
 * MPCI_PULSE_PWL, F(t) = scale * env(t) * sin(w t + cepd) on [0, t_end]
 * with env interpolating n_env samples placed uniformly on [0, t_end]. */
typedef struct mpci_pulse {
  int kind;
  double Eo, w, cepd, a, b;
  double t_end, scale;
  int n_env;
  const double *env;
} mpci_pulse;

typedef struct mpci_stats {
  int accepted;   /* accepted steps */
  int rejected;   /* rejected attempts */
  int attempts;
  int status;     /* 0 ok, MPCI_ERR_LIMIT if max_attempts was reached */
  double t;       /* final time */
  double h;       /* next proposed step */
} mpci_stats;

typedef struct mpci_adjoint_stats {
  int steps_max;   /* most accepted steps of any pulse */
  int segments;    /* checkpoint segments swept backward */
  double fwd_ms;   /* forward pass (adaptive, with checkpoints) */
  double bwd_ms;   /* segment recomputation + backward sweep */
  long long peak_bytes;  /* most device memory the call's own buffers held */
  int ckpt_slots;  /* checkpoint slots allocated (grown with the run) */
} mpci_adjoint_stats;

/* Rows of the observable weights (mpci_obs_weights). */
#define MPCI_OBS_GROUND 0    /* ground-state population |c_0|^2 */
#define MPCI_OBS_YIELD 1     /* ionisation yield */
#define MPCI_OBS_EXCITED 2   /* excited population (two electrons) */
#define MPCI_OBS_NORM 3      /* sum_k |c_k|^2 */
#define MPCI_OBS_ROWS 4

typedef struct mpci_system mpci_system;
typedef struct mpci_prop mpci_prop;

const char *mpci_last_error(void);
/* Number of visible GPUs (0 without a GPU). */
int mpci_device_count(int *n);

/* ---- pulses: src/tdse/pulse.cpp ------------------------------------- */
/* pulse::toAU (pulse.cpp:3-6): W/cm^2, eV -> atomic units. */
void mpci_to_au(double Io_W, double w_eV, double *Io_au, double *w_au);
/* pulse::period (pulse.cpp:8). */
double mpci_period(double w, int cycles);
/* pulse::sineASetup (gauge 'v') / sineESetup (gauge 'l'), pulse.cpp:41-75. */
int mpci_pulse_sine(double Io, double w, double cepd, int cycles, char gauge,
                    mpci_pulse *out);
/* pulse::gaussASetup / gaussESetup, pulse.cpp:10-32. */
int mpci_pulse_gauss(double Io, double w, double tau, int cycles, double cepd,
                     char gauge, mpci_pulse *out);
/* Host evaluation: sineAA / sineEE / gauss (pulse.cpp:34-39,52-56,93-97). */
double mpci_pulse_eval(const mpci_pulse *p, double t);

/* ---- system ---------------------------------------------------------- */
/* Replaces the MatVecV / MatVecL set-up of tdse::propV / propL
 * (src/tdse/tdse.cpp:160-169, 243-252): nb = L_max + 1 blocks of sizes
 * s[L]; E[L] has s[L] energies; D[L] (L < nb - 1) is the s[L] x s[L+1]
 * row-major dipole block read by tdrd::readDipoles (td_read.cpp:88-120).
 * gauge 'v' = MatVecV (velocity), 'l' = MatVecL (length). */
int mpci_system_create(int nb, const int *s, char gauge,
                       const double *const *E, const double *const *D,
                       mpci_system **out);
int mpci_system_destroy(mpci_system *sys);
/* Tuning: number of workgroups of the persistent single-pulse coupling
 * kernel (each streams every G-th tile of the dipole store); 0 restores the
 * default of 2 per compute unit.  Results do not depend on it. */
int mpci_system_set_gemv_grid(mpci_system *sys, int workgroups);
/* N = sum of s; bytes = device memory held by the system. */
int mpci_system_info(const mpci_system *sys, int *N, long long *bytes);

/* ---- propagator ------------------------------------------------------ */
/* P pulses (1 <= P <= 256).  P = 1 streams D once per RHS through the fused
 * GEMV kernel (HBM-bound); P >= 2 runs the coupling as FP64 MFMA GEMMs with
 * the tile shape picked by an on-device autotuner at the first creation for
 * this system and P; later propagators of the same system and P reuse it,
 * so they sum in the same order and give bitwise-equal results. */
int mpci_prop_create(mpci_system *sys, int P, mpci_prop **out);
int mpci_prop_destroy(mpci_prop *prop);
/* ct: 2N doubles for pulse p (tdse_main.cpp:148-160 initial state). */
int mpci_prop_set_state(mpci_prop *prop, int p, const double *ct);
int mpci_prop_get_state(mpci_prop *prop, int p, double *ct);
int mpci_prop_set_pulse(mpci_prop *prop, int p, const mpci_pulse *pulse);

/* One RHS evaluation per pulse at frozen fields F[p]: dxdt = f(x) for the
 * 2N-double vectors x[p] (MatVecV/MatVecL::operator(), tdse.cpp:14,85).
 * x and dxdt hold P consecutive 2N-double vectors.  Overwrites the state. */
int mpci_prop_rhs(mpci_prop *prop, const double *F, const double *x,
                  double *dxdt);

/* No counterpart in the reference.  For callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops): the
 * RHS of one pulse, dxdt = f(x) at constant field F, as mpci_prop_rhs with
 * P = 1 computes it on the large-system path (the same kernels and
 * summation order), but with x and dxdt the caller's own device buffers of
 * 2N doubles (16-byte aligned) and no hipMalloc, hipFree or allocating
 * library call.  The work is enqueued on `stream` (a hipStream_t; NULL is
 * the null stream) and the call returns without waiting.  The coupling
 * partials live in a fixed per-device scratch array of 10 MiB inside the
 * library (the He shape of 4 blocks of 4000 states needs 9.2 MB); a system
 * that needs more is refused with MPCI_ERR_ARG.  Every call on one device
 * shares that array: calls whose work overlaps on a device (other streams,
 * other threads) race and must be ordered by the caller. */
int mpci_system_rhs_noalloc(const mpci_system *sys, double F, const double *x,
                            double *dxdt, void *stream);

/* Reference mode (tdse::propV / propL loop, tdse.cpp:186-204): `steps`
 * fixed RKF78 steps of size dt with the field frozen per step at
 * F[p * steps + st] (the caller evaluates field(pars, t + dt)), each step
 * followed by division by the 2-norm.  pop (P x steps, may be NULL) gets
 * |c_track|^2 after each step.  t0 only labels the steps. */
int mpci_prop_fixed(mpci_prop *prop, double t0, double dt, int steps,
                    const double *F, int track, double *pop);

/* Adaptive mode (extension; the reference has none): embedded-error
 * controlled RKF78 from t0[p] to t1[p] (last step clipped) with initial
 * step h0[p], field evaluated from the pulse at the stage times, error
 * ratio max_i |err_i| / (atol + rtol (|x_i| + h |k1_i|)), renormalised
 * after each accepted step.  stats: P entries.  log (may be NULL): per
 * pulse log_cap entries of (t, h, |c_track|^2) for the accepted steps. */
int mpci_prop_adaptive(mpci_prop *prop, const double *t0, const double *t1,
                       const double *h0, double atol, double rtol,
                       int max_attempts, int track, mpci_stats *stats,
                       double *log, int log_cap);

/* Multi-instance adaptive propagation (extension; the reference runs one
 * system per process, src/tdse/tdse_main.cpp:169-194).  Runs
 * mpci_prop_adaptive(props[i], t0[i], t1[i], h0[i], atol, rtol,
 * max_attempts, track 0, stats[i], no log) for n distinct propagators, which
 * may hold different systems and batch sizes, concurrently on one device:
 * each member keeps its own stream and its next chunk is launched as soon
 * as its previous one drains.  Results equal n separate calls.  stats[i] may
 * be NULL.  Returns MPCI_ERR_LIMIT when any pulse hit max_attempts. */
int mpci_multi_adaptive(int n, mpci_prop *const *props,
                        const double *const *t0, const double *const *t1,
                        const double *const *h0, double atol, double rtol,
                        int max_attempts, mpci_stats *const *stats);

/* The same propagation with every member's attempt in the same launches
 * (extension): one grouped GEMM launch per stage over an interleaved list
 * of all members' panels, one stage launch over all members' rows, one
 * step-control and one apply launch over all their pulses, captured once
 * as a graph.  Each member keeps its own K split, so its results equal a
 * separate mpci_prop_adaptive call bit for bit.  1 <= n <= 8; every
 * member must be a batched propagator (P >= 2) with two or more blocks on
 * the same device, otherwise MPCI_ERR_NOTSUP (use mpci_multi_adaptive). */
int mpci_multi_adaptive_grouped(int n, mpci_prop *const *props,
                                const double *const *t0,
                                const double *const *t1,
                                const double *const *h0, double atol,
                                double rtol, int max_attempts,
                                mpci_stats *const *stats);

/* Replays the adaptive map over n given steps hs[p * n + k] from t0[p]
 * (no controller): used for frozen-schedule finite differences. */
int mpci_prop_schedule(mpci_prop *prop, const double *t0, int n,
                       const double *hs);

/* Discrete adjoint (extension; the reference has none).  All P pulses must
 * be MPCI_PULSE_PWL.  Runs the adaptive forward propagation of
 * mpci_prop_adaptive from the current states, keeping the state after every
 * ckpt_every accepted steps, then sweeps backward through the same discrete
 * map (accepted steps only, schedule frozen, renormalisation included),
 * recomputing each checkpoint segment.  Objective per pulse:
 * J[p] = sum_k W[k] |c_{k,p}(t1[p])|^2 (e.g. the ground-state population or
 * the yield mask of pes::genPES2e, src/pes/pes.cpp:199-213).  Output:
 * grad[p * grad_stride + k] = dJ[p] / d env_p[k]; nsteps[p] and hlog
 * (P x hlog_cap, may be NULL) return the accepted step sizes for
 * frozen-schedule finite differences.  Leaves the propagator at the final
 * states.  Fails with MPCI_ERR_LIMIT beyond max_steps accepted steps. */
int mpci_prop_adjoint(mpci_prop *prop, const double *t0, const double *t1,
                      const double *h0, double atol, double rtol,
                      int max_steps, const double *W, int ckpt_every,
                      double *J, double *grad, int grad_stride, int *nsteps,
                      double *hlog, int hlog_cap, mpci_adjoint_stats *stats);

/* On-device weighted populations: out[o * P + p] = sum_k W[o*N + k]
 * |c_{k,p}|^2 (ground population, yield and excitation masks of
 * pes::genPES2e, src/pes/pes.cpp:199-213; all-ones for the norm). */
int mpci_prop_populations(mpci_prop *prop, int nobs, const double *W,
                          double *out);

/* ---- observables: src/pes/pes.cpp ------------------------------------ */
/* The scalars pes prints, as rows of per-state weights: observable
 * o = sum_k W[o * N + k] |c_k|^2, N = sum of s, rows MPCI_OBS_*.
 *   GROUND   |c_0|^2 (pes.cpp:145, 211-213).
 *   YIELD    electrons 2 (genPES2e, pes.cpp:199-205): the states i in
 *            [1, s_L - 2] of every block L with E_L[i] - threshold > 0,
 *            threshold = the lowest one-electron l = 0 energy, which pes
 *            reads from dat/<pot>0.h5:/En[0] (pes.cpp:168-176).
 *            electrons 1 (genPES1e, pes.cpp:90-110): the states i >= 1 of
 *            every block whose energy falls in a bin k in [1, s_0 - 2] of
 *            the l = 1 energies (bin k spans (E_1[k] - |E_1[k]-E_1[k-1]|/2,
 *            E_1[k] + |E_1[k+1]-E_1[k]|/2), the first such bin counts) with
 *            E_1[k] > 0; threshold is unused; needs nb >= 2, s_0 <= s_1.
 *   EXCITED  electrons 2: the states i in [1, s_L - 2] with o_L + i > 0
 *            (pes.cpp:206-208); electrons 1: all zero (pes -e 1 prints no
 *            excited population).
 *   NORM     all ones (pes.cpp:138-142, 250-254).
 * E[L] holds the s[L] energies of block L as tdse reads them
 * (tdrd::readEnergies).  W: MPCI_OBS_ROWS x N doubles.  Host code only. */
int mpci_obs_weights(int nb, const int *s, const double *const *E,
                     int electrons, double threshold, double *W);
/* The same observables of every pulse's current state, evaluated on the
 * device from the system's energies: out[o * P + p] for o < MPCI_OBS_ROWS.
 * The weights stay on the device between calls of the same kind. */
int mpci_prop_observables(mpci_prop *prop, int electrons, double threshold,
                          double *out);

/* Launch the 13-stage step as one captured hipGraph (1, default) or as
 * individual launches (0). */
int mpci_prop_set_graph(mpci_prop *prop, int enable);
/* One pulse on a small system: each stage as one fused launch (coupling
 * from row-major copies of D_L and D_L^T, one wave per row, then the stage
 * algebra) instead of the GEMV tile kernel and the stage kernel.  mode -1
 * (default): on when P = 1, the system has more than one block and its
 * dipoles take at most 12 MiB; 0: off; 1: on whenever P = 1 and there is
 * more than one block.  active (may be NULL) reports the outcome.  Both
 * paths give the same result to rounding (the coupling sums in another
 * order). */
int mpci_prop_set_small_path(mpci_prop *prop, int mode, int *active);
/* Instrumentation: when on, steps are launched individually and every
 * coupling kernel carries a pair of HIP events on its own dispatch
 * (hipExtLaunchKernelGGL), so each interval is that kernel's execution;
 * the durations of the attempts that did work accumulate into
 * mpci_prop_kernel_time.  Off by default. */
int mpci_prop_set_instrument(mpci_prop *prop, int enable);
int mpci_prop_kernel_time(mpci_prop *prop, double *total_ms,
                          long long *launches, long long *bytes_per_launch,
                          long long *flops_per_launch);
/* Batched form (P >= 2): the MFMA tile the autotuner chose at
 * mpci_prop_create -- wm x wn waves per workgroup, each wave fm x fn blocks
 * of 16 x 16 (a panel of 16 fm wm rows by 16 fn wn real columns) -- its K
 * split, kernel (0: single LDS buffer, 1: double-buffered with 16-byte
 * loads through registers, 2: double-buffered with direct-to-LDS loads,
 * 3: both coupling directions from one read of each dipole block, 16 waves,
 * ks then being the workgroups per compute unit its split aims at) and
 * K chunk, and its measured time per coupling evaluation in ms (the GEMM
 * and the stage kernel that reads its partial sums back).
 * Any pointer may be NULL. */
int mpci_prop_gemm_config(mpci_prop *prop, int *wm, int *wn, int *ks,
                          double *ms, int *kind, int *tk, int *fm, int *fn);
/* The autotuner's candidate tiles: their number, and pinning one of them
 * (0 <= index < count) on a batched propagator in place of the autotuned
 * choice (it is timed as the autotuner times it; mpci_prop_gemm_config
 * then reports it).  For tests and tuning studies; the pin holds for this
 * propagator only. */
int mpci_prop_gemm_candidates(int *count);
int mpci_prop_set_gemm_candidate(mpci_prop *prop, int index);

/* ---- measurement probes --------------------------------------------- */
/* No counterpart in the reference: the ceilings the current device reaches
 * in practice, reported beside the published peaks by the benchmark.
 * mpci_probe_hbm_read: a buffer of `bytes` (>= 1 MiB) streamed once per
 * launch with 16-byte non-temporal loads, `reps` timed launches at 2, 4 and
 * 8 workgroups per CU; the best rate in GB/s (1e9 bytes/s).
 * mpci_probe_mfma_f64: v_mfma_f64_16x16x4f64 back to back (four
 * independent chains per wave, two waves per SIMD, `iters` >= 16 steps per
 * chain, no memory traffic); the best of three launches in TFLOP/s.
 * Both allocate their own device memory and use the null stream. */
int mpci_probe_hbm_read(long long bytes, int reps, double *gbps);
int mpci_probe_mfma_f64(int iters, double *tflops);

#ifdef __cplusplus
}
#endif
#endif /* MPCI_TDSE_H_ */
