// Kernel declarations of the MI355X tdse engine.
#pragma once

#include "common.hpp"

namespace mpci {

struct StageArgs {
  const double2 *X;      // committed state (N x P)
  double2 *Xs;           // stage input (stages 1..12; a valid buffer of
                         // N x P at stage 0 too, where it is not used)
  double2 *Xs_next;      // where the next stage's input goes: Xs itself
                         // when the coupling ran in an earlier kernel, the
                         // other of two buffers on the fused path
  double2 *K;            // 13 stage derivatives, stride NP
  double2 *Y;            // candidate state of the attempt
  const double2 *part;   // coupling partial sums
  const double *E;       // energies (N)
  double *red;           // per-block (error max, |y|^2) partials
  Ctl *ctl;              // per-pulse control (P), with the stage fields
  double2 *xs_store;     // optional: stage inputs of every step
  long long NP;
  int s;                 // stage 0..12
  int pad;
};

// Backward (adjoint) sweep of one step for all pulses.
struct AdjArgs {
  const double2 *part;   // coupling partials of mu_s
  const double *E;
  const PulseDev *pulses;
  const double *env;
  double2 *mu;           // 13 stage adjoints, stride NP
  double2 *lam;          // adjoint of the state after the step (input)
  double2 *lamx;         // adjoint of the state before the step (output)
  const double2 *x1;     // normalised state after the step
  const double2 *xn;     // state before the step (stage 0 input)
  const double2 *xs;     // stage inputs 1..12 of the step, stride NP
  const double *dot;     // <x1, lam> per pulse
  const double *nrm;     // ||y|| per pulse
  const double *tn;      // step start time per pulse
  const double *hn;      // step size per pulse (<= 0: pulse inactive)
  double *red;           // per-block partials of dJ/dF_s
  long long NP;
  int s;
  int pad;
};

// The coupling operands of every member of a grouped GEMM launch, passed
// in the kernel arguments (so the compiler may reload them at will instead
// of holding them in registers across the K loop).
constexpr int kGroupMax = 8;
struct GroupGemm {
  const double *Dr[kGroupMax];
  const long long *dr_off[kGroupMax];
  const double *X[kGroupMax];   // stage input of stage 0
  const double *Xs[kGroupMax];  // stage input of stages 1..12
  double *part[kGroupMax];
  const SysDesc *d[kGroupMax];
  const Ctl *ctl[kGroupMax];
};

// One member of a grouped launch (several batched propagators stepped with
// one launch per kernel of the attempt): its descriptor, dipole copy,
// stage buffers and step-control inputs, and where its workgroups start in
// each grouped launch.
struct GroupMember {
  const SysDesc *d;
  const double *Dr;          // row-major dipole copy
  const long long *dr_off;
  StageArgs sa;              // X, Xs, K, Y, part, E, red, ctl, NP (s unset)
  double2 *X;                // sa.X, written by the apply step
  double *part;              // sa.part, written by the coupling
  const double *hs;
  const double *Ftab;
  const PulseDev *pulses;
  const double *env;
  LogEnt *log;
  StepStore st;
  int P, N, logcap, track;
  int nblk_stage, wg_stage0;  // stage / finalize partial workgroups
  int nblk_apply, wg_apply0;
  int pulse0;                 // first k_finalize_g workgroup
  int pad;
};

__global__ void k_relayout_gemv(const double *D, int sL, int sL1, int ncb,
                                long long nelem, int job0, int njobs,
                                double *out);
__global__ void k_gemv_tile(const TileJob *jobs, int njobs, const double *Dg,
                            const double2 *X, double2 *part, const SysDesc *d,
                            const Ctl *ctl, int rev);
template <int WM, int WN>
__global__ void k_gemm_lds(const GemmJob *jobs, const double *Dr,
                           const long long *dr_off, const double *X,
                           double *part, const SysDesc *d, const Ctl *ctl);
template <int WM, int WN>
__global__ void k_gemm_lds_g(const GemmJob *jobs, GroupGemm g, int s);
__global__ void k_gemm_pair(const GemmJob *jobs, const double *Dr,
                            const long long *dr_off, const double *X,
                            double *part, const SysDesc *d, const Ctl *ctl);
template <int WM, int WN, int TK, int FM, int FN>
__global__ void k_gemm_db(const GemmJob *jobs, const double *Dr,
                          const long long *dr_off, const double *X,
                          double *part, const SysDesc *d, const Ctl *ctl);
template <int WM, int WN>
__global__ void k_gemm_coupling(const GemmJob *jobs, const double *Dr,
                                const long long *dr_off, const double *X,
                                double *part, const SysDesc *d, const Ctl *ctl);
template <int SUB, bool LAST>
__global__ void k_stage(StageArgs a, const SysDesc *d);
template <bool LAST>
__global__ void k_stage_g(const GroupMember *m, int nm, int s);
__global__ void k_finalize_g(const GroupMember *m, int nm);
__global__ void k_apply_g(const GroupMember *m, int nm);
template <bool LAST>
__global__ void k_stage1(StageArgs a, const SysDesc *d);
template <bool LAST>
__global__ void k_small_stage(StageArgs a, const SysDesc *d, const double *Dr,
                              const long long *dr_off, const int *dr_ld,
                              const double *DrT, const long long *drt_off,
                              const int *drt_ld);
__global__ void k_finalize(Ctl *ctl, const double *red, int nblk, int P,
                           const double *hs, const double *Ftab,
                           const PulseDev *pulses, const double *env);
// The stage fields Ctl::Fs of every live pulse from its Ctl as the host set
// it up (the first attempt of a run; k_finalize sets them for the next).
__global__ void k_ctl_fields(Ctl *ctl, int P, const double *Ftab,
                             const PulseDev *pulses, const double *env);
__global__ void k_apply(double2 *X, const double2 *Y, const Ctl *ctl,
                        LogEnt *log, int logcap, int track, int N, int P,
                        StepStore st);
__global__ void k_adj_seed(const double2 *X, const double *W, int N, int P,
                           double2 *lam);
__global__ void k_dot_partial(const double2 *a, const double2 *b, int N, int P,
                              double *red);
__global__ void k_reduce_final(const double *red, int nblk, int P,
                               double *out);
__global__ void k_adj_init(AdjArgs a, const SysDesc *d);
__global__ void k_adj_stage(AdjArgs a, const SysDesc *d);
__global__ void k_weighted_pop(const double2 *X, const double *W, int N, int P,
                               double *out);
// Ctl record of one RHS evaluation at constant field F (mode RHS).
__global__ void k_rhs_ctl(Ctl *ctl, double F);
__global__ void k_copy_strided(double2 *dst, int dst_stride,
                               const double2 *src, int src_stride, int n);
__global__ void k_probe_read(const double2 *a, long long n, double *out);
__global__ void k_probe_mfma(double *out, int iters);

}  // namespace mpci
