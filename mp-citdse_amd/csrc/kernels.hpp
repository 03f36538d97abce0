// Kernel declarations of the MI355X tdse engine.
#pragma once

#include "common.hpp"

namespace mpci {

struct StageArgs {
  const double2 *X;      // committed state (N x P)
  double2 *Xs;           // stage input, rewritten for the next stage
  double2 *K;            // 13 stage derivatives, stride NP
  double2 *Y;            // candidate state of the attempt
  const double2 *part;   // coupling partial sums
  const double *E;       // energies (N)
  double *red;           // per-block (error max, |y|^2) partials
  Ctl *ctl;              // per-pulse control (P)
  const PulseDev *pulses;
  const double *env;     // envelope samples of PWL pulses
  const double *Ftab;    // frozen per-step fields (P x nsteps)
  long long NP;
  int s;                 // stage 0..12
  int pad;
};

__global__ void k_relayout_gemv(const double *D, int sL, int sL1, int ncb,
                                long long nelem, double *out);
__global__ void k_gemv_tile(const TileJob *jobs, const double *Dg,
                            const double2 *X, double2 *part, SysDesc d,
                            const Ctl *ctl);
__global__ void k_stage(StageArgs a, SysDesc d);
__global__ void k_finalize(Ctl *ctl, const double *red, int nblk, int P,
                           const double *hs);
__global__ void k_apply(double2 *X, const double2 *Y, const Ctl *ctl,
                        LogEnt *log, int logcap, int track, int N, int P);
__global__ void k_weighted_pop(const double2 *X, const double *W, int N, int P,
                               double *out);

}  // namespace mpci
