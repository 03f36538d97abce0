// Shared definitions of the MI355X tdse engine (host and device).
//
// The engine propagates dc/dt = -i[diag(E) + F(t) K] c for P pulses that
// share one block-tridiagonal system (blocks L = 0..nb-1, couplings only
// between L and L+1), the operator of MatVecV / MatVecL in the reference
// (src/tdse/tdse.cpp:14-143).  All arithmetic is FP64.
//
// State layout in HBM: complex coefficient (row i, pulse p) at double2
// index i * P + p, i.e. for P = 1 exactly the reference's interleaved
// (Re, Im) vector (tdse_main.cpp:148).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace mpci {

constexpr int kMaxBlocks = 32;   // L_max <= 31
constexpr int kStages = 13;      // Fehlberg 7(8)

// ---- fused GEMV tile geometry (single pulse, HBM-bound form) ----------
// A workgroup of 512 lanes streams TR x TC tiles of the dipole blocks D_L
// (persistently, several per workgroup) and produces for each, in one
// pass, the row partials of D_L x_{L+1} and the column partials of
// D_L^T x_L.  Lane t covers rows (t & 15) + 16 i and
// column pairs (t >> 4) + 32 j.
constexpr int kGvThreads = 512;
constexpr int kGvRowIter = 8;                  // i
constexpr int kGvColPairIter = 4;              // j
constexpr int kGvTR = 16 * kGvRowIter;         // 128 rows
constexpr int kGvTC = 2 * 32 * kGvColPairIter; // 256 cols
constexpr int kGvTileElems = kGvTR * kGvTC;

struct TileJob {
  int L;        // pair index (D_L couples block L and L+1)
  int rb;       // row tile
  int cb;       // column tile
  int pad;
  long long doff;  // job index (the tile's slot in the interleaved store)
};

// ---- batched GEMM form (P >= 2 pulses, FP64 MFMA) ----------------------
// C = op(D_L) X over a K range, op = identity (dir 0: rows of block L, from
// X_{L+1}) or transpose (dir 1: rows of block L+1, from X_L).  X is the
// state viewed as a row-major matrix with 2P real columns.
constexpr int kGmTK = 32;  // K chunk staged in LDS
// Lanes per row of the single-pulse stage kernel (k_stage1): they split a
// row's partial sums and its stage-vector loads, all issued in one round.
constexpr int kStageSub1 = 16;

struct GemmJob {
  int L;
  int dir;
  int m0;       // first output row of the panel
  int n0;       // first real column
  int k0, k1;   // K range [k0, k1)
  int q;        // partial slot (K-split index)
  int ld;       // row stride of D_L in the row-major copy (even)
  int member;   // grouped launches: the member the panel belongs to
  int qt;       // k_gemm_pair: the D^T x_L partial slot (strip index); q
                // is then the D x_{L+1} slot (column-range index)
};

// k_gemm_pair geometry: a workgroup of 16 waves takes a strip of kPrR rows
// of D_L over a range of its columns, 16 columns (one LDS chunk) at a time,
// for one 32-column panel of X.
constexpr int kPrR = 256;
constexpr int kPrTK = 16;
constexpr int kPrTN = 32;

// Per-propagator description of the blocks and the partial-sum buffers.
struct SysDesc {
  int nb;
  int N;              // sum of block sizes
  int P;              // pulses
  int gauge;          // 0 velocity ('v'), 1 length ('l')
  int off[kMaxBlocks + 1];
  // pair L: nrow_part[L] partial vectors for D_L x_{L+1} (length s_L) and
  // ncol_part[L] partial vectors for D_L^T x_L (length s_{L+1}).
  int nrow_part[kMaxBlocks];
  int ncol_part[kMaxBlocks];
  long long row_part_off[kMaxBlocks];  // double2 offset (already times P)
  long long col_part_off[kMaxBlocks];
};

enum FieldMode : int { kFieldTable = 0, kFieldPulse = 1, kFieldConst = 2 };
enum StepMode : int { kModeFixed = 0, kModeAdaptive = 1, kModeSchedule = 2,
                      kModeRhs = 3 };
enum PulseKind : int { kSine2 = 0, kGauss = 1, kPwl = 2 };

// Device-side pulse::params (src/tdse/include/pulse.hpp:20-26) + PWL.
struct PulseDev {
  int kind;
  int n_env;
  long long env_off;   // offset into the propagator's envelope store
  double Eo, w, cepd, a, b;
  double t_end, scale;
};

// Per-pulse step control, resident in device memory and updated on device.
struct Ctl {
  double t;        // time at the start of the current attempt
  double h;        // proposed step (controller state)
  double hc;       // step used by the current attempt
  double t_end;
  double atol, rtol;
  double Fc;       // constant field (kFieldConst)
  double nrm;      // 2-norm of the last candidate state
  double err;      // last error ratio
  double hacc;     // size of the last accepted step
  int mode;
  int fmode;
  int done;
  int last;        // current attempt is the clipped final step
  int accept;      // outcome of the last attempt
  int step;        // fixed/schedule: steps taken
  int nsteps;      // fixed/schedule: total steps
  int accepted;
  int rejected;
  int attempts;
  int max_attempts;
  int status;      // 0 ok, -1 attempt limit hit
  int nlog;        // log entries written in the current chunk
  int hstride;     // row stride of the per-pulse step / field tables
  double Fs[kStages];  // the field at the current attempt's 13 stage times
                       // (frozen-field modes: the same value 13 times),
                       // set on the device when the attempt is set up
};

// Optional stores written while stepping (adjoint support).
struct StepStore {
  double2 *xs;      // stage inputs 1..12 of every step: [step][12][N x P]
  double2 *xn;      // normalised states after every step: [step + 1][N x P]
  double *nrm;      // ||y|| of every step: [step][P]
  double2 *ckpt;    // adaptive mode: state after every ckpt_every accepted
  int ckpt_every;   //   steps, slot accepted / ckpt_every, ckpt_cap slots
  int ckpt_cap;
};

struct LogEnt {
  double t;    // time after the accepted step
  double h;    // accepted step size
  double pop;  // |c_track|^2 after renormalisation
};

}  // namespace mpci
