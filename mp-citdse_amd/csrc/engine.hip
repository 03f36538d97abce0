// Host side of the MI355X tdse engine and its C ABI (include/mpci_tdse.h).
//
// One RKF78 attempt = 13 x (coupling, stage kernel) + k_finalize + k_apply,
// the coupling being k_gemv_tile for one pulse and an MFMA GEMM for a batch;
// one pulse on a small system takes 13 fused k_small_stage launches
// instead.  k_finalize also sets the next attempt's 13 stage fields.  The attempt is captured once into a hipGraph and replayed.  Step control lives in
// device memory (Ctl), so the host only re-launches the graph in chunks and
// drains the per-step log between chunks.
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include "../../include/mpci_tdse.h"
#include "kernels.hpp"

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg) {
  g_err = msg;
  return code;
}

#define HIPCHK(call)                                                       \
  do {                                                                     \
    hipError_t e_ = (call);                                                \
    if (e_ != hipSuccess)                                                  \
      return fail(MPCI_ERR_HIP, std::string(#call) + ": " +                \
                                    hipGetErrorString(e_));                \
  } while (0)

template <class T>
struct DevBuf {
  T *p = nullptr;
  size_t n = 0;
  DevBuf() = default;
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  int alloc(size_t count) {
    if (p) {
      (void)hipFree(p);
      p = nullptr;
    }
    n = count;
    if (count == 0) return MPCI_OK;
    HIPCHK(hipMalloc(&p, sizeof(T) * count));
    return MPCI_OK;
  }
  void swap(DevBuf &o) {
    std::swap(p, o.p);
    std::swap(n, o.n);
  }
  size_t bytes() const { return sizeof(T) * n; }
};

#define TRY(x)                  \
  do {                          \
    int rc_ = (x);              \
    if (rc_ != MPCI_OK) return rc_; \
  } while (0)

constexpr double kPi = 3.14159265358979323846;

}  // namespace

using namespace mpci;

struct mpci_system {
  int nb = 0, N = 0;
  int gauge = 0;
  std::vector<int> s, off;
  DevBuf<double> E;
  DevBuf<double> Dg;       // GEMV tiles
  DevBuf<SysDesc> ddesc1;  // the one-pulse layout of the partials (device)
  std::vector<TileJob> jobs;
  DevBuf<TileJob> djobs;
  std::vector<int> nrb, ncb;   // per pair: row tiles, column tiles
  long long S = 0;             // sum_L s_L s_{L+1} (dipole elements)
  int device = 0;
  int gv_grid = 512;  // persistent GEMV workgroups: 2 per compute unit
  int ncu = 256;      // compute units of the device
  // row-major copy of every D_L for the batched (GEMM) form, built on the
  // first propagator with P >= 2 from the host blocks kept until then
  DevBuf<double> Dr;
  DevBuf<long long> dr_off;
  std::vector<int> dr_ld;  // even row stride of each block in Dr
  DevBuf<int> dr_ld_d;     // the same on the device
  // row-major D_L^T (s_{L+1} x s_L) of every pair, for the single-pulse
  // small-system path, built with it
  DevBuf<double> DrT;
  DevBuf<long long> drt_off;
  DevBuf<int> drt_ld_d;
  std::vector<std::vector<double>> hostD;
  // the energies as given, for the observable weights (pes's masks)
  std::vector<std::vector<double>> hostE;
  // autotuned GEMM tile per batch size P: every propagator of this system
  // with the same P uses the same tiles, hence the same summation order
  struct Tuned {
    int kind, wm, wn, fm, fn, tk, ks;
    double ms;
  };
  std::map<int, Tuned> tuned;
};

struct mpci_prop {
  mpci_system *sys = nullptr;
  int P = 1;
  hipStream_t stream = nullptr;
  SysDesc desc{};
  DevBuf<SysDesc> ddesc;  // desc on the device (kernels take a pointer)
  long long NP = 0;
  int nblk = 0;           // workgroups of the per-row kernels (256 rows / P)
  int nblk_stage = 0;     // stage-kernel workgroups (16 rows for one pulse)
  // one pulse on a small system: one fused launch per stage
  // (k_small_stage, four rows per workgroup) instead of GEMV tile + stage
  bool small = false;
  int nblk_small = 0;
  DevBuf<double2> X, Xs, K, Y, part;
  DevBuf<double2> xstage;  // one pulse's contiguous state (set/get_state)
  DevBuf<double2> Xs2;  // second stage-input buffer of the fused path
  DevBuf<double> red;
  DevBuf<Ctl> ctl;
  DevBuf<PulseDev> pulses;
  DevBuf<double> env;
  DevBuf<double> Ftab;
  DevBuf<double> hs;
  DevBuf<LogEnt> log;
  int logcap = 4096;
  int track = 0;
  StepStore store{};  // adjoint support: checkpoints and recompute stores
  std::vector<PulseDev> hpulses;
  std::vector<std::vector<double>> henv;
  bool use_graph = true;
  hipGraphExec_t gexec = nullptr;
  hipGraph_t graph = nullptr;
  int graph_track = -1;
  // batched form: chosen MFMA tile (waves along M, N) and K splits
  int gm_kind = 0, gm_wm = 2, gm_wn = 2, gm_fm = 2, gm_fn = 2, gm_tk = 32,
      gm_ks = 1;
  double gemm_ms = 0.0;  // autotuned time of one coupling evaluation
  std::vector<GemmJob> gjobs;
  DevBuf<GemmJob> dgjobs;
  // pes observable weights on the device (mpci_prop_observables)
  DevBuf<double> obsW, obsOut;
  int obs_electrons = 0;
  double obs_threshold = 0.0;
  bool instrument = false;
  std::vector<hipEvent_t> ev;
  int ev_base = 0;
  double kernel_ms = 0.0;
  long long kernel_launches = 0;
  ~mpci_prop() {
    if (gexec) (void)hipGraphExecDestroy(gexec);
    if (graph) (void)hipGraphDestroy(graph);
    for (auto e : ev) (void)hipEventDestroy(e);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

// ---- copies on the propagator's stream ----------------------------------
// Every host <-> device copy that touches buffers the propagator's kernels
// use goes through its own stream and is waited for: a synchronous
// hipMemcpy from pageable memory may return before its DMA has landed and
// is not ordered with a non-blocking stream.
static int h2d(mpci_prop *pr, void *dst, const void *src, size_t bytes) {
  if (bytes == 0) return MPCI_OK;
  HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, pr->stream));
  HIPCHK(hipStreamSynchronize(pr->stream));
  return MPCI_OK;
}

static int d2h(mpci_prop *pr, void *dst, const void *src, size_t bytes) {
  if (bytes == 0) return MPCI_OK;
  HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, pr->stream));
  HIPCHK(hipStreamSynchronize(pr->stream));
  return MPCI_OK;
}

// ---------------------------------------------------------------------
// pulses (host): src/tdse/pulse.cpp
// ---------------------------------------------------------------------
extern "C" const char *mpci_last_error(void) { return g_err.c_str(); }

extern "C" int mpci_device_count(int *n) {
  if (!n) return fail(MPCI_ERR_ARG, "n is NULL");
  int c = 0;
  hipError_t e = hipGetDeviceCount(&c);
  *n = (e == hipSuccess) ? c : 0;
  return MPCI_OK;
}

extern "C" void mpci_to_au(double Io_W, double w_eV, double *Io_au,
                           double *w_au) {
  *Io_au = Io_W * ((1. / 3.509338) * 1.e-16);
  *w_au = w_eV * (1. / 27.211396181);
}

extern "C" double mpci_period(double w, int cycles) {
  return cycles * 2.0 * kPi / w;
}

extern "C" int mpci_pulse_sine(double Io, double w, double cepd, int cycles,
                               char gauge, mpci_pulse *out) {
  if (!out || (gauge != 'v' && gauge != 'l') || cycles <= 0 || w == 0.0)
    return fail(MPCI_ERR_ARG, "mpci_pulse_sine: bad argument");
  std::memset(out, 0, sizeof(*out));
  out->kind = MPCI_PULSE_SINE2;
  out->Eo = (gauge == 'v') ? std::sqrt(Io) / w : std::sqrt(Io);
  out->w = w;
  out->cepd = cepd * (2.0 * kPi / w);
  out->a = w / (2.0 * cycles);
  return MPCI_OK;
}

extern "C" int mpci_pulse_gauss(double Io, double w, double tau, int cycles,
                                double cepd, char gauge, mpci_pulse *out) {
  if (!out || (gauge != 'v' && gauge != 'l') || w == 0.0)
    return fail(MPCI_ERR_ARG, "mpci_pulse_gauss: bad argument");
  constexpr double width = 0.58870501125773734549;
  std::memset(out, 0, sizeof(*out));
  out->kind = MPCI_PULSE_GAUSS;
  out->Eo = (gauge == 'v') ? std::sqrt(Io) / w : std::sqrt(Io);
  out->w = w;
  double To = 2 * kPi / w;
  int cycles_fwhm = 2 + int(cycles / 10);
  out->cepd = cepd;
  out->a = 0.5 * To * cycles_fwhm / width;
  out->b = tau;
  return MPCI_OK;
}

extern "C" double mpci_pulse_eval(const mpci_pulse *p, double t) {
  if (!p) return 0.0;
  if (p->kind == MPCI_PULSE_SINE2) {
    double sinWt = std::sin(p->a * t);
    return (double)(t >= 0.0 && t <= kPi / p->a) * p->Eo * sinWt * sinWt *
           std::sin(p->w * t + p->cepd);
  }
  if (p->kind == MPCI_PULSE_GAUSS) {
    double t_diff = t - 0.5 * p->b;
    return (double)(t >= 0.0 && t <= p->b) * p->Eo *
           std::exp(-t_diff * t_diff / (p->a * p->a)) *
           std::sin(p->w * t + p->cepd);
  }
  if (!(t >= 0.0 && t <= p->t_end) || p->n_env < 2 || !p->env) return 0.0;
  double dx = p->t_end / (p->n_env - 1);
  double u = t / dx;
  int k = (int)std::floor(u);
  if (k > p->n_env - 2) k = p->n_env - 2;
  if (k < 0) k = 0;
  double f = u - k;
  double en = (1.0 - f) * p->env[k] + f * p->env[k + 1];
  return p->scale * en * std::sin(p->w * t + p->cepd);
}

// ---------------------------------------------------------------------
// system
// ---------------------------------------------------------------------
static long long fill_desc(const mpci_system *sys, int P, int ks,
                           SysDesc &d);

static int build_system(mpci_system *sys, const double *const *E,
                        const double *const *D) {
  TRY(sys->E.alloc(sys->N));
  for (int L = 0; L < sys->nb; ++L)
    HIPCHK(hipMemcpy(sys->E.p + sys->off[L], E[L], sizeof(double) * sys->s[L],
                     hipMemcpyHostToDevice));
  const int npair = sys->nb - 1;
  sys->nrb.assign(std::max(npair, 0), 0);
  sys->ncb.assign(std::max(npair, 0), 0);
  std::vector<int> job0(std::max(npair, 0));
  sys->jobs.clear();
  for (int L = 0; L < npair; ++L) {
    sys->nrb[L] = (sys->s[L] + kGvTR - 1) / kGvTR;
    sys->ncb[L] = (sys->s[L + 1] + kGvTC - 1) / kGvTC;
    sys->S += (long long)sys->s[L] * sys->s[L + 1];
    job0[L] = (int)sys->jobs.size();
    // Column blocks in an order rotated by the row block: workgroup w
    // streams units w, w + G, ..., so without the rotation the narrow last
    // column block of every row block would fall to the workgroups with
    // w = ncb - 1 (mod ncb), all on one XCD when ncb is a multiple of 8.
    for (int rb = 0; rb < sys->nrb[L]; ++rb)
      for (int k = 0; k < sys->ncb[L]; ++k) {
        TileJob j;
        j.L = L;
        j.rb = rb;
        j.cb = (k + rb) % sys->ncb[L];
        j.pad = 0;
        j.doff = (long long)sys->jobs.size();
        sys->jobs.push_back(j);
      }
  }
  const int njobs = (int)sys->jobs.size();
  TRY(sys->Dg.alloc((size_t)njobs * kGvTileElems));
  for (int L = 0; L < npair; ++L) {
    const long long ne = (long long)sys->s[L] * sys->s[L + 1];
    DevBuf<double> tmp;
    TRY(tmp.alloc(ne));
    HIPCHK(hipMemcpy(tmp.p, D[L], sizeof(double) * ne, hipMemcpyHostToDevice));
    const long long nout = (long long)sys->nrb[L] * sys->ncb[L] * kGvTileElems;
    const int tpb = 256;
    const long long nblocks = (nout + tpb - 1) / tpb;
    hipLaunchKernelGGL(k_relayout_gemv, dim3((unsigned)nblocks), dim3(tpb), 0,
                       0, tmp.p, sys->s[L], sys->s[L + 1], sys->ncb[L], nout,
                       job0[L], njobs, sys->Dg.p);
    HIPCHK(hipGetLastError());
    HIPCHK(hipDeviceSynchronize());
  }
  {  // the one-pulse descriptor, for mpci_system_rhs_noalloc
    SysDesc d1;
    fill_desc(sys, 1, 1, d1);
    TRY(sys->ddesc1.alloc(1));
    HIPCHK(hipMemcpy(sys->ddesc1.p, &d1, sizeof(SysDesc),
                     hipMemcpyHostToDevice));
  }
  TRY(sys->djobs.alloc(sys->jobs.size()));
  if (!sys->jobs.empty())
    HIPCHK(hipMemcpy(sys->djobs.p, sys->jobs.data(),
                     sizeof(TileJob) * sys->jobs.size(),
                     hipMemcpyHostToDevice));
  // pageable copies may still be in flight: finish them before any stream
  // of a propagator reads the system
  HIPCHK(hipDeviceSynchronize());
  return MPCI_OK;
}

extern "C" int mpci_system_create(int nb, const int *s, char gauge,
                                  const double *const *E,
                                  const double *const *D, mpci_system **out) {
  if (!out) return fail(MPCI_ERR_ARG, "out is NULL");
  *out = nullptr;
  if (nb < 1 || nb > kMaxBlocks || !s || !E || (nb > 1 && !D))
    return fail(MPCI_ERR_ARG, "mpci_system_create: need 1 <= nb <= 32 and "
                              "non-NULL s, E, D");
  if (gauge != 'v' && gauge != 'l')
    return fail(MPCI_ERR_ARG, "gauge must be 'v' or 'l'");
  long long N = 0;
  for (int L = 0; L < nb; ++L) {
    if (s[L] < 1 || !E[L])
      return fail(MPCI_ERR_ARG, "block sizes must be >= 1 with energies");
    if (L + 1 < nb && !D[L]) return fail(MPCI_ERR_ARG, "NULL dipole block");
    N += s[L];
  }
  if (N > (1LL << 30)) return fail(MPCI_ERR_ARG, "state too large");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
    return fail(MPCI_ERR_HIP, "no HIP device visible: the tdse engine needs "
                              "an MI355X (gfx950) GPU");
  auto sys = new (std::nothrow) mpci_system();
  if (!sys) return fail(MPCI_ERR_HIP, "host allocation failed");
  sys->nb = nb;
  sys->N = (int)N;
  sys->gauge = (gauge == 'v') ? 0 : 1;
  sys->s.assign(s, s + nb);
  sys->off.assign(nb + 1, 0);
  for (int L = 0; L < nb; ++L) sys->off[L + 1] = sys->off[L] + s[L];
  (void)hipGetDevice(&sys->device);
  int ncu = 0;
  if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount,
                            sys->device) == hipSuccess && ncu > 0) {
    sys->gv_grid = 2 * ncu;
    sys->ncu = ncu;
  }
  int rc = build_system(sys, E, D);
  if (rc != MPCI_OK) {
    delete sys;
    return rc;
  }
  // keep the host blocks for a later batched propagator
  sys->hostD.resize(nb - 1);
  for (int L = 0; L + 1 < nb; ++L)
    sys->hostD[L].assign(D[L], D[L] + (size_t)s[L] * s[L + 1]);
  sys->hostE.resize(nb);
  for (int L = 0; L < nb; ++L) sys->hostE[L].assign(E[L], E[L] + s[L]);
  *out = sys;
  return MPCI_OK;
}

extern "C" int mpci_system_set_gemv_grid(mpci_system *sys, int workgroups) {
  if (!sys || workgroups < 0)
    return fail(MPCI_ERR_ARG, "mpci_system_set_gemv_grid: bad argument");
  if (workgroups == 0) {
    int ncu = 0;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount,
                              sys->device) != hipSuccess || ncu <= 0)
      ncu = 256;
    workgroups = 2 * ncu;
  }
  sys->gv_grid = workgroups;
  return MPCI_OK;
}

extern "C" int mpci_system_destroy(mpci_system *sys) {
  delete sys;
  return MPCI_OK;
}

extern "C" int mpci_system_info(const mpci_system *sys, int *N,
                                long long *bytes) {
  if (!sys) return fail(MPCI_ERR_ARG, "sys is NULL");
  if (N) *N = sys->N;
  if (bytes)
    *bytes = (long long)(sys->Dg.n + sys->E.n) * 8 +
             (long long)sys->djobs.n * sizeof(TileJob);
  return MPCI_OK;
}

// ---------------------------------------------------------------------
// propagator
// ---------------------------------------------------------------------
// Partial-sum layout: the GEMV form writes ncb(L) row partials and nrb(L)
// column partials per pair; the GEMM form writes `ks` K-split partials per
// pair and direction.
// Workgroups of the single-pulse stage kernel k_stage1: ceil(s_L / 16)
// per block, each workgroup inside one block.
static int stage1_blocks(const mpci_system *sys) {
  const int rpb = 256 / kStageSub1;
  int n = 0;
  for (int L = 0; L < sys->nb; ++L) n += (sys->s[L] + rpb - 1) / rpb;
  return n;
}

// Fills d for P pulses (K split ks in the batched form) and returns the
// number of double2 partials the layout needs.
// Partial-slot layout: pair L keeps nrow[L] partial vectors of D_L x_{L+1}
// (length s_L) and ncol[L] of D_L^T x_L (length s_{L+1}); returns the
// double2 count.
static long long fill_desc_counts(const mpci_system *sys, int P,
                                  const int *nrow, const int *ncol,
                                  SysDesc &d) {
  std::memset(&d, 0, sizeof(d));
  d.nb = sys->nb;
  d.N = sys->N;
  d.P = P;
  d.gauge = sys->gauge;
  for (int L = 0; L <= sys->nb; ++L) d.off[L] = sys->off[L];
  long long o = 0;
  for (int L = 0; L + 1 < sys->nb; ++L) {
    d.nrow_part[L] = nrow[L];
    d.ncol_part[L] = ncol[L];
    d.row_part_off[L] = o;
    o += (long long)d.nrow_part[L] * sys->s[L] * P;
    d.col_part_off[L] = o;
    o += (long long)d.ncol_part[L] * sys->s[L + 1] * P;
  }
  return o;
}

// One pulse: one slot per GEMV unit column / row; batched: ks K-split
// slots per direction.
static long long fill_desc(const mpci_system *sys, int P, int ks,
                           SysDesc &d) {
  std::vector<int> nr(std::max(sys->nb - 1, 1)), nc(nr.size());
  for (int L = 0; L + 1 < sys->nb; ++L) {
    nr[L] = P == 1 ? sys->ncb[L] : ks;
    nc[L] = P == 1 ? sys->nrb[L] : ks;
  }
  return fill_desc_counts(sys, P, nr.data(), nc.data(), d);
}

static int upload_desc(mpci_prop *pr, long long o) {
  TRY(pr->part.alloc(std::max<long long>(o, 1)));
  if (!pr->ddesc.p) TRY(pr->ddesc.alloc(1));
  TRY(h2d(pr, pr->ddesc.p, &pr->desc, sizeof(SysDesc)));
  return MPCI_OK;
}

static int init_desc(mpci_prop *pr, int ks) {
  return upload_desc(pr, fill_desc(pr->sys, pr->P, ks, pr->desc));
}

// k_gemm_pair's split of pair L: strips of kPrR rows (nS[L], one D^T x_L
// slot each) and column ranges of crange[L] columns (nC[L], one D x_{L+1}
// slot each), the ranges sized so that the jobs of all pairs and 32-column
// panels number about f x (compute units): one workgroup per CU holds the
// kernel's LDS.
static void pair_split(const mpci_prop *pr, int f, std::vector<int> &nS,
                       std::vector<int> &nC, std::vector<int> &crange) {
  const mpci_system *sys = pr->sys;
  const int np = std::max(sys->nb - 1, 1);
  nS.assign(np, 0);
  nC.assign(np, 0);
  crange.assign(np, 0);
  const int panels = (2 * pr->P + kPrTN - 1) / kPrTN;
  long long strips = 0;
  for (int L = 0; L + 1 < sys->nb; ++L) {
    nS[L] = (sys->s[L] + kPrR - 1) / kPrR;
    strips += nS[L];
  }
  const long long target = (long long)f * sys->ncu / std::max(panels, 1);
  const int nc = (int)std::max<long long>(1, target / std::max(strips, 1LL));
  for (int L = 0; L + 1 < sys->nb; ++L) {
    const int cols = sys->s[L + 1];
    const int cr = ((cols + nc - 1) / nc + kPrTK - 1) / kPrTK * kPrTK;
    crange[L] = cr;
    nC[L] = (cols + cr - 1) / cr;  // every range holds columns
  }
}

static int init_desc_pair(mpci_prop *pr, int f) {
  std::vector<int> nS, nC, cr;
  pair_split(pr, f, nS, nC, cr);
  return upload_desc(pr, fill_desc_counts(pr->sys, pr->P, nC.data(),
                                          nS.data(), pr->desc));
}

// the partial layout of candidate kind (3: k_gemm_pair, ks its f)
static int init_desc_kind(mpci_prop *pr, int kind, int ks) {
  return kind == 3 ? init_desc_pair(pr, ks) : init_desc(pr, ks);
}

// Row-major copy of every D_L with an even row stride ld_L (the padding
// column, when there is one, holds zeros), so the batched kernels can load
// 16-byte pairs from any even column.
static int ensure_rowmajor(mpci_system *sys) {
  if (sys->Dr.p || sys->nb < 2) return MPCI_OK;
  std::vector<long long> off(sys->nb - 1);
  sys->dr_ld.assign(sys->nb - 1, 0);
  long long tot = 0;
  for (int L = 0; L + 1 < sys->nb; ++L) {
    sys->dr_ld[L] = (sys->s[L + 1] + 1) & ~1;
    off[L] = tot;
    tot += (long long)sys->s[L] * sys->dr_ld[L];
  }
  TRY(sys->Dr.alloc(tot));
  TRY(sys->dr_off.alloc(off.size()));
  TRY(sys->dr_ld_d.alloc(off.size()));
  HIPCHK(hipMemcpy(sys->dr_ld_d.p, sys->dr_ld.data(),
                   sizeof(int) * off.size(), hipMemcpyHostToDevice));
  HIPCHK(hipMemset(sys->Dr.p, 0, sizeof(double) * tot));
  HIPCHK(hipMemcpy(sys->dr_off.p, off.data(), sizeof(long long) * off.size(),
                   hipMemcpyHostToDevice));
  for (int L = 0; L + 1 < sys->nb; ++L)
    HIPCHK(hipMemcpy2D(sys->Dr.p + off[L], sizeof(double) * sys->dr_ld[L],
                       sys->hostD[L].data(), sizeof(double) * sys->s[L + 1],
                       sizeof(double) * sys->s[L + 1], sys->s[L],
                       hipMemcpyHostToDevice));
  HIPCHK(hipDeviceSynchronize());
  return MPCI_OK;
}

// Row-major D_L^T of every pair (row stride s_L) for k_small_stage.
static int ensure_transposed(mpci_system *sys) {
  if (sys->DrT.p || sys->nb < 2) return MPCI_OK;
  std::vector<long long> off(sys->nb - 1);
  std::vector<int> ld(sys->nb - 1);
  long long tot = 0;
  for (int L = 0; L + 1 < sys->nb; ++L) {
    ld[L] = sys->s[L];
    off[L] = tot;
    tot += (long long)sys->s[L + 1] * sys->s[L];
  }
  std::vector<double> h((size_t)tot);
  for (int L = 0; L + 1 < sys->nb; ++L) {
    const int a = sys->s[L], b = sys->s[L + 1];
    const double *D = sys->hostD[L].data();
    double *T = h.data() + off[L];
    for (int i = 0; i < a; ++i)
      for (int j = 0; j < b; ++j) T[(size_t)j * a + i] = D[(size_t)i * b + j];
  }
  TRY(sys->DrT.alloc(tot));
  TRY(sys->drt_off.alloc(off.size()));
  TRY(sys->drt_ld_d.alloc(ld.size()));
  HIPCHK(hipMemcpy(sys->DrT.p, h.data(), sizeof(double) * tot,
                   hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(sys->drt_off.p, off.data(), sizeof(long long) * off.size(),
                   hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(sys->drt_ld_d.p, ld.data(), sizeof(int) * ld.size(),
                   hipMemcpyHostToDevice));
  return MPCI_OK;
}

// GEMM jobs: for every pair and direction, output panels of 16*fm*wm rows
// x 16*fn*wn real columns (wm x wn waves, each fm x fn MFMA blocks of
// 16 x 16), each over one of `ks` equal K ranges.

static std::vector<GemmJob> gemm_job_list(const mpci_prop *pr, int wm, int wn,
                                          int fm, int fn, int ks, int member) {
  const mpci_system *sys = pr->sys;
  const int TM = 16 * fm * wm, TN = 16 * fn * wn, P2 = 2 * pr->P;
  std::vector<GemmJob> out;
  for (int L = 0; L + 1 < sys->nb; ++L)
    for (int dir = 0; dir < 2; ++dir) {
      const int M = dir == 0 ? sys->s[L] : sys->s[L + 1];
      const int K = dir == 0 ? sys->s[L + 1] : sys->s[L];
      const int kchunk = ((K + ks - 1) / ks + kGmTK - 1) / kGmTK * kGmTK;
      for (int n0 = 0; n0 < P2; n0 += TN)
        for (int q = 0; q < ks; ++q)
          for (int m0 = 0; m0 < M; m0 += TM) {
            GemmJob j{};
            j.L = L;
            j.dir = dir;
            j.m0 = m0;
            j.n0 = n0;
            j.k0 = std::min(K, q * kchunk);
            j.k1 = std::min(K, (q + 1) * kchunk);
            j.q = q;
            j.ld = sys->dr_ld[L];
            j.member = member;
            out.push_back(j);
          }
    }
  return out;
}

// k_gemm_pair's jobs: per 32-column panel of X, pair, strip and column
// range (pair_split).
static std::vector<GemmJob> pair_job_list(const mpci_prop *pr, int f,
                                          int member) {
  const mpci_system *sys = pr->sys;
  std::vector<int> nS, nC, cr;
  pair_split(pr, f, nS, nC, cr);
  std::vector<GemmJob> out;
  for (int n0 = 0; n0 < 2 * pr->P; n0 += kPrTN)
    for (int L = 0; L + 1 < sys->nb; ++L)
      for (int st = 0; st < nS[L]; ++st)
        for (int ci = 0; ci < nC[L]; ++ci) {
          GemmJob j{};
          j.L = L;
          j.dir = 2;
          j.m0 = st * kPrR;
          j.n0 = n0;
          j.k0 = ci * cr[L];
          j.k1 = std::min(sys->s[L + 1], (ci + 1) * cr[L]);
          j.q = ci;
          j.qt = st;
          j.ld = sys->dr_ld[L];
          j.member = member;
          out.push_back(j);
        }
  return out;
}

static int build_gemm_jobs(mpci_prop *pr, int kind, int wm, int wn, int fm,
                           int fn, int tk, int ks) {
  pr->gjobs = kind == 3 ? pair_job_list(pr, ks, 0)
                        : gemm_job_list(pr, wm, wn, fm, fn, ks, 0);
  TRY(pr->dgjobs.alloc(pr->gjobs.size()));
  if (!pr->gjobs.empty())
    TRY(h2d(pr, pr->dgjobs.p, pr->gjobs.data(),
            sizeof(GemmJob) * pr->gjobs.size()));
  pr->gm_kind = kind;
  pr->gm_wm = wm;
  pr->gm_wn = wn;
  pr->gm_fm = fm;
  pr->gm_fn = fn;
  pr->gm_tk = tk;
  pr->gm_ks = ks;
  return MPCI_OK;
}

// e0 / e1 (may be null): HIP events attached to the kernel's own dispatch,
// so their interval is the kernel's execution alone.
static int launch_gemm(mpci_prop *pr, const double2 *xin,
                       hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr) {
  mpci_system *sys = pr->sys;
  const dim3 grid((unsigned)pr->gjobs.size());
  const double *X = reinterpret_cast<const double *>(xin);
  double *part = reinterpret_cast<double *>(pr->part.p);
#define MPCI_GEMM(WM, WN)                                                  \
  hipExtLaunchKernelGGL((k_gemm_coupling<WM, WN>), grid, dim3(64 * WM * WN), \
                        0, pr->stream, e0, e1, 0, pr->dgjobs.p, sys->Dr.p,  \
                        sys->dr_off.p, X, part, pr->ddesc.p, pr->ctl.p)
#define MPCI_GEMM_DB(WM, WN, TK, FM, FN)                                    \
  hipExtLaunchKernelGGL((k_gemm_db<WM, WN, TK, FM, FN>), grid,              \
                        dim3(64 * WM * WN), 0, pr->stream, e0, e1, 0,       \
                        pr->dgjobs.p, sys->Dr.p, sys->dr_off.p, X, part,    \
                        pr->ddesc.p, pr->ctl.p)
#define MPCI_GEMM_LDS(WM, WN)                                                \
  hipExtLaunchKernelGGL((k_gemm_lds<WM, WN>), grid, dim3(64 * WM * WN), 0,   \
                        pr->stream, e0, e1, 0, pr->dgjobs.p, sys->Dr.p,      \
                        sys->dr_off.p, X, part, pr->ddesc.p, pr->ctl.p)
  const int wm = pr->gm_wm, wn = pr->gm_wn, tk = pr->gm_tk;
  const int f = pr->gm_fm * 10 + pr->gm_fn;
  if (pr->gm_kind == 3) {
    hipExtLaunchKernelGGL(k_gemm_pair, grid, dim3(1024), 0, pr->stream, e0, e1,
                          0, pr->dgjobs.p, sys->Dr.p, sys->dr_off.p, X, part,
                          pr->ddesc.p, pr->ctl.p);
  } else if (pr->gm_kind == 2) {
    if (wm == 2 && wn == 1)
      MPCI_GEMM_LDS(2, 1);
    else if (wm == 4 && wn == 1)
      MPCI_GEMM_LDS(4, 1);
    else if (wm == 2 && wn == 2)
      MPCI_GEMM_LDS(2, 2);
    else if (wm == 4 && wn == 2)
      MPCI_GEMM_LDS(4, 2);
    else if (wm == 2 && wn == 4)
      MPCI_GEMM_LDS(2, 4);
    else
      MPCI_GEMM_LDS(4, 4);
  } else if (pr->gm_kind == 0) {
    if (wm == 2 && wn == 2)
      MPCI_GEMM(2, 2);
    else if (wm == 2 && wn == 4)
      MPCI_GEMM(2, 4);
    else if (wm == 4 && wn == 2)
      MPCI_GEMM(4, 2);
    else
      MPCI_GEMM(4, 4);
  } else if (f == 44) {
    if (tk == 32)
      MPCI_GEMM_DB(2, 2, 32, 4, 4);
    else if (wm == 2 && wn == 2)
      MPCI_GEMM_DB(2, 2, 16, 4, 4);
    else if (wm == 1 && wn == 2)
      MPCI_GEMM_DB(1, 2, 16, 4, 4);
    else if (wm == 2 && wn == 1)
      MPCI_GEMM_DB(2, 1, 16, 4, 4);
    else
      MPCI_GEMM_DB(1, 1, 16, 4, 4);
  } else if (f == 24) {
    if (wm == 4 && wn == 1)
      MPCI_GEMM_DB(4, 1, 16, 2, 4);
    else
      MPCI_GEMM_DB(2, 2, 16, 2, 4);
  } else if (f == 42) {
    MPCI_GEMM_DB(2, 2, 16, 4, 2);
  } else if (tk == 32) {
    if (wm == 2 && wn == 4)
      MPCI_GEMM_DB(2, 4, 32, 2, 2);
    else
      MPCI_GEMM_DB(4, 2, 32, 2, 2);
  } else {
    if (wm == 2 && wn == 1)
      MPCI_GEMM_DB(2, 1, 16, 2, 2);
    else if (wm == 4 && wn == 1)
      MPCI_GEMM_DB(4, 1, 16, 2, 2);
    else if (wm == 2 && wn == 2)
      MPCI_GEMM_DB(2, 2, 16, 2, 2);
    else if (wm == 4 && wn == 2)
      MPCI_GEMM_DB(4, 2, 16, 2, 2);
    else if (wm == 2 && wn == 4)
      MPCI_GEMM_DB(2, 4, 16, 2, 2);
    else
      MPCI_GEMM_DB(4, 4, 16, 2, 2);
  }
#undef MPCI_GEMM
#undef MPCI_GEMM_DB
#undef MPCI_GEMM_LDS
  HIPCHK(hipGetLastError());
  return MPCI_OK;
}

// Tile candidates of the batched form.  kind 0: k_gemm_coupling (single LDS
// buffer, K chunks of 32, 2 x 2 blocks per wave); kind 1: k_gemm_db
// (double-buffered, 16-byte loads, K chunks of tk, fm x fn blocks per wave);
// kind 2: k_gemm_lds (double-buffered direct-to-LDS loads, K chunks of 16,
// 2 x 2 blocks per wave); kind 3: k_gemm_pair (both directions from one
// read of each dipole block, 16 waves; ks is the number of workgroups per
// compute unit its split aims at).
struct GemmCand {
  int kind, wm, wn, fm, fn, tk, ks;
};
static const GemmCand kGemmCands[] = {
    {0, 2, 2, 2, 2, 32, 1}, {0, 2, 2, 2, 2, 32, 2}, {0, 2, 4, 2, 2, 32, 1},
    {0, 2, 4, 2, 2, 32, 2}, {0, 4, 2, 2, 2, 32, 1}, {0, 4, 2, 2, 2, 32, 2},
    {0, 4, 4, 2, 2, 32, 1}, {0, 4, 4, 2, 2, 32, 2}, {0, 4, 4, 2, 2, 32, 4},
    {0, 2, 4, 2, 2, 32, 4}, {1, 2, 1, 2, 2, 16, 1}, {1, 2, 1, 2, 2, 16, 2},
    {1, 4, 1, 2, 2, 16, 1}, {1, 4, 1, 2, 2, 16, 2}, {1, 2, 2, 2, 2, 16, 1},
    {1, 2, 2, 2, 2, 16, 2}, {1, 4, 2, 2, 2, 16, 1}, {1, 4, 2, 2, 2, 16, 2},
    {1, 2, 4, 2, 2, 16, 1}, {1, 2, 4, 2, 2, 16, 2}, {1, 2, 4, 2, 2, 16, 4},
    {1, 4, 4, 2, 2, 16, 1}, {1, 4, 4, 2, 2, 16, 2}, {1, 2, 4, 2, 2, 32, 1},
    {1, 2, 4, 2, 2, 32, 2}, {1, 4, 2, 2, 2, 32, 1}, {1, 4, 2, 2, 2, 32, 2},
    {1, 4, 2, 2, 2, 16, 4}, {1, 2, 2, 4, 4, 16, 1}, {1, 2, 2, 4, 4, 16, 2},
    {1, 2, 2, 4, 4, 16, 4}, {1, 1, 2, 4, 4, 16, 1}, {1, 1, 2, 4, 4, 16, 2},
    {1, 1, 2, 4, 4, 16, 4}, {1, 2, 1, 4, 4, 16, 2}, {1, 2, 1, 4, 4, 16, 4},
    {1, 1, 1, 4, 4, 16, 2}, {1, 1, 1, 4, 4, 16, 4}, {1, 4, 1, 2, 4, 16, 2},
    {1, 4, 1, 2, 4, 16, 4}, {1, 2, 2, 2, 4, 16, 2}, {1, 2, 2, 2, 4, 16, 4},
    {1, 2, 2, 4, 2, 16, 2}, {1, 2, 2, 4, 2, 16, 4}, {1, 2, 2, 4, 4, 32, 2},
    {1, 2, 2, 4, 4, 32, 4}, {2, 2, 1, 2, 2, 16, 1}, {2, 2, 1, 2, 2, 16, 2},
    {2, 2, 1, 2, 2, 16, 4}, {2, 4, 1, 2, 2, 16, 1}, {2, 4, 1, 2, 2, 16, 2},
    {2, 4, 1, 2, 2, 16, 4}, {2, 2, 2, 2, 2, 16, 1}, {2, 2, 2, 2, 2, 16, 2},
    {2, 2, 2, 2, 2, 16, 4}, {2, 4, 2, 2, 2, 16, 1}, {2, 4, 2, 2, 2, 16, 2},
    {2, 4, 2, 2, 2, 16, 4}, {2, 2, 4, 2, 2, 16, 1}, {2, 2, 4, 2, 2, 16, 2},
    {2, 2, 4, 2, 2, 16, 4}, {2, 4, 4, 2, 2, 16, 1}, {2, 4, 4, 2, 2, 16, 2},
    {2, 4, 4, 2, 2, 16, 4}, {3, 16, 1, 2, 2, 16, 1}, {3, 16, 1, 2, 2, 16, 2}};

// k_gemm_lds addresses its operands through 32-bit buffer offsets: every
// D_L copy and the X slab of the propagator must stay below 2 GiB.
static bool gemm_lds_fits(const mpci_prop *pr) {
  const mpci_system *sys = pr->sys;
  const long long lim = 1LL << 31;
  if ((long long)sys->N * 2 * pr->P * 8 >= lim) return false;
  for (int L = 0; L + 1 < sys->nb; ++L)
    if ((long long)sys->s[L] * sys->dr_ld[L] * 8 >= lim) return false;
  return true;
}

static int launch_stage(mpci_prop *pr, const StageArgs &a);
static StageArgs stage_args(mpci_prop *pr, int s);

// Sets up candidate cd and times the coupling evaluations of a step as the
// step runs them: stages 0 .. 11 in order, each the GEMM (on the state at
// stage 0, on the stage input the previous stage kernel wrote after that)
// and then the stage kernel, which reads the candidate's partial slots
// back (their number differs from candidate to candidate).  Between HIP
// events on the propagator's stream, all pulses live; the stage kernels
// write only k_s and the stage input buffer, so the state is untouched.
// (Timing one stage over and over, on the state, ranked k_gemm_pair
// behind the best K-split tile at 16 pulses on most devices, while whole
// attempts run it 2-5 % faster.)
static int time_gemm_cand(mpci_prop *pr, const GemmCand &cd, hipEvent_t e0,
                          hipEvent_t e1, float *ms) {
  TRY(init_desc_kind(pr, cd.kind, cd.ks));
  TRY(build_gemm_jobs(pr, cd.kind, cd.wm, cd.wn, cd.fm, cd.fn, cd.tk, cd.ks));
  auto eval = [&](int s) {
    TRY(launch_gemm(pr, s == 0 ? pr->X.p : pr->Xs.p));
    StageArgs a = stage_args(pr, s);
    a.xs_store = nullptr;  // leave any adjoint stage store alone
    return launch_stage(pr, a);
  };
  // two warm evaluations, then the twelve stages timed
  constexpr int kReps = kStages - 1;
  for (int s = 0; s < 2; ++s) TRY(eval(s));
  HIPCHK(hipEventRecord(e0, pr->stream));
  for (int s = 0; s < kReps; ++s) TRY(eval(s));
  HIPCHK(hipEventRecord(e1, pr->stream));
  HIPCHK(hipEventSynchronize(e1));
  HIPCHK(hipEventElapsedTime(ms, e0, e1));
  *ms /= (float)kReps;
  return MPCI_OK;
}

static int set_all_live(mpci_prop *pr, int done) {
  std::vector<Ctl> c(pr->P);
  std::memset(c.data(), 0, sizeof(Ctl) * pr->P);
  for (auto &x : c) x.done = done;
  return h2d(pr, pr->ctl.p, c.data(), sizeof(Ctl) * pr->P);
}

// Times every candidate of kGemmCands (or only `only`, when given), sets up
// the fastest (GEMM plus the stage kernel's reading of its partial sums)
// and, for a full search, records it in the system's tune cache for this
// pulse count.
static int tune_gemm(mpci_prop *pr, const GemmCand *only) {
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  int rc = set_all_live(pr, 0);
  float best = 1e30f;
  GemmCand bc = only ? *only : kGemmCands[0];
  if (rc == MPCI_OK) {
    if (only) {
      rc = time_gemm_cand(pr, *only, e0, e1, &best);
    } else {
      const bool lds_ok = gemm_lds_fits(pr);
      for (const GemmCand &cd : kGemmCands) {
        if (cd.kind >= 2 && !lds_ok) continue;
        float ms = 0.f;
        rc = time_gemm_cand(pr, cd, e0, e1, &ms);
        if (rc != MPCI_OK) break;
        if (ms < best) {
          best = ms;
          bc = cd;
        }
      }
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  TRY(rc);
  TRY(set_all_live(pr, 1));
  TRY(init_desc_kind(pr, bc.kind, bc.ks));
  TRY(build_gemm_jobs(pr, bc.kind, bc.wm, bc.wn, bc.fm, bc.fn, bc.tk, bc.ks));
  pr->gemm_ms = best;
  if (!only)
    pr->sys->tuned[pr->P] = {bc.kind, bc.wm, bc.wn, bc.fm,  bc.fn,
                             bc.tk,   bc.ks, pr->gemm_ms};
  return MPCI_OK;
}

// Autotuner of the batched form: times every candidate tile on this system
// and pulse count and keeps the fastest; later propagators of the same
// system and pulse count reuse the choice.
static int autotune_gemm(mpci_prop *pr) {
  auto it = pr->sys->tuned.find(pr->P);
  if (it != pr->sys->tuned.end()) {
    const auto &t = it->second;
    TRY(init_desc_kind(pr, t.kind, t.ks));
    TRY(build_gemm_jobs(pr, t.kind, t.wm, t.wn, t.fm, t.fn, t.tk, t.ks));
    pr->gemm_ms = t.ms;
    return MPCI_OK;
  }
  // timed on a state of nonzero components (the bit pattern 0x3FF00000 in
  // every 32-bit word: 1.0000002), then cleared again
  HIPCHK(hipMemsetD32Async(pr->X.p, (int)0x3FF00000u, (size_t)pr->NP * 4,
                           pr->stream));
  const int rc = tune_gemm(pr, nullptr);
  HIPCHK(hipMemsetAsync(pr->X.p, 0, sizeof(double2) * pr->NP, pr->stream));
  HIPCHK(hipStreamSynchronize(pr->stream));
  return rc;
}

// Dipole bytes (8 S) up to which one pulse takes the small-system path by
// default: there every stage is a handful of microseconds of launch and
// latency, and the fused kernel halves the launches per attempt.  Measured
// on four equal blocks: 1.95x the tile path at 0.9 MB of dipoles, 1.30x at
// 8.6 MB, even at 15 MB, 0.85x at 24 MB.
constexpr long long kSmallPathBytes = 12LL << 20;

// mode -1: automatic (one pulse, more than one block, 8 S within
// kSmallPathBytes); 0: off; 1: on wherever it applies (one pulse, more
// than one block).
static int apply_small_path(mpci_prop *pr, int mode) {
  mpci_system *sys = pr->sys;
  bool on = false;
  if (pr->P == 1 && sys->nb > 1)
    on = mode == 1 || (mode < 0 && 8 * sys->S <= kSmallPathBytes);
  if (on) {
    TRY(ensure_rowmajor(sys));
    TRY(ensure_transposed(sys));
    if (!pr->Xs2.p) TRY(pr->Xs2.alloc(pr->NP));
  }
  pr->small = on;
  return MPCI_OK;
}

extern "C" int mpci_prop_create(mpci_system *sys, int P, mpci_prop **out) {
  if (!out || !sys) return fail(MPCI_ERR_ARG, "NULL argument");
  *out = nullptr;
  if (P < 1 || P > 256)
    return fail(MPCI_ERR_ARG, "P must be in [1, 256]");
  std::unique_ptr<mpci_prop> pr(new (std::nothrow) mpci_prop());
  if (!pr) return fail(MPCI_ERR_HIP, "host allocation failed");
  pr->sys = sys;
  pr->P = P;
  HIPCHK(hipSetDevice(sys->device));
  HIPCHK(hipStreamCreateWithFlags(&pr->stream, hipStreamNonBlocking));
  pr->NP = (long long)sys->N * P;
  const int rpb = 256 / P;
  pr->nblk = (sys->N + rpb - 1) / rpb;
  pr->nblk_stage = P == 1 ? stage1_blocks(sys) : pr->nblk;
  TRY(init_desc(pr.get(), 1));
  TRY(pr->X.alloc(pr->NP));
  TRY(pr->Xs.alloc(pr->NP));
  TRY(pr->Y.alloc(pr->NP));
  TRY(pr->K.alloc(pr->NP * kStages));
  pr->nblk_small = (sys->N + 3) / 4;
  TRY(pr->red.alloc(
      (size_t)std::max({pr->nblk, pr->nblk_stage, pr->nblk_small}) * P * 2));
  TRY(pr->ctl.alloc(P));
  TRY(pr->pulses.alloc(P));
  TRY(pr->log.alloc((size_t)pr->logcap * P));
  HIPCHK(hipMemsetAsync(pr->X.p, 0, sizeof(double2) * pr->NP, pr->stream));
  HIPCHK(hipMemsetAsync(pr->part.p, 0, sizeof(double2) * pr->part.n,
                        pr->stream));
  HIPCHK(hipStreamSynchronize(pr->stream));
  std::vector<Ctl> c(P);
  std::memset(c.data(), 0, sizeof(Ctl) * P);
  for (auto &x : c) x.done = 1;
  TRY(h2d(pr.get(), pr->ctl.p, c.data(), sizeof(Ctl) * P));
  pr->hpulses.assign(P, PulseDev{});
  pr->henv.assign(P, {});
  if (P > 1 && sys->nb > 1) {
    TRY(ensure_rowmajor(sys));
    TRY(autotune_gemm(pr.get()));
  }
  TRY(apply_small_path(pr.get(), -1));
  *out = pr.release();
  return MPCI_OK;
}

extern "C" int mpci_prop_destroy(mpci_prop *pr) {
  if (pr) {
    (void)hipStreamSynchronize(pr->stream);
    delete pr;
  }
  return MPCI_OK;
}

extern "C" int mpci_prop_set_state(mpci_prop *pr, int p, const double *ct) {
  if (!pr || !ct || p < 0 || p >= pr->P)
    return fail(MPCI_ERR_ARG, "mpci_prop_set_state: bad argument");
  if (pr->P == 1) {
    return h2d(pr, pr->X.p, ct, sizeof(double2) * pr->sys->N);
  }
  // one contiguous copy, then a strided kernel copy into the batch layout
  // (a 2-D copy of N 16-byte rows from host memory moves row by row)
  const int N = pr->sys->N;
  if (!pr->xstage.p) TRY(pr->xstage.alloc(N));
  HIPCHK(hipMemcpyAsync(pr->xstage.p, ct, sizeof(double2) * N,
                        hipMemcpyHostToDevice, pr->stream));
  hipLaunchKernelGGL(k_copy_strided, dim3((N + 255) / 256), dim3(256), 0,
                     pr->stream, pr->X.p + p, pr->P, pr->xstage.p, 1, N);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(pr->stream));
  return MPCI_OK;
}

extern "C" int mpci_prop_get_state(mpci_prop *pr, int p, double *ct) {
  if (!pr || !ct || p < 0 || p >= pr->P)
    return fail(MPCI_ERR_ARG, "mpci_prop_get_state: bad argument");
  HIPCHK(hipStreamSynchronize(pr->stream));
  if (pr->P == 1) {
    return d2h(pr, ct, pr->X.p, sizeof(double2) * pr->sys->N);
  }
  const int N = pr->sys->N;
  if (!pr->xstage.p) TRY(pr->xstage.alloc(N));
  hipLaunchKernelGGL(k_copy_strided, dim3((N + 255) / 256), dim3(256), 0,
                     pr->stream, pr->xstage.p, 1, pr->X.p + p, pr->P, N);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(ct, pr->xstage.p, sizeof(double2) * N,
                        hipMemcpyDeviceToHost, pr->stream));
  HIPCHK(hipStreamSynchronize(pr->stream));
  return MPCI_OK;
}

static void drop_graph(mpci_prop *pr);

// The envelope store only grows.  A captured step holds its address in the
// stage kernels' arguments, so when it has to move, the graph is dropped
// (after the stream has drained) and recaptured at the next run; samples
// rewritten in place need no recapture.
static int upload_pulses(mpci_prop *pr) {
  size_t tot = 0;
  for (int p = 0; p < pr->P; ++p) {
    pr->hpulses[p].env_off = (long long)tot;
    tot += pr->henv[p].size();
  }
  if (tot > pr->env.n) {
    HIPCHK(hipStreamSynchronize(pr->stream));
    drop_graph(pr);
    TRY(pr->env.alloc(tot));
  }
  if (tot > 0) {
    std::vector<double> flat;
    flat.reserve(tot);
    for (auto &v : pr->henv) flat.insert(flat.end(), v.begin(), v.end());
    TRY(h2d(pr, pr->env.p, flat.data(), sizeof(double) * tot));
  }
  TRY(h2d(pr, pr->pulses.p, pr->hpulses.data(), sizeof(PulseDev) * pr->P));
  return MPCI_OK;
}

extern "C" int mpci_prop_set_pulse(mpci_prop *pr, int p,
                                   const mpci_pulse *pl) {
  if (!pr || !pl || p < 0 || p >= pr->P)
    return fail(MPCI_ERR_ARG, "mpci_prop_set_pulse: bad argument");
  if (pl->kind < MPCI_PULSE_SINE2 || pl->kind > MPCI_PULSE_PWL)
    return fail(MPCI_ERR_ARG, "unknown pulse kind");
  if (pl->kind == MPCI_PULSE_PWL && (pl->n_env < 2 || !pl->env))
    return fail(MPCI_ERR_ARG, "PWL pulse needs n_env >= 2 samples");
  PulseDev d{};
  d.kind = pl->kind;
  d.Eo = pl->Eo;
  d.w = pl->w;
  d.cepd = pl->cepd;
  d.a = pl->a;
  d.b = pl->b;
  d.t_end = pl->t_end;
  d.scale = pl->scale;
  d.n_env = (pl->kind == MPCI_PULSE_PWL) ? pl->n_env : 0;
  pr->hpulses[p] = d;
  pr->henv[p].assign(pl->env, pl->env + d.n_env);
  return upload_pulses(pr);
}

// ---- one attempt ------------------------------------------------------
static int enqueue_tile(mpci_prop *pr, const double2 *xin, int s) {
  mpci_system *sys = pr->sys;
  if (sys->jobs.empty()) return MPCI_OK;
  // instrumented: the event pair rides on the kernel's dispatch packet
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (pr->instrument) {
    e0 = pr->ev[pr->ev_base + 2 * s];
    e1 = pr->ev[pr->ev_base + 2 * s + 1];
  }
  if (pr->P == 1) {
    const int njobs = (int)sys->jobs.size();
    hipExtLaunchKernelGGL(k_gemv_tile,
                          dim3((unsigned)std::min(njobs, sys->gv_grid)),
                          dim3(kGvThreads), 0, pr->stream, e0, e1, 0,
                          sys->djobs.p, njobs, sys->Dg.p, xin, pr->part.p,
                          pr->ddesc.p, pr->ctl.p,
                          s & 1);  // stages alternate sweep directions
    HIPCHK(hipGetLastError());
  } else {
    TRY(launch_gemm(pr, xin, e0, e1));
  }
  return MPCI_OK;
}

static int launch_stage(mpci_prop *pr, const StageArgs &a) {
  const dim3 grid(pr->nblk_stage), block(256);
  const bool last = a.s == kStages - 1;
  if (pr->P == 1) {
    if (last)
      hipLaunchKernelGGL((k_stage1<true>), grid, block, 0, pr->stream, a,
                         pr->ddesc.p);
    else
      hipLaunchKernelGGL((k_stage1<false>), grid, block, 0, pr->stream, a,
                         pr->ddesc.p);
  } else if (last) {
    hipLaunchKernelGGL((k_stage<1, true>), grid, block, 0, pr->stream, a,
                       pr->ddesc.p);
  } else {
    hipLaunchKernelGGL((k_stage<1, false>), grid, block, 0, pr->stream, a,
                       pr->ddesc.p);
  }
  HIPCHK(hipGetLastError());
  return MPCI_OK;
}

static StageArgs stage_args(mpci_prop *pr, int s) {
  StageArgs a{};
  a.X = pr->X.p;
  a.Xs = pr->Xs.p;
  a.Xs_next = pr->Xs.p;
  if (pr->small) {
    // stage s reads its input from one buffer while it writes stage s + 1's
    // into the other (stage 0 reads X)
    double2 *buf[2] = {pr->Xs.p, pr->Xs2.p};
    a.Xs = buf[s & 1];
    a.Xs_next = buf[(s + 1) & 1];
  }
  a.K = pr->K.p;
  a.Y = pr->Y.p;
  a.part = pr->part.p;
  a.E = pr->sys->E.p;
  a.red = pr->red.p;
  a.ctl = pr->ctl.p;
  a.xs_store = pr->store.xs;
  a.NP = pr->NP;
  a.s = s;
  return a;
}

// Stage s on the small-system path (the instrumented event pair rides on
// the fused kernel's dispatch).
static int launch_small_stage(mpci_prop *pr, const StageArgs &a, int s) {
  mpci_system *sys = pr->sys;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (pr->instrument) {
    e0 = pr->ev[pr->ev_base + 2 * s];
    e1 = pr->ev[pr->ev_base + 2 * s + 1];
  }
  const dim3 grid(pr->nblk_small), block(256);
  if (a.s == kStages - 1)
    hipExtLaunchKernelGGL((k_small_stage<true>), grid, block, 0, pr->stream,
                          e0, e1, 0, a, pr->ddesc.p, sys->Dr.p, sys->dr_off.p,
                          sys->dr_ld_d.p, sys->DrT.p, sys->drt_off.p,
                          sys->drt_ld_d.p);
  else
    hipExtLaunchKernelGGL((k_small_stage<false>), grid, block, 0, pr->stream,
                          e0, e1, 0, a, pr->ddesc.p, sys->Dr.p, sys->dr_off.p,
                          sys->dr_ld_d.p, sys->DrT.p, sys->drt_off.p,
                          sys->drt_ld_d.p);
  HIPCHK(hipGetLastError());
  return MPCI_OK;
}

static int enqueue_attempt(mpci_prop *pr) {
  for (int s = 0; s < kStages; ++s) {
    if (pr->small) {
      TRY(launch_small_stage(pr, stage_args(pr, s), s));
    } else {
      TRY(enqueue_tile(pr, s == 0 ? pr->X.p : pr->Xs.p, s));
      TRY(launch_stage(pr, stage_args(pr, s)));
    }
    HIPCHK(hipGetLastError());
  }
  hipLaunchKernelGGL(k_finalize, dim3(pr->P), dim3(256), 0, pr->stream,
                     pr->ctl.p, pr->red.p,
                     pr->small ? pr->nblk_small : pr->nblk_stage, pr->P,
                     pr->hs.p, pr->Ftab.p, pr->pulses.p, pr->env.p);
  HIPCHK(hipGetLastError());
  hipLaunchKernelGGL(k_apply, dim3(pr->nblk), dim3(256), 0, pr->stream,
                     pr->X.p, pr->Y.p, pr->ctl.p, pr->log.p, pr->logcap,
                     pr->track, pr->sys->N, pr->P, pr->store);
  HIPCHK(hipGetLastError());
  return MPCI_OK;
}

static void drop_graph(mpci_prop *pr) {
  if (pr->gexec) (void)hipGraphExecDestroy(pr->gexec);
  if (pr->graph) (void)hipGraphDestroy(pr->graph);
  pr->gexec = nullptr;
  pr->graph = nullptr;
}

// Kernel arguments that change between runs (Ftab, hs pointers, track) are
// baked into the graph, so it is rebuilt when they change.
static int ensure_graph(mpci_prop *pr) {
  if (pr->gexec && pr->graph_track == pr->track) return MPCI_OK;
  drop_graph(pr);
  HIPCHK(hipStreamBeginCapture(pr->stream, hipStreamCaptureModeThreadLocal));
  int rc = enqueue_attempt(pr);
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(pr->stream, &g);
  if (rc != MPCI_OK) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (e != hipSuccess)
    return fail(MPCI_ERR_HIP, std::string("hipStreamEndCapture: ") +
                                  hipGetErrorString(e));
  pr->graph = g;
  HIPCHK(hipGraphInstantiate(&pr->gexec, g, nullptr, nullptr, 0));
  pr->graph_track = pr->track;
  return MPCI_OK;
}

static int read_ctl(mpci_prop *pr, std::vector<Ctl> &c);

static int max_attempts_done(const std::vector<Ctl> &c) {
  int m = 0;
  for (auto &x : c) m = std::max(m, x.attempts);
  return m;
}

// Instrumented attempts: launched individually with a HIP event pair around
// every coupling kernel on the propagator's stream.  The host synchronises
// once per kInstrChunk attempts; only attempts that did work (the device
// attempt counter advanced) are added to the kernel time.
static int run_attempts_instrumented(mpci_prop *pr, int n) {
  constexpr int kInstrChunk = 64;
  if (pr->ev.empty()) {
    pr->ev.resize(2 * kStages * kInstrChunk);
    for (auto &e : pr->ev) HIPCHK(hipEventCreate(&e));
  }
  std::vector<Ctl> c;
  int launched = 0;
  while (launched < n) {
    const int m = std::min(kInstrChunk, n - launched);
    TRY(read_ctl(pr, c));
    const int before = max_attempts_done(c);
    for (int k = 0; k < m; ++k) {
      pr->ev_base = 2 * kStages * k;
      TRY(enqueue_attempt(pr));
    }
    TRY(read_ctl(pr, c));
    const int real = max_attempts_done(c) - before;
    if (!pr->sys->jobs.empty())
      for (int k = 0; k < real && k < m; ++k)
        for (int s = 0; s < kStages; ++s) {
          const int b = 2 * kStages * k + 2 * s;
          float ms = 0.f;
          HIPCHK(hipEventElapsedTime(&ms, pr->ev[b], pr->ev[b + 1]));
          pr->kernel_ms += ms;
          pr->kernel_launches++;
        }
    launched += m;
    if (real < m) break;  // every pulse finished inside this chunk
  }
  pr->ev_base = 0;
  return MPCI_OK;
}

static int run_attempts(mpci_prop *pr, int n) {
  if (pr->instrument) return run_attempts_instrumented(pr, n);
  if (pr->use_graph) {
    TRY(ensure_graph(pr));
    for (int i = 0; i < n; ++i) HIPCHK(hipGraphLaunch(pr->gexec, pr->stream));
  } else {
    for (int i = 0; i < n; ++i) TRY(enqueue_attempt(pr));
  }
  return MPCI_OK;
}

static int read_ctl(mpci_prop *pr, std::vector<Ctl> &c) {
  c.resize(pr->P);
  HIPCHK(hipStreamSynchronize(pr->stream));
  TRY(d2h(pr, c.data(), pr->ctl.p, sizeof(Ctl) * pr->P));
  return MPCI_OK;
}

static int write_ctl(mpci_prop *pr, const std::vector<Ctl> &c) {
  TRY(h2d(pr, pr->ctl.p, c.data(), sizeof(Ctl) * pr->P));
  return MPCI_OK;
}

// Writes the Ctl records a run starts from, then sets their first
// attempt's stage fields on the device (k_finalize sets the later ones).
static int start_ctl(mpci_prop *pr, const std::vector<Ctl> &c) {
  TRY(write_ctl(pr, c));
  hipLaunchKernelGGL(k_ctl_fields, dim3(pr->P), dim3(64), 0, pr->stream,
                     pr->ctl.p, pr->P, pr->Ftab.p, pr->pulses.p, pr->env.p);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(pr->stream));
  return MPCI_OK;
}

// Drain the per-chunk log into the caller's arrays and reset the counters.
static int drain_log(mpci_prop *pr, std::vector<Ctl> &c,
                     std::vector<std::vector<LogEnt>> &acc) {
  int maxn = 0;
  for (auto &x : c) maxn = std::max(maxn, x.nlog);
  if (maxn > pr->logcap)
    return fail(MPCI_ERR_STATE, "internal: log overflow");
  if (maxn > 0) {
    std::vector<LogEnt> h((size_t)pr->logcap * pr->P);
    TRY(d2h(pr, h.data(), pr->log.p, sizeof(LogEnt) * h.size()));
    for (int p = 0; p < pr->P; ++p)
      acc[p].insert(acc[p].end(), h.begin() + (size_t)p * pr->logcap,
                    h.begin() + (size_t)p * pr->logcap + c[p].nlog);
  }
  for (auto &x : c) x.nlog = 0;
  return write_ctl(pr, c);
}

extern "C" int mpci_prop_rhs(mpci_prop *pr, const double *F, const double *x,
                             double *dxdt) {
  if (!pr || !F || !x || !dxdt) return fail(MPCI_ERR_ARG, "NULL argument");
  for (int p = 0; p < pr->P; ++p)
    TRY(mpci_prop_set_state(pr, p, x + (size_t)p * 2 * pr->sys->N));
  std::vector<Ctl> c(pr->P);
  std::memset(c.data(), 0, sizeof(Ctl) * pr->P);
  for (int p = 0; p < pr->P; ++p) {
    c[p].mode = kModeRhs;
    c[p].fmode = kFieldConst;
    c[p].Fc = F[p];
  }
  TRY(start_ctl(pr, c));
  StageArgs a = stage_args(pr, 0);
  a.red = nullptr;
  if (pr->small) {
    TRY(launch_small_stage(pr, a, 0));
  } else {
    TRY(enqueue_tile(pr, pr->X.p, 0));
    TRY(launch_stage(pr, a));
  }
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(pr->stream));
  std::vector<double2> k(pr->NP);
  TRY(d2h(pr, k.data(), pr->K.p, sizeof(double2) * pr->NP));
  const int N = pr->sys->N;
  for (int p = 0; p < pr->P; ++p)
    for (int i = 0; i < N; ++i) {
      dxdt[(size_t)p * 2 * N + 2 * i] = k[(size_t)i * pr->P + p].x;
      dxdt[(size_t)p * 2 * N + 2 * i + 1] = k[(size_t)i * pr->P + p].y;
    }
  for (auto &x : c) x.done = 1;
  return write_ctl(pr, c);
}

// ---------------------------------------------------------------------
// One RHS evaluation that allocates nothing (mpci_system_rhs_noalloc).
// Its scratch, device memory beyond the caller's input and output, is this
// file-scope array: a Ctl record at offset 0 (mode RHS, constant field),
// then the single-pulse GEMV partial sums.  Every device the library runs
// this on holds one copy until the process ends.  Host code never uses the
// array's name as an address (on the host it is only a stand-in for the
// symbol): the device address comes from hipGetSymbolAddress, once per
// device with that device current, and is kept in g_scratch_ptr.
// ---------------------------------------------------------------------
constexpr size_t kNoAllocScratchBytes = 10u << 20;  // 10 MiB
constexpr size_t kNoAllocPartOffset = 512;          // partials after the Ctl
static_assert(sizeof(Ctl) <= kNoAllocPartOffset, "Ctl must fit its slot");
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {
constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxScratchDevices] = {};

// The device address of `scratch` on device dev, which must be current.
int scratch_ptr(int dev, void **out) {
  if (dev < 0 || dev >= kMaxScratchDevices)
    return fail(MPCI_ERR_ARG, "mpci_system_rhs_noalloc: device ordinal " +
                                  std::to_string(dev) + " beyond 63");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(MPCI_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") +
                                    hipGetErrorString(e));
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return MPCI_OK;
}
}  // namespace

extern "C" int mpci_system_rhs_noalloc(const mpci_system *sys, double F,
                                       const double *x, double *dxdt,
                                       void *stream) {
  if (!sys || !x || !dxdt)
    return fail(MPCI_ERR_ARG, "mpci_system_rhs_noalloc: NULL argument");
  if ((reinterpret_cast<uintptr_t>(x) | reinterpret_cast<uintptr_t>(dxdt)) &
      15)
    return fail(MPCI_ERR_ARG, "mpci_system_rhs_noalloc: x and dxdt must be "
                              "16-byte aligned device pointers");
  SysDesc d;
  const long long npart = fill_desc(sys, 1, 1, d);
  const long long need =
      (long long)kNoAllocPartOffset + npart * (long long)sizeof(double2);
  if (need > (long long)kNoAllocScratchBytes)
    return fail(MPCI_ERR_ARG,
                "mpci_system_rhs_noalloc: this system's coupling partials "
                "need " + std::to_string(need) + " bytes of scratch; the "
                "fixed scratch holds " +
                std::to_string(kNoAllocScratchBytes) +
                " (use mpci_prop_rhs, which allocates)");
  HIPCHK(hipSetDevice(sys->device));
  void *base = nullptr;
  TRY(scratch_ptr(sys->device, &base));
  Ctl *ctl = static_cast<Ctl *>(base);
  double2 *part =
      reinterpret_cast<double2 *>(static_cast<char *>(base) + kNoAllocPartOffset);
  hipStream_t st = static_cast<hipStream_t>(stream);
  // the Ctl record is rewritten by a kernel (no memset node under capture);
  // every partial slot is written by exactly one GEMV unit before it is read
  hipLaunchKernelGGL(k_rhs_ctl, dim3(1), dim3(64), 0, st, ctl, F);
  HIPCHK(hipGetLastError());
  const double2 *xv = reinterpret_cast<const double2 *>(x);
  if (!sys->jobs.empty()) {
    const int njobs = (int)sys->jobs.size();
    hipLaunchKernelGGL(k_gemv_tile,
                       dim3((unsigned)std::min(njobs, sys->gv_grid)),
                       dim3(kGvThreads), 0, st, sys->djobs.p, njobs, sys->Dg.p,
                       xv, part, sys->ddesc1.p, ctl, 0);
    HIPCHK(hipGetLastError());
  }
  StageArgs a{};
  a.X = xv;
  a.Xs = const_cast<double2 *>(xv);  // read (not used) at stage 0
  a.K = reinterpret_cast<double2 *>(dxdt);
  a.part = part;
  a.E = sys->E.p;
  a.ctl = ctl;
  a.NP = sys->N;
  a.s = 0;
  hipLaunchKernelGGL((k_stage1<false>), dim3((unsigned)stage1_blocks(sys)),
                     dim3(256), 0, st, a, sys->ddesc1.p);
  HIPCHK(hipGetLastError());
  return MPCI_OK;
}

extern "C" int mpci_prop_fixed(mpci_prop *pr, double t0, double dt, int steps,
                               const double *F, int track, double *pop) {
  if (!pr || !F || steps < 0) return fail(MPCI_ERR_ARG, "bad argument");
  if (track < 0 || track >= pr->sys->N)
    return fail(MPCI_ERR_ARG, "track index out of range");
  if (steps == 0) return MPCI_OK;
  TRY(pr->Ftab.alloc((size_t)pr->P * steps));
  TRY(h2d(pr, pr->Ftab.p, F, sizeof(double) * pr->P * steps));
  pr->track = track;
  drop_graph(pr);  // Ftab moved
  std::vector<Ctl> c(pr->P);
  std::memset(c.data(), 0, sizeof(Ctl) * pr->P);
  for (auto &x : c) {
    x.mode = kModeFixed;
    x.fmode = kFieldTable;
    x.t = t0;
    x.h = dt;
    x.hc = dt;
    x.nsteps = steps;
    x.hstride = steps;
    x.max_attempts = steps;
  }
  TRY(start_ctl(pr, c));
  std::vector<std::vector<LogEnt>> acc(pr->P);
  int left = steps;
  while (left > 0) {
    const int n = std::min(left, pr->logcap);
    TRY(run_attempts(pr, n));
    TRY(read_ctl(pr, c));
    TRY(drain_log(pr, c, acc));
    left -= n;
  }
  if (pop)
    for (int p = 0; p < pr->P; ++p)
      for (int st = 0; st < steps && st < (int)acc[p].size(); ++st)
        pop[(size_t)p * steps + st] = acc[p][st].pop;
  return MPCI_OK;
}

// Attempts to launch in the next chunk (0: every pulse has finished).
static int plan_chunk(const mpci_prop *pr, const std::vector<Ctl> &c) {
  int est = 0;
  bool any = false;
  for (auto &x : c) {
    if (x.done) continue;
    any = true;
    int e;
    if (x.mode == kModeAdaptive) {
      // Launch about half of the steps still expected: the proposed h
      // only grows by 4.5x per step, so the chunk rarely overshoots and
      // the number of host synchronisations stays logarithmic.
      double r = 0.5 * (x.t_end - x.t) / std::max(x.h, 1e-300);
      e = (int)std::min(1e6, std::ceil(r));
    } else {
      e = x.nsteps - x.step;
    }
    // never launch attempts past a pulse's attempt limit (they would only
    // return at once)
    e = std::min(e, x.max_attempts - x.attempts);
    est = std::max(est, e);
  }
  if (!any) return 0;
  return std::max(1, std::min(est, pr->logcap));
}

static int run_controlled(mpci_prop *pr, std::vector<Ctl> &c,
                          std::vector<std::vector<LogEnt>> &acc) {
  TRY(start_ctl(pr, c));
  for (int n; (n = plan_chunk(pr, c)) > 0;) {
    TRY(run_attempts(pr, n));
    TRY(read_ctl(pr, c));
    TRY(drain_log(pr, c, acc));
  }
  return MPCI_OK;
}

static int setup_adaptive(mpci_prop *pr, const double *t0, const double *t1,
                          const double *h0, double atol, double rtol,
                          int max_attempts, int track, std::vector<Ctl> &c,
                          std::vector<std::vector<LogEnt>> &acc) {
  if (!pr || !t0 || !t1 || !h0) return fail(MPCI_ERR_ARG, "NULL argument");
  if (!(atol >= 0.0) || !(rtol >= 0.0) || atol + rtol <= 0.0 ||
      max_attempts < 1)
    return fail(MPCI_ERR_ARG, "need atol, rtol >= 0 (not both 0) and "
                              "max_attempts >= 1");
  if (track < 0 || track >= pr->sys->N)
    return fail(MPCI_ERR_ARG, "track index out of range");
  pr->track = track;
  c.assign(pr->P, Ctl{});
  std::memset(c.data(), 0, sizeof(Ctl) * pr->P);
  for (int p = 0; p < pr->P; ++p) {
    if (!(h0[p] > 0.0)) return fail(MPCI_ERR_ARG, "h0 must be > 0");
    Ctl &x = c[p];
    x.mode = kModeAdaptive;
    x.fmode = kFieldPulse;
    x.t = t0[p];
    x.t_end = t1[p];
    x.h = h0[p];
    x.atol = atol;
    x.rtol = rtol;
    x.max_attempts = max_attempts;
    x.done = !(t0[p] < t1[p]);
    x.last = (x.t + x.h >= x.t_end);
    x.hc = x.last ? (x.t_end - x.t) : x.h;
  }
  acc.assign(pr->P, {});
  return MPCI_OK;
}

static int run_adaptive(mpci_prop *pr, const double *t0, const double *t1,
                        const double *h0, double atol, double rtol,
                        int max_attempts, int track, std::vector<Ctl> &c,
                        std::vector<std::vector<LogEnt>> &acc) {
  TRY(setup_adaptive(pr, t0, t1, h0, atol, rtol, max_attempts, track, c, acc));
  return run_controlled(pr, c, acc);
}

static void fill_stats(const std::vector<Ctl> &c, mpci_stats *stats) {
  for (size_t p = 0; p < c.size(); ++p) {
    stats[p].accepted = c[p].accepted;
    stats[p].rejected = c[p].rejected;
    stats[p].attempts = c[p].attempts;
    stats[p].status = c[p].status ? MPCI_ERR_LIMIT : MPCI_OK;
    stats[p].t = c[p].t;
    stats[p].h = c[p].h;
  }
}

extern "C" int mpci_prop_adaptive(mpci_prop *pr, const double *t0,
                                  const double *t1, const double *h0,
                                  double atol, double rtol, int max_attempts,
                                  int track, mpci_stats *stats, double *log,
                                  int log_cap) {
  std::vector<Ctl> c;
  std::vector<std::vector<LogEnt>> acc;
  TRY(run_adaptive(pr, t0, t1, h0, atol, rtol, max_attempts, track, c, acc));
  int rc = MPCI_OK;
  if (stats) fill_stats(c, stats);
  for (int p = 0; p < pr->P; ++p) {
    if (c[p].status) rc = fail(MPCI_ERR_LIMIT, "adaptive: max_attempts hit");
    if (log)
      for (int k = 0; k < (int)acc[p].size() && k < log_cap; ++k) {
        double *e = log + ((size_t)p * log_cap + k) * 3;
        e[0] = acc[p][k].t;
        e[1] = acc[p][k].h;
        e[2] = acc[p][k].pop;
      }
  }
  return rc;
}

// Several propagators (instances: different systems, different P) stepped
// concurrently.  Each member runs on its own stream; whenever a member's
// chunk has drained, its log is collected and its next chunk is enqueued at
// once, so a small instance never waits for a large one and the device
// always has the work of every unfinished member queued.
extern "C" int mpci_multi_adaptive(int n, mpci_prop *const *props,
                                   const double *const *t0,
                                   const double *const *t1,
                                   const double *const *h0, double atol,
                                   double rtol, int max_attempts,
                                   mpci_stats *const *stats) {
  if (n < 1 || !props || !t0 || !t1 || !h0)
    return fail(MPCI_ERR_ARG, "mpci_multi_adaptive: bad argument");
  for (int i = 0; i < n; ++i) {
    if (!props[i] || !t0[i] || !t1[i] || !h0[i])
      return fail(MPCI_ERR_ARG, "mpci_multi_adaptive: NULL member");
    for (int j = 0; j < i; ++j)
      if (props[j] == props[i])
        return fail(MPCI_ERR_ARG, "mpci_multi_adaptive: a propagator is "
                                  "listed twice");
  }
  struct Member {
    mpci_prop *pr;
    std::vector<Ctl> c;
    std::vector<std::vector<LogEnt>> acc;
    bool busy = false, finished = false;
  };
  std::vector<Member> m(n);
  for (int i = 0; i < n; ++i) {
    m[i].pr = props[i];
    TRY(setup_adaptive(props[i], t0[i], t1[i], h0[i], atol, rtol,
                       max_attempts, 0, m[i].c, m[i].acc));
    TRY(start_ctl(props[i], m[i].c));
  }
  for (;;) {
    int live = 0;
    for (auto &x : m) {
      if (x.finished || x.busy) {
        live += !x.finished;
        continue;
      }
      const int k = plan_chunk(x.pr, x.c);
      if (k == 0) {
        x.finished = true;
        continue;
      }
      TRY(run_attempts(x.pr, k));
      x.busy = true;
      ++live;
    }
    if (live == 0) break;
    // wait for any member's stream to drain
    for (bool got = false; !got;) {
      for (auto &x : m) {
        if (!x.busy) continue;
        const hipError_t q = hipStreamQuery(x.pr->stream);
        if (q == hipErrorNotReady) continue;
        if (q != hipSuccess)
          return fail(MPCI_ERR_HIP, std::string("hipStreamQuery: ") +
                                        hipGetErrorString(q));
        TRY(read_ctl(x.pr, x.c));
        TRY(drain_log(x.pr, x.c, x.acc));
        x.busy = false;
        got = true;
      }
      if (!got) std::this_thread::yield();
    }
  }
  int rc = MPCI_OK;
  for (int i = 0; i < n; ++i) {
    if (stats && stats[i]) fill_stats(m[i].c, stats[i]);
    for (auto &x : m[i].c)
      if (x.status) rc = fail(MPCI_ERR_LIMIT, "adaptive: max_attempts hit");
  }
  return rc;
}

// ---------------------------------------------------------------------
// Grouped multi-instance propagation: every member's attempt in the same
// launches.  One job list interleaves all members' GEMM panels (each job
// names its member; every member keeps its own K split, so each element
// sums its K range in the same order as in a separate run), and one stage,
// one finalize and one apply launch cover all members' rows and pulses.
// The attempt is captured once into a graph on the first member's stream.
// ---------------------------------------------------------------------
namespace {
int launch_gemm_group(hipStream_t st, int wm, int wn, unsigned njobs,
                      const GemmJob *jobs, const GroupGemm &m, int s) {
#define MPCI_GEMM_G(WM, WN)                                                  \
  hipLaunchKernelGGL((k_gemm_lds_g<WM, WN>), dim3(njobs), dim3(64 * WM * WN), \
                     0, st, jobs, m, s)
  if (wm == 2 && wn == 1)
    MPCI_GEMM_G(2, 1);
  else if (wm == 4 && wn == 1)
    MPCI_GEMM_G(4, 1);
  else if (wm == 2 && wn == 2)
    MPCI_GEMM_G(2, 2);
  else if (wm == 4 && wn == 2)
    MPCI_GEMM_G(4, 2);
  else if (wm == 2 && wn == 4)
    MPCI_GEMM_G(2, 4);
  else
    MPCI_GEMM_G(4, 4);
#undef MPCI_GEMM_G
  HIPCHK(hipGetLastError());
  return MPCI_OK;
}
}  // namespace

extern "C" int mpci_multi_adaptive_grouped(int n, mpci_prop *const *props,
                                           const double *const *t0,
                                           const double *const *t1,
                                           const double *const *h0,
                                           double atol, double rtol,
                                           int max_attempts,
                                           mpci_stats *const *stats) {
  if (n < 1 || n > kGroupMax || !props || !t0 || !t1 || !h0)
    return fail(MPCI_ERR_ARG, "mpci_multi_adaptive_grouped: bad argument "
                              "(1 to 8 members)");
  for (int i = 0; i < n; ++i) {
    if (!props[i] || !t0[i] || !t1[i] || !h0[i])
      return fail(MPCI_ERR_ARG, "mpci_multi_adaptive_grouped: NULL member");
    for (int j = 0; j < i; ++j)
      if (props[j] == props[i])
        return fail(MPCI_ERR_ARG, "mpci_multi_adaptive_grouped: a "
                                  "propagator is listed twice");
    const mpci_prop *pr = props[i];
    if (pr->P < 2 || pr->sys->nb < 2 || !gemm_lds_fits(pr))
      return fail(MPCI_ERR_NOTSUP,
                  "mpci_multi_adaptive_grouped: every member must be a "
                  "batched propagator (P >= 2, two or more blocks, operands "
                  "below 2 GiB); use mpci_multi_adaptive");
    if (pr->sys->device != props[0]->sys->device)
      return fail(MPCI_ERR_NOTSUP, "mpci_multi_adaptive_grouped: members on "
                                   "different devices");
  }
  // the group's tile: that of the member with the most coupling work when
  // it is a direct-to-LDS tile, else 4 x 2 waves
  int lead = 0;
  double most = -1.0;
  for (int i = 0; i < n; ++i) {
    const double w = (double)props[i]->P * (double)props[i]->sys->S;
    if (w > most) {
      most = w;
      lead = i;
    }
  }
  int wm = 4, wn = 2;
  if (props[lead]->gm_kind == 2) {
    wm = props[lead]->gm_wm;
    wn = props[lead]->gm_wn;
  }
  for (int i = 0; i < n; ++i) HIPCHK(hipStreamSynchronize(props[i]->stream));
  hipStream_t st = props[0]->stream;
  // member table and the interleaved job list
  std::vector<GroupMember> tab(n);
  GroupGemm gg{};
  std::vector<std::vector<GemmJob>> jl(n);
  int wg_stage = 0, wg_apply = 0, pulses = 0;
  size_t njobs = 0;
  for (int i = 0; i < n; ++i) {
    mpci_prop *pr = props[i];
    GroupMember g{};
    g.d = pr->ddesc.p;
    g.Dr = pr->sys->Dr.p;
    g.dr_off = pr->sys->dr_off.p;
    g.sa = stage_args(pr, 0);
    g.X = pr->X.p;
    g.part = reinterpret_cast<double *>(pr->part.p);
    g.hs = pr->hs.p;
    g.Ftab = pr->Ftab.p;
    g.pulses = pr->pulses.p;
    g.env = pr->env.p;
    g.log = pr->log.p;
    g.st = pr->store;
    g.P = pr->P;
    g.N = pr->sys->N;
    g.logcap = pr->logcap;
    g.track = 0;
    g.nblk_stage = pr->nblk_stage;
    g.wg_stage0 = wg_stage;
    g.nblk_apply = pr->nblk;
    g.wg_apply0 = wg_apply;
    g.pulse0 = pulses;
    wg_stage += pr->nblk_stage;
    wg_apply += pr->nblk;
    pulses += pr->P;
    tab[i] = g;
    gg.Dr[i] = g.Dr;
    gg.dr_off[i] = g.dr_off;
    gg.X[i] = reinterpret_cast<const double *>(pr->X.p);
    gg.Xs[i] = reinterpret_cast<const double *>(pr->Xs.p);
    gg.part[i] = g.part;
    gg.d[i] = g.d;
    gg.ctl[i] = pr->ctl.p;
    // a member tuned to k_gemm_pair keeps its own launch (its sums run in
    // another order than the direct-to-LDS kernel's)
    if (pr->gm_kind == 3) continue;
    jl[i] = gemm_job_list(pr, wm, wn, 2, 2, pr->gm_ks, i);
    njobs += jl[i].size();
  }
  // shortest K range first: the short panels of small-block members run
  // while the first long panels start, and the long panels are not left
  // queued behind them at the end (measured on 21 x 998 + 4 x 4000 at
  // P = 32: 4.76 ms per attempt, against 5.26 for an even interleave and
  // 4.90 for longest first)
  std::vector<std::pair<double, GemmJob>> order;
  order.reserve(njobs);
  for (int i = 0; i < n; ++i)
    for (size_t k = 0; k < jl[i].size(); ++k)
      order.push_back({(double)(jl[i][k].k1 - jl[i][k].k0) + 1e-9 * i,
                       jl[i][k]});
  std::stable_sort(order.begin(), order.end(),
                   [](const std::pair<double, GemmJob> &a,
                      const std::pair<double, GemmJob> &b) {
                     return a.first < b.first;
                   });
  std::vector<GemmJob> jobs;
  jobs.reserve(njobs);
  for (auto &o : order) jobs.push_back(o.second);
  DevBuf<GroupMember> dtab;
  DevBuf<GemmJob> djobs;
  TRY(dtab.alloc(n));
  TRY(djobs.alloc(jobs.size()));
  TRY(h2d(props[0], dtab.p, tab.data(), sizeof(GroupMember) * n));
  TRY(h2d(props[0], djobs.p, jobs.data(), sizeof(GemmJob) * jobs.size()));
  // per-member set-up, as mpci_prop_adaptive does it
  struct Member {
    std::vector<Ctl> c;
    std::vector<std::vector<LogEnt>> acc;
  };
  std::vector<Member> m(n);
  for (int i = 0; i < n; ++i) {
    TRY(setup_adaptive(props[i], t0[i], t1[i], h0[i], atol, rtol,
                       max_attempts, 0, m[i].c, m[i].acc));
    TRY(start_ctl(props[i], m[i].c));
  }
  // the grouped attempt, captured once
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  int rc = MPCI_OK;
  HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
  for (int s = 0; s < kStages && rc == MPCI_OK; ++s) {
    if (!jobs.empty())
      rc = launch_gemm_group(st, wm, wn, (unsigned)jobs.size(), djobs.p, gg, s);
    if (rc != MPCI_OK) break;
    for (int i = 0; i < n; ++i) {
      mpci_prop *pr = props[i];
      if (pr->gm_kind != 3) continue;
      hipLaunchKernelGGL(k_gemm_pair, dim3((unsigned)pr->gjobs.size()),
                         dim3(1024), 0, st, pr->dgjobs.p, pr->sys->Dr.p,
                         pr->sys->dr_off.p,
                         reinterpret_cast<const double *>(s == 0 ? pr->X.p
                                                                 : pr->Xs.p),
                         reinterpret_cast<double *>(pr->part.p), pr->ddesc.p,
                         pr->ctl.p);
    }
    if (s == kStages - 1)
      hipLaunchKernelGGL(k_stage_g<true>, dim3(wg_stage), dim3(256), 0, st,
                         dtab.p, n, s);
    else
      hipLaunchKernelGGL(k_stage_g<false>, dim3(wg_stage), dim3(256), 0, st,
                         dtab.p, n, s);
    if (hipGetLastError() != hipSuccess)
      rc = fail(MPCI_ERR_HIP, "mpci_multi_adaptive_grouped: stage launch");
  }
  if (rc == MPCI_OK) {
    hipLaunchKernelGGL(k_finalize_g, dim3(pulses), dim3(256), 0, st, dtab.p, n);
    hipLaunchKernelGGL(k_apply_g, dim3(wg_apply), dim3(256), 0, st, dtab.p, n);
    if (hipGetLastError() != hipSuccess)
      rc = fail(MPCI_ERR_HIP, "mpci_multi_adaptive_grouped: launch");
  }
  const hipError_t ec = hipStreamEndCapture(st, &graph);
  if (rc != MPCI_OK || ec != hipSuccess) {
    if (graph) (void)hipGraphDestroy(graph);
    return rc != MPCI_OK ? rc
                         : fail(MPCI_ERR_HIP, std::string("hipStreamEndCapture: ") +
                                                  hipGetErrorString(ec));
  }
  if (hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0) != hipSuccess) {
    (void)hipGraphDestroy(graph);
    return fail(MPCI_ERR_HIP, "hipGraphInstantiate failed (grouped attempt)");
  }
  // chunks of attempts until every pulse of every member has finished
  for (;;) {
    int k = 0;
    for (int i = 0; i < n; ++i) k = std::max(k, plan_chunk(props[i], m[i].c));
    if (k == 0) break;
    for (int a = 0; a < k && rc == MPCI_OK; ++a)
      if (hipGraphLaunch(gexec, st) != hipSuccess)
        rc = fail(MPCI_ERR_HIP, "mpci_multi_adaptive_grouped: graph launch");
    if (hipStreamSynchronize(st) != hipSuccess && rc == MPCI_OK)
      rc = fail(MPCI_ERR_HIP, "mpci_multi_adaptive_grouped: synchronize");
    for (int i = 0; i < n && rc == MPCI_OK; ++i) {
      rc = read_ctl(props[i], m[i].c);
      if (rc == MPCI_OK) rc = drain_log(props[i], m[i].c, m[i].acc);
    }
    if (rc != MPCI_OK) break;
  }
  (void)hipGraphExecDestroy(gexec);
  (void)hipGraphDestroy(graph);
  TRY(rc);
  for (int i = 0; i < n; ++i) {
    if (stats && stats[i]) fill_stats(m[i].c, stats[i]);
    for (auto &x : m[i].c)
      if (x.status) rc = fail(MPCI_ERR_LIMIT, "adaptive: max_attempts hit");
  }
  return rc;
}

// Schedule replay with per-pulse step counts n[p] (<= stride); hs is
// [P][stride].  Used directly by mpci_prop_schedule and by the adjoint's
// segment recompute.
static int run_schedule(mpci_prop *pr, const double *t0, const int *n,
                        int stride, const double *hs) {
  TRY(pr->hs.alloc((size_t)pr->P * std::max(stride, 1)));
  TRY(h2d(pr, pr->hs.p, hs, sizeof(double) * pr->P * stride));
  drop_graph(pr);  // hs moved
  std::vector<Ctl> c(pr->P);
  std::memset(c.data(), 0, sizeof(Ctl) * pr->P);
  for (int p = 0; p < pr->P; ++p) {
    Ctl &x = c[p];
    x.mode = kModeSchedule;
    x.fmode = kFieldPulse;
    x.t = t0[p];
    x.nsteps = n[p];
    x.hstride = stride;
    x.max_attempts = n[p];
    x.done = n[p] <= 0;
    x.hc = n[p] > 0 ? hs[(size_t)p * stride] : 0.0;
    x.h = x.hc;
  }
  std::vector<std::vector<LogEnt>> acc(pr->P);
  return run_controlled(pr, c, acc);
}

extern "C" int mpci_prop_schedule(mpci_prop *pr, const double *t0, int n,
                                  const double *hs) {
  if (!pr || !t0 || (n > 0 && !hs) || n < 0)
    return fail(MPCI_ERR_ARG, "bad argument");
  if (n == 0) return MPCI_OK;
  std::vector<int> nn(pr->P, n);
  return run_schedule(pr, t0, nn.data(), n, hs);
}

// ---------------------------------------------------------------------
// Discrete adjoint (extension).
// ---------------------------------------------------------------------
namespace {
// Fehlberg 7(8) nodes c_s, as in the device tableau (stage times t + c_s h).
const double kHostC[kStages] = {0, 2.0 / 27.0, 1.0 / 9.0, 1.0 / 6.0,
                                5.0 / 12.0, 0.5, 5.0 / 6.0, 1.0 / 6.0,
                                2.0 / 3.0, 1.0 / 3.0, 1.0, 0, 1.0};

// dF/d env for the piecewise-linear pulse at time t (the derivative of the
// device's field_at with respect to the two samples it interpolates).
void pwl_sensitivity(const PulseDev &pl, double t, int *k, double *w0,
                     double *w1) {
  *k = -1;
  if (!(t >= 0.0 && t <= pl.t_end) || pl.n_env < 2) return;
  const double dx = pl.t_end / (pl.n_env - 1);
  const double u = t / dx;
  int kk = (int)std::floor(u);
  if (kk > pl.n_env - 2) kk = pl.n_env - 2;
  if (kk < 0) kk = 0;
  const double f = u - kk;
  const double sv = pl.scale * std::sin(pl.w * t + pl.cepd);
  *k = kk;
  *w0 = (1.0 - f) * sv;
  *w1 = f * sv;
}

double ms_since(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double, std::milli>(
             std::chrono::steady_clock::now() - t0)
      .count();
}
}  // namespace

extern "C" int mpci_prop_adjoint(mpci_prop *pr, const double *t0,
                                 const double *t1, const double *h0,
                                 double atol, double rtol, int max_steps,
                                 const double *W, int ckpt_every, double *J,
                                 double *grad, int grad_stride, int *nsteps,
                                 double *hlog, int hlog_cap,
                                 mpci_adjoint_stats *stats) {
  if (!pr || !t0 || !t1 || !h0 || !W || !J || !grad || max_steps < 1 ||
      ckpt_every < 1)
    return fail(MPCI_ERR_ARG, "mpci_prop_adjoint: bad argument");
  const int P = pr->P, N = pr->sys->N;
  const long long NP = pr->NP;
  for (int p = 0; p < P; ++p) {
    if (pr->hpulses[p].kind != kPwl)
      return fail(MPCI_ERR_ARG, "the adjoint needs piecewise-linear pulses "
                                "(gradients are taken w.r.t. their samples)");
    if (pr->hpulses[p].n_env > grad_stride)
      return fail(MPCI_ERR_ARG, "grad_stride smaller than n_env");
  }
  const int K = ckpt_every;
  DevBuf<double2> ckpt, xs_store, xn_store, mu, lam, lamx, xfinal;
  DevBuf<double> nrm_store, tn_tab, hn_tab, gdev, red2, dotv, Wd;
  // device bytes held by the adjoint's own buffers, and their peak
  long long held = 0, peak = 0;
  auto hold = [&](long long bytes) {
    held += bytes;
    peak = std::max(peak, held);
  };
  // Checkpoint slot j holds the state after j K accepted steps.  The store
  // grows with the forward pass, which runs in chunks of at most logcap
  // attempts: before each chunk it is made large enough for every step the
  // chunk could accept, so it is sized by the run, not by max_steps.
  int ckcap = 0;
  auto grow_ckpt = [&](int need) -> int {
    if (need <= ckcap) return MPCI_OK;
    const int cap = std::max(need, ckcap + ckcap / 2);
    DevBuf<double2> bigger;
    TRY(bigger.alloc((size_t)cap * NP));
    hold((long long)bigger.bytes());
    if (ckcap > 0)
      HIPCHK(hipMemcpyAsync(bigger.p, ckpt.p, ckpt.bytes(),
                            hipMemcpyDeviceToDevice, pr->stream));
    HIPCHK(hipStreamSynchronize(pr->stream));
    held -= (long long)ckpt.bytes();
    ckpt.swap(bigger);
    ckcap = cap;
    pr->store.ckpt = ckpt.p;
    pr->store.ckpt_cap = ckcap;
    return MPCI_OK;
  };
  TRY(xfinal.alloc(NP));
  TRY(Wd.alloc(N));
  hold((long long)(xfinal.bytes() + Wd.bytes()));
  TRY(h2d(pr, Wd.p, W, sizeof(double) * N));
  const bool graph_was = pr->use_graph;
  pr->use_graph = false;
  drop_graph(pr);
  struct Restore {
    mpci_prop *pr;
    bool g;
    ~Restore() {
      pr->store = StepStore{};
      pr->use_graph = g;
      drop_graph(pr);
    }
  } restore{pr, graph_was};

  // ---- forward: adaptive, checkpoint every K accepted steps ----------
  HIPCHK(hipStreamSynchronize(pr->stream));
  auto tf = std::chrono::steady_clock::now();
  pr->store = StepStore{};
  pr->store.ckpt_every = K;
  TRY(grow_ckpt(2));
  HIPCHK(hipMemcpyAsync(ckpt.p, pr->X.p, sizeof(double2) * NP,
                        hipMemcpyDeviceToDevice, pr->stream));
  std::vector<Ctl> c;
  std::vector<std::vector<LogEnt>> acc;
  TRY(setup_adaptive(pr, t0, t1, h0, atol, rtol, max_steps * 8 + 64, 0, c,
                     acc));
  TRY(start_ctl(pr, c));
  for (int n; (n = plan_chunk(pr, c)) > 0;) {
    int done_max = 0;
    for (const Ctl &x : c) done_max = std::max(done_max, x.accepted);
    if (done_max > max_steps)
      return fail(MPCI_ERR_LIMIT, "adjoint forward took more than max_steps "
                                  "accepted steps");
    TRY(grow_ckpt((done_max + n) / K + 2));
    TRY(run_attempts(pr, n));
    TRY(read_ctl(pr, c));
    TRY(drain_log(pr, c, acc));
  }
  HIPCHK(hipStreamSynchronize(pr->stream));
  const double fwd_ms = ms_since(tf);
  std::vector<int> nst(P);
  int nmax = 0;
  for (int p = 0; p < P; ++p) {
    if (c[p].status)
      return fail(MPCI_ERR_LIMIT, "adjoint forward: attempt limit reached");
    nst[p] = (int)acc[p].size();
    if (nst[p] > max_steps)
      return fail(MPCI_ERR_LIMIT, "adjoint forward took more than max_steps "
                                  "accepted steps");
    nmax = std::max(nmax, nst[p]);
    if (nsteps) nsteps[p] = nst[p];
    if (hlog)
      for (int k = 0; k < nst[p] && k < hlog_cap; ++k)
        hlog[(size_t)p * hlog_cap + k] = acc[p][k].h;
  }
  auto t_before = [&](int p, int n) {  // time at the start of step n
    return n == 0 ? t0[p] : acc[p][n - 1].t;
  };
  HIPCHK(hipMemcpyAsync(xfinal.p, pr->X.p, sizeof(double2) * NP,
                        hipMemcpyDeviceToDevice, pr->stream));

  // ---- objective and its gradient w.r.t. the final state ---------------
  auto tb = std::chrono::steady_clock::now();
  TRY(lam.alloc(NP));
  TRY(lamx.alloc(NP));
  {
    const unsigned nb = (unsigned)((NP + 255) / 256);
    hipLaunchKernelGGL(k_adj_seed, dim3(nb), dim3(256), 0, pr->stream,
                       pr->X.p, Wd.p, N, P, lam.p);
    HIPCHK(hipGetLastError());
    TRY(red2.alloc((size_t)pr->nblk * P));
    TRY(dotv.alloc(P));
    // J_p = <x, W x>: reuse the dot kernels on (x, x * W) = (x, lam) / 2
    hipLaunchKernelGGL(k_dot_partial, dim3(pr->nblk), dim3(256), 0,
                       pr->stream, pr->X.p, lam.p, N, P, red2.p);
    hipLaunchKernelGGL(k_reduce_final, dim3(P), dim3(256), 0, pr->stream,
                       red2.p, pr->nblk, P, dotv.p);
    HIPCHK(hipGetLastError());
    std::vector<double> jj(P);
    HIPCHK(hipStreamSynchronize(pr->stream));
    TRY(d2h(pr, jj.data(), dotv.p, sizeof(double) * P));
    for (int p = 0; p < P; ++p) J[p] = 0.5 * jj[p];
  }

  // ---- backward, one checkpoint segment at a time ----------------------
  TRY(xs_store.alloc((size_t)K * (kStages - 1) * NP));
  TRY(xn_store.alloc((size_t)(K + 1) * NP));
  TRY(nrm_store.alloc((size_t)K * P));
  TRY(mu.alloc((size_t)kStages * NP));
  TRY(tn_tab.alloc((size_t)K * P));
  TRY(hn_tab.alloc((size_t)K * P));
  TRY(gdev.alloc((size_t)K * kStages * P));
  hold((long long)(lam.bytes() + lamx.bytes() + red2.bytes() + dotv.bytes() +
                   xs_store.bytes() + xn_store.bytes() + nrm_store.bytes() +
                   mu.bytes() + tn_tab.bytes() + hn_tab.bytes() +
                   gdev.bytes()));
  for (int p = 0; p < P; ++p)
    for (int k = 0; k < grad_stride; ++k) grad[(size_t)p * grad_stride + k] = 0;

  const int nseg = (nmax + K - 1) / K;
  std::vector<double> hs((size_t)P * K), tn((size_t)K * P), hn((size_t)K * P),
      g((size_t)K * kStages * P), ts(P);
  std::vector<int> nloc(P);
  for (int j = nseg - 1; j >= 0; --j) {
    int kmax = 0;
    for (int p = 0; p < P; ++p) {
      nloc[p] = std::max(0, std::min(K, nst[p] - j * K));
      kmax = std::max(kmax, nloc[p]);
    }
    if (kmax == 0) continue;
    // recompute the segment from checkpoint j, storing what the sweep needs
    HIPCHK(hipMemcpyAsync(pr->X.p, ckpt.p + (size_t)j * NP,
                          sizeof(double2) * NP, hipMemcpyDeviceToDevice,
                          pr->stream));
    HIPCHK(hipMemcpyAsync(xn_store.p, pr->X.p, sizeof(double2) * NP,
                          hipMemcpyDeviceToDevice, pr->stream));
    std::fill(hs.begin(), hs.end(), 0.0);
    std::fill(hn.begin(), hn.end(), 0.0);
    std::fill(tn.begin(), tn.end(), 0.0);
    for (int p = 0; p < P; ++p) {
      ts[p] = t_before(p, std::min(j * K, std::max(nst[p] - 1, 0)));
      for (int k = 0; k < nloc[p]; ++k) {
        const int n = j * K + k;
        hs[(size_t)p * K + k] = acc[p][n].h;
        hn[(size_t)k * P + p] = acc[p][n].h;
        tn[(size_t)k * P + p] = t_before(p, n);
      }
    }
    pr->store = StepStore{};
    pr->store.xs = xs_store.p;
    pr->store.xn = xn_store.p;
    pr->store.nrm = nrm_store.p;
    TRY(run_schedule(pr, ts.data(), nloc.data(), K, hs.data()));
    pr->store = StepStore{};
    TRY(h2d(pr, tn_tab.p, tn.data(), sizeof(double) * K * P));
    TRY(h2d(pr, hn_tab.p, hn.data(), sizeof(double) * K * P));
    // coupling kernels skip finished pulses: mark all live for the sweep
    std::vector<Ctl> live(P);
    std::memset(live.data(), 0, sizeof(Ctl) * P);
    for (auto &x : live) x.mode = kModeRhs;
    TRY(write_ctl(pr, live));
    for (int k = kmax - 1; k >= 0; --k) {
      AdjArgs a{};
      a.part = pr->part.p;
      a.E = pr->sys->E.p;
      a.pulses = pr->pulses.p;
      a.env = pr->env.p;
      a.mu = mu.p;
      a.lam = lam.p;
      a.lamx = lamx.p;
      a.x1 = xn_store.p + (size_t)(k + 1) * NP;
      a.xn = xn_store.p + (size_t)k * NP;
      a.xs = xs_store.p + (size_t)k * (kStages - 1) * NP;
      a.dot = dotv.p;
      a.nrm = nrm_store.p + (size_t)k * P;
      a.tn = tn_tab.p + (size_t)k * P;
      a.hn = hn_tab.p + (size_t)k * P;
      a.red = red2.p;
      a.NP = NP;
      hipLaunchKernelGGL(k_dot_partial, dim3(pr->nblk), dim3(256), 0,
                         pr->stream, a.x1, lam.p, N, P, red2.p);
      hipLaunchKernelGGL(k_reduce_final, dim3(P), dim3(256), 0, pr->stream,
                         red2.p, pr->nblk, P, dotv.p);
      hipLaunchKernelGGL(k_adj_init, dim3(pr->nblk), dim3(256), 0,
                         pr->stream, a, pr->ddesc.p);
      HIPCHK(hipGetLastError());
      for (int st = kStages - 1; st >= 0; --st) {
        TRY(enqueue_tile(pr, mu.p + (size_t)st * NP, st));
        a.s = st;
        hipLaunchKernelGGL(k_adj_stage, dim3(pr->nblk), dim3(256), 0,
                           pr->stream, a, pr->ddesc.p);
        hipLaunchKernelGGL(k_reduce_final, dim3(P), dim3(256), 0, pr->stream,
                           red2.p, pr->nblk, P,
                           gdev.p + ((size_t)k * kStages + st) * P);
        HIPCHK(hipGetLastError());
      }
      std::swap(lam.p, lamx.p);
    }
    HIPCHK(hipStreamSynchronize(pr->stream));
    TRY(d2h(pr, g.data(), gdev.p, sizeof(double) * K * kStages * P));
    for (int p = 0; p < P; ++p)
      for (int k = 0; k < nloc[p]; ++k)
        for (int st = 0; st < kStages; ++st) {
          const double gv = g[((size_t)k * kStages + st) * P + p];
          const double h = hn[(size_t)k * P + p];
          int kk;
          double w0, w1;
          pwl_sensitivity(pr->hpulses[p], tn[(size_t)k * P + p] + kHostC[st] * h,
                          &kk, &w0, &w1);
          if (kk >= 0) {
            grad[(size_t)p * grad_stride + kk] += gv * w0;
            grad[(size_t)p * grad_stride + kk + 1] += gv * w1;
          }
        }
  }
  const double bwd_ms = ms_since(tb);
  // leave the propagator at the forward's final state, all pulses finished
  HIPCHK(hipMemcpyAsync(pr->X.p, xfinal.p, sizeof(double2) * NP,
                        hipMemcpyDeviceToDevice, pr->stream));
  HIPCHK(hipStreamSynchronize(pr->stream));
  for (auto &x : c) x.done = 1;
  TRY(write_ctl(pr, c));
  if (stats) {
    stats->steps_max = nmax;
    stats->segments = nseg;
    stats->fwd_ms = fwd_ms;
    stats->bwd_ms = bwd_ms;
    stats->peak_bytes = peak;
    stats->ckpt_slots = ckcap;
  }
  return MPCI_OK;
}

extern "C" int mpci_prop_populations(mpci_prop *pr, int nobs, const double *W,
                                     double *out) {
  if (!pr || nobs < 1 || !W || !out) return fail(MPCI_ERR_ARG, "bad argument");
  const int N = pr->sys->N;
  DevBuf<double> dW, dout;
  TRY(dW.alloc((size_t)nobs * N));
  TRY(dout.alloc((size_t)nobs * pr->P));
  TRY(h2d(pr, dW.p, W, sizeof(double) * nobs * N));
  hipLaunchKernelGGL(k_weighted_pop, dim3(pr->P, nobs), dim3(256), 0,
                     pr->stream, pr->X.p, dW.p, N, pr->P, dout.p);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(pr->stream));
  TRY(d2h(pr, out, dout.p, sizeof(double) * nobs * pr->P));
  return MPCI_OK;
}

// ---------------------------------------------------------------------
// pes observables (src/pes/pes.cpp): every scalar that pes prints is a sum
// of |c_k|^2 over a set of states, so each is one row of weights, and the
// device evaluates all rows for all pulses with one k_weighted_pop launch.
// ---------------------------------------------------------------------
extern "C" int mpci_obs_weights(int nb, const int *s, const double *const *E,
                                int electrons, double threshold, double *W) {
  if (nb < 1 || nb > kMaxBlocks || !s || !E || !W ||
      (electrons != 1 && electrons != 2))
    return fail(MPCI_ERR_ARG, "mpci_obs_weights: need 1 <= nb <= 32, s, E, "
                              "W and electrons 1 or 2");
  long long N = 0;
  for (int L = 0; L < nb; ++L) {
    if (s[L] < 1 || !E[L])
      return fail(MPCI_ERR_ARG, "mpci_obs_weights: empty block");
    N += s[L];
  }
  if (electrons == 1 && (nb < 2 || s[0] > s[1]))
    return fail(MPCI_ERR_ARG,
                "mpci_obs_weights: the one-electron yield bins onto the l = 1 "
                "energies and needs at least two blocks with s[0] <= s[1] "
                "(pes.cpp:97-100 reads l = 1 energies up to index s[0] - 1)");
  std::fill(W, W + MPCI_OBS_ROWS * N, 0.0);
  double *ground = W + MPCI_OBS_GROUND * N, *yield = W + MPCI_OBS_YIELD * N,
         *excited = W + MPCI_OBS_EXCITED * N, *norm = W + MPCI_OBS_NORM * N;
  ground[0] = 1.0;
  std::fill(norm, norm + N, 1.0);
  long long off = 0;
  for (int L = 0; L < nb; ++L) {
    const double *e = E[L];
    if (electrons == 2) {
      // interior states of each L above the one-electron threshold
      // (pes.cpp:199-208): the first and last state of a block take part
      // in neither sum
      for (int i = 1; i + 1 < s[L]; ++i) {
        if (e[i] - threshold > 0.0) yield[off + i] = 1.0;
        if (off + i > 0) excited[off + i] = 1.0;
      }
    } else {
      // each state i >= 1 of block l goes to the first bin k in
      // [1, s[0] - 2] of the l = 1 spectrum that holds it, each bin reaching
      // halfway to its neighbours (open at both ends); the bins of positive
      // energy make the yield (pes.cpp:96-110)
      const double *bin = E[1];
      for (int i = 1; i < s[L]; ++i) {
        for (int k = 1; k + 1 < s[0]; ++k) {
          const double below = std::fabs(bin[k] - bin[k - 1]);
          const double above = std::fabs(bin[k + 1] - bin[k]);
          if (e[i] > bin[k] - 0.5 * below && e[i] < bin[k] + 0.5 * above) {
            if (bin[k] > 0) yield[off + i] = 1.0;
            break;
          }
        }
      }
    }
    off += s[L];
  }
  return MPCI_OK;
}

extern "C" int mpci_prop_observables(mpci_prop *pr, int electrons,
                                     double threshold, double *out) {
  if (!pr || !out || (electrons != 1 && electrons != 2))
    return fail(MPCI_ERR_ARG, "mpci_prop_observables: bad argument");
  mpci_system *sys = pr->sys;
  const int N = sys->N;
  // the weights stay on the device for repeated calls with the same kind
  if (!pr->obsW.p || pr->obs_electrons != electrons ||
      !(pr->obs_threshold == threshold)) {
    std::vector<const double *> E(sys->nb);
    for (int L = 0; L < sys->nb; ++L) E[L] = sys->hostE[L].data();
    std::vector<double> W((size_t)MPCI_OBS_ROWS * N);
    TRY(mpci_obs_weights(sys->nb, sys->s.data(), E.data(), electrons,
                         threshold, W.data()));
    if (!pr->obsW.p) {
      TRY(pr->obsW.alloc(W.size()));
      TRY(pr->obsOut.alloc((size_t)MPCI_OBS_ROWS * pr->P));
    }
    TRY(h2d(pr, pr->obsW.p, W.data(), sizeof(double) * W.size()));
    pr->obs_electrons = electrons;
    pr->obs_threshold = threshold;
  }
  hipLaunchKernelGGL(k_weighted_pop, dim3(pr->P, MPCI_OBS_ROWS), dim3(256), 0,
                     pr->stream, pr->X.p, pr->obsW.p, N, pr->P,
                     pr->obsOut.p);
  HIPCHK(hipGetLastError());
  return d2h(pr, out, pr->obsOut.p, sizeof(double) * MPCI_OBS_ROWS * pr->P);
}

extern "C" int mpci_prop_set_graph(mpci_prop *pr, int enable) {
  if (!pr) return fail(MPCI_ERR_ARG, "NULL argument");
  pr->use_graph = enable != 0;
  if (!pr->use_graph) drop_graph(pr);
  return MPCI_OK;
}

extern "C" int mpci_prop_set_instrument(mpci_prop *pr, int enable) {
  if (!pr) return fail(MPCI_ERR_ARG, "NULL argument");
  pr->instrument = enable != 0;
  pr->kernel_ms = 0.0;
  pr->kernel_launches = 0;
  return MPCI_OK;
}

extern "C" int mpci_prop_kernel_time(mpci_prop *pr, double *total_ms,
                                     long long *launches,
                                     long long *bytes_per_launch,
                                     long long *flops_per_launch) {
  if (!pr) return fail(MPCI_ERR_ARG, "NULL argument");
  if (total_ms) *total_ms = pr->kernel_ms;
  if (launches) *launches = pr->kernel_launches;
  // algorithmic work of one coupling evaluation: every D_L element streamed
  // once (GEMV form), or 2 products x 2 flops x 2P real columns (GEMM form)
  if (bytes_per_launch) *bytes_per_launch = 8LL * pr->sys->S;
  if (flops_per_launch) *flops_per_launch = 8LL * pr->P * pr->sys->S;
  return MPCI_OK;
}

extern "C" int mpci_prop_gemm_config(mpci_prop *pr, int *wm, int *wn, int *ks,
                                     double *ms, int *kind, int *tk, int *fm,
                                     int *fn) {
  if (!pr) return fail(MPCI_ERR_ARG, "NULL argument");
  if (wm) *wm = pr->gm_wm;
  if (wn) *wn = pr->gm_wn;
  if (ks) *ks = pr->gm_ks;
  if (ms) *ms = pr->gemm_ms;
  if (kind) *kind = pr->gm_kind;
  if (tk) *tk = pr->gm_tk;
  if (fm) *fm = pr->gm_fm;
  if (fn) *fn = pr->gm_fn;
  return MPCI_OK;
}

extern "C" int mpci_prop_gemm_candidates(int *count) {
  if (!count) return fail(MPCI_ERR_ARG, "NULL argument");
  *count = (int)(sizeof(kGemmCands) / sizeof(kGemmCands[0]));
  return MPCI_OK;
}

extern "C" int mpci_prop_set_gemm_candidate(mpci_prop *pr, int index) {
  if (!pr) return fail(MPCI_ERR_ARG, "NULL argument");
  const int n = (int)(sizeof(kGemmCands) / sizeof(kGemmCands[0]));
  if (index < 0 || index >= n)
    return fail(MPCI_ERR_ARG, "GEMM candidate index out of range");
  if (pr->P < 2 || pr->sys->nb < 2)
    return fail(MPCI_ERR_ARG, "this propagator has no batched coupling");
  if (kGemmCands[index].kind >= 2 && !gemm_lds_fits(pr))
    return fail(MPCI_ERR_ARG, "candidate needs buffers below 2 GiB");
  HIPCHK(hipSetDevice(pr->sys->device));
  // the captured step refers to the job list and partial buffers, which
  // are reallocated here
  HIPCHK(hipStreamSynchronize(pr->stream));
  drop_graph(pr);
  return tune_gemm(pr, &kGemmCands[index]);
}

extern "C" int mpci_prop_set_small_path(mpci_prop *pr, int mode,
                                        int *active) {
  if (!pr || mode < -1 || mode > 1)
    return fail(MPCI_ERR_ARG, "mpci_prop_set_small_path: bad argument");
  HIPCHK(hipSetDevice(pr->sys->device));
  HIPCHK(hipStreamSynchronize(pr->stream));
  drop_graph(pr);  // the captured step holds the other path's kernels
  TRY(apply_small_path(pr, mode));
  if (active) *active = pr->small ? 1 : 0;
  return MPCI_OK;
}

// ---------------------------------------------------------------------
// measurement probes (no counterpart in the reference): the ceilings this
// device reaches, beside the published peaks
// ---------------------------------------------------------------------
namespace {
struct EventPair {
  hipEvent_t a = nullptr, b = nullptr;
  ~EventPair() {
    if (a) (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
  }
  int create() {
    HIPCHK(hipEventCreate(&a));
    HIPCHK(hipEventCreate(&b));
    return MPCI_OK;
  }
  // milliseconds between the two events, after both have completed
  int elapsed(float *ms) {
    HIPCHK(hipEventSynchronize(b));
    HIPCHK(hipEventElapsedTime(ms, a, b));
    return MPCI_OK;
  }
};

int device_cus() {
  int dev = 0, ncu = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount,
                            dev) != hipSuccess || ncu <= 0)
    return 256;
  return ncu;
}
}  // namespace

extern "C" int mpci_probe_hbm_read(long long bytes, int reps, double *gbps) {
  if (!gbps || bytes < (1LL << 20) || reps < 1)
    return fail(MPCI_ERR_ARG, "mpci_probe_hbm_read: need gbps, bytes >= "
                              "1 MiB and reps >= 1");
  *gbps = 0.0;
  const long long n = bytes / 16 / 1024 * 1024;  // double2, whole grid steps
  DevBuf<double2> buf;
  DevBuf<double> out;
  TRY(buf.alloc((size_t)n));
  TRY(out.alloc(1));
  HIPCHK(hipMemset(buf.p, 0, (size_t)n * 16));
  EventPair ev;
  TRY(ev.create());
  const int ncu = device_cus();
  double best = 0.0;
  for (int per_cu : {2, 4, 8}) {
    const dim3 grid((unsigned)(per_cu * ncu));
    for (int r = 0; r < 2; ++r)
      hipLaunchKernelGGL(k_probe_read, grid, dim3(256), 0, 0, buf.p, n,
                         out.p);
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(ev.a, 0));
    for (int r = 0; r < reps; ++r)
      hipLaunchKernelGGL(k_probe_read, grid, dim3(256), 0, 0, buf.p, n,
                         out.p);
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(ev.b, 0));
    float ms = 0.f;
    TRY(ev.elapsed(&ms));
    if (ms > 0.f)
      best = std::max(best, (double)n * 16 * reps / (ms * 1e-3) / 1e9);
  }
  *gbps = best;
  return MPCI_OK;
}

extern "C" int mpci_probe_mfma_f64(int iters, double *tflops) {
  if (!tflops || iters < 16)
    return fail(MPCI_ERR_ARG, "mpci_probe_mfma_f64: need tflops and "
                              "iters >= 16");
  *tflops = 0.0;
  DevBuf<double> out;
  TRY(out.alloc(1));
  EventPair ev;
  TRY(ev.create());
  // two workgroups of four waves per CU: two waves per SIMD, four
  // independent accumulator chains each
  const int ncu = device_cus();
  const dim3 grid((unsigned)(2 * ncu));
  const double flops = (double)grid.x * 4 * 4 * iters * (16.0 * 16 * 4 * 2);
  hipLaunchKernelGGL(k_probe_mfma, grid, dim3(256), 0, 0, out.p, iters);
  HIPCHK(hipGetLastError());
  double best = 0.0;
  for (int r = 0; r < 3; ++r) {
    HIPCHK(hipEventRecord(ev.a, 0));
    hipLaunchKernelGGL(k_probe_mfma, grid, dim3(256), 0, 0, out.p, iters);
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(ev.b, 0));
    float ms = 0.f;
    TRY(ev.elapsed(&ms));
    if (ms > 0.f) best = std::max(best, flops / (ms * 1e-3) / 1e12);
  }
  *tflops = best;
  return MPCI_OK;
}
