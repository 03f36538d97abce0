// Host side of the MI355X tdse engine and its C ABI (include/mpci_tdse.h).
//
// One RKF78 attempt = 13 x (k_gemv_tile, k_stage) + k_finalize + k_apply,
// captured once into a hipGraph and replayed.  Step control lives in
// device memory (Ctl), so the host only re-launches the graph in chunks and
// drains the per-step log between chunks.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <new>
#include <string>
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
  std::vector<TileJob> jobs;
  DevBuf<TileJob> djobs;
  std::vector<int> nrb, ncb;   // per pair: row tiles, column tiles
  long long S = 0;             // sum_L s_L s_{L+1} (dipole elements)
  int device = 0;
};

struct mpci_prop {
  mpci_system *sys = nullptr;
  int P = 1;
  hipStream_t stream = nullptr;
  SysDesc desc{};
  long long NP = 0;
  int nblk = 0;           // k_stage workgroups
  DevBuf<double2> X, Xs, K, Y, part;
  DevBuf<double> red;
  DevBuf<Ctl> ctl;
  DevBuf<PulseDev> pulses;
  DevBuf<double> env;
  DevBuf<double> Ftab;
  DevBuf<double> hs;
  DevBuf<LogEnt> log;
  int logcap = 4096;
  int track = 0;
  std::vector<PulseDev> hpulses;
  std::vector<std::vector<double>> henv;
  bool use_graph = true;
  hipGraphExec_t gexec = nullptr;
  hipGraph_t graph = nullptr;
  int graph_track = -1;
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
static int build_system(mpci_system *sys, const double *const *E,
                        const double *const *D) {
  TRY(sys->E.alloc(sys->N));
  for (int L = 0; L < sys->nb; ++L)
    HIPCHK(hipMemcpy(sys->E.p + sys->off[L], E[L], sizeof(double) * sys->s[L],
                     hipMemcpyHostToDevice));
  const int npair = sys->nb - 1;
  sys->nrb.assign(std::max(npair, 0), 0);
  sys->ncb.assign(std::max(npair, 0), 0);
  long long total = 0;
  std::vector<long long> base(std::max(npair, 0));
  for (int L = 0; L < npair; ++L) {
    sys->nrb[L] = (sys->s[L] + kGvTR - 1) / kGvTR;
    sys->ncb[L] = (sys->s[L + 1] + kGvTC - 1) / kGvTC;
    base[L] = total;
    total += (long long)sys->nrb[L] * sys->ncb[L] * kGvTileElems;
    sys->S += (long long)sys->s[L] * sys->s[L + 1];
  }
  TRY(sys->Dg.alloc(total));
  sys->jobs.clear();
  for (int L = 0; L < npair; ++L) {
    const long long ne = (long long)sys->s[L] * sys->s[L + 1];
    DevBuf<double> tmp;
    TRY(tmp.alloc(ne));
    HIPCHK(hipMemcpy(tmp.p, D[L], sizeof(double) * ne, hipMemcpyHostToDevice));
    const long long nout = (long long)sys->nrb[L] * sys->ncb[L] * kGvTileElems;
    const int tpb = 256;
    const long long nb = (nout + tpb - 1) / tpb;
    hipLaunchKernelGGL(k_relayout_gemv, dim3((unsigned)nb), dim3(tpb), 0, 0,
                       tmp.p, sys->s[L], sys->s[L + 1], sys->ncb[L], nout,
                       sys->Dg.p + base[L]);
    HIPCHK(hipGetLastError());
    HIPCHK(hipDeviceSynchronize());
    for (int rb = 0; rb < sys->nrb[L]; ++rb)
      for (int cb = 0; cb < sys->ncb[L]; ++cb) {
        TileJob j;
        j.L = L;
        j.rb = rb;
        j.cb = cb;
        j.pad = 0;
        j.doff = base[L] + ((long long)rb * sys->ncb[L] + cb) * kGvTileElems;
        sys->jobs.push_back(j);
      }
  }
  TRY(sys->djobs.alloc(sys->jobs.size()));
  if (!sys->jobs.empty())
    HIPCHK(hipMemcpy(sys->djobs.p, sys->jobs.data(),
                     sizeof(TileJob) * sys->jobs.size(),
                     hipMemcpyHostToDevice));
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
  int rc = build_system(sys, E, D);
  if (rc != MPCI_OK) {
    delete sys;
    return rc;
  }
  *out = sys;
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
static int init_desc(mpci_prop *pr) {
  mpci_system *sys = pr->sys;
  SysDesc &d = pr->desc;
  std::memset(&d, 0, sizeof(d));
  d.nb = sys->nb;
  d.N = sys->N;
  d.P = pr->P;
  d.gauge = sys->gauge;
  for (int L = 0; L <= sys->nb; ++L) d.off[L] = sys->off[L];
  long long o = 0;
  for (int L = 0; L + 1 < sys->nb; ++L) {
    d.nrow_part[L] = sys->ncb[L];
    d.ncol_part[L] = sys->nrb[L];
    d.row_part_off[L] = o;
    o += (long long)sys->ncb[L] * sys->s[L] * pr->P;
    d.col_part_off[L] = o;
    o += (long long)sys->nrb[L] * sys->s[L + 1] * pr->P;
  }
  TRY(pr->part.alloc(std::max<long long>(o, 1)));
  return MPCI_OK;
}

extern "C" int mpci_prop_create(mpci_system *sys, int P, mpci_prop **out) {
  if (!out || !sys) return fail(MPCI_ERR_ARG, "NULL argument");
  *out = nullptr;
  if (P != 1)
    return fail(MPCI_ERR_NOTSUP, "this build propagates P = 1 pulse per "
                                 "propagator");
  std::unique_ptr<mpci_prop> pr(new (std::nothrow) mpci_prop());
  if (!pr) return fail(MPCI_ERR_HIP, "host allocation failed");
  pr->sys = sys;
  pr->P = P;
  HIPCHK(hipSetDevice(sys->device));
  HIPCHK(hipStreamCreateWithFlags(&pr->stream, hipStreamNonBlocking));
  pr->NP = (long long)sys->N * P;
  const int rpb = 256 / P;
  pr->nblk = (sys->N + rpb - 1) / rpb;
  TRY(init_desc(pr.get()));
  TRY(pr->X.alloc(pr->NP));
  TRY(pr->Xs.alloc(pr->NP));
  TRY(pr->Y.alloc(pr->NP));
  TRY(pr->K.alloc(pr->NP * kStages));
  TRY(pr->red.alloc((size_t)pr->nblk * P * 2));
  TRY(pr->ctl.alloc(P));
  TRY(pr->pulses.alloc(P));
  TRY(pr->log.alloc((size_t)pr->logcap * P));
  HIPCHK(hipMemset(pr->X.p, 0, sizeof(double2) * pr->NP));
  HIPCHK(hipMemset(pr->part.p, 0, sizeof(double2) * pr->part.n));
  std::vector<Ctl> c(P);
  std::memset(c.data(), 0, sizeof(Ctl) * P);
  for (auto &x : c) x.done = 1;
  HIPCHK(hipMemcpy(pr->ctl.p, c.data(), sizeof(Ctl) * P,
                   hipMemcpyHostToDevice));
  pr->hpulses.assign(P, PulseDev{});
  pr->henv.assign(P, {});
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
    HIPCHK(hipMemcpy(pr->X.p, ct, sizeof(double2) * pr->sys->N,
                     hipMemcpyHostToDevice));
    return MPCI_OK;
  }
  HIPCHK(hipMemcpy2D(pr->X.p + p, sizeof(double2) * pr->P, ct,
                     sizeof(double2), sizeof(double2), pr->sys->N,
                     hipMemcpyHostToDevice));
  return MPCI_OK;
}

extern "C" int mpci_prop_get_state(mpci_prop *pr, int p, double *ct) {
  if (!pr || !ct || p < 0 || p >= pr->P)
    return fail(MPCI_ERR_ARG, "mpci_prop_get_state: bad argument");
  HIPCHK(hipStreamSynchronize(pr->stream));
  if (pr->P == 1) {
    HIPCHK(hipMemcpy(ct, pr->X.p, sizeof(double2) * pr->sys->N,
                     hipMemcpyDeviceToHost));
    return MPCI_OK;
  }
  HIPCHK(hipMemcpy2D(ct, sizeof(double2), pr->X.p + p,
                     sizeof(double2) * pr->P, sizeof(double2), pr->sys->N,
                     hipMemcpyDeviceToHost));
  return MPCI_OK;
}

static int upload_pulses(mpci_prop *pr) {
  size_t tot = 0;
  for (int p = 0; p < pr->P; ++p) {
    pr->hpulses[p].env_off = (long long)tot;
    tot += pr->henv[p].size();
  }
  if (tot > 0) {
    TRY(pr->env.alloc(tot));
    std::vector<double> flat;
    flat.reserve(tot);
    for (auto &v : pr->henv) flat.insert(flat.end(), v.begin(), v.end());
    HIPCHK(hipMemcpy(pr->env.p, flat.data(), sizeof(double) * tot,
                     hipMemcpyHostToDevice));
  }
  HIPCHK(hipMemcpy(pr->pulses.p, pr->hpulses.data(),
                   sizeof(PulseDev) * pr->P, hipMemcpyHostToDevice));
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
  const bool ev = pr->instrument;
  if (ev) HIPCHK(hipEventRecord(pr->ev[pr->ev_base + 2 * s], pr->stream));
  hipLaunchKernelGGL(k_gemv_tile, dim3((unsigned)sys->jobs.size()),
                     dim3(kGvThreads), 0, pr->stream, sys->djobs.p, sys->Dg.p,
                     xin, pr->part.p, pr->desc, pr->ctl.p);
  HIPCHK(hipGetLastError());
  if (ev) HIPCHK(hipEventRecord(pr->ev[pr->ev_base + 2 * s + 1], pr->stream));
  return MPCI_OK;
}

static StageArgs stage_args(mpci_prop *pr, int s) {
  StageArgs a{};
  a.X = pr->X.p;
  a.Xs = pr->Xs.p;
  a.K = pr->K.p;
  a.Y = pr->Y.p;
  a.part = pr->part.p;
  a.E = pr->sys->E.p;
  a.red = pr->red.p;
  a.ctl = pr->ctl.p;
  a.pulses = pr->pulses.p;
  a.env = pr->env.p;
  a.Ftab = pr->Ftab.p;
  a.NP = pr->NP;
  a.s = s;
  return a;
}

static int enqueue_attempt(mpci_prop *pr) {
  for (int s = 0; s < kStages; ++s) {
    TRY(enqueue_tile(pr, s == 0 ? pr->X.p : pr->Xs.p, s));
    hipLaunchKernelGGL(k_stage, dim3(pr->nblk), dim3(256), 0, pr->stream,
                       stage_args(pr, s), pr->desc);
    HIPCHK(hipGetLastError());
  }
  hipLaunchKernelGGL(k_finalize, dim3(pr->P), dim3(256), 0, pr->stream,
                     pr->ctl.p, pr->red.p, pr->nblk, pr->P, pr->hs.p);
  HIPCHK(hipGetLastError());
  hipLaunchKernelGGL(k_apply, dim3(pr->nblk), dim3(256), 0, pr->stream,
                     pr->X.p, pr->Y.p, pr->ctl.p, pr->log.p, pr->logcap,
                     pr->track, pr->sys->N, pr->P);
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
  HIPCHK(hipMemcpy(c.data(), pr->ctl.p, sizeof(Ctl) * pr->P,
                   hipMemcpyDeviceToHost));
  return MPCI_OK;
}

static int write_ctl(mpci_prop *pr, const std::vector<Ctl> &c) {
  HIPCHK(hipMemcpy(pr->ctl.p, c.data(), sizeof(Ctl) * pr->P,
                   hipMemcpyHostToDevice));
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
    HIPCHK(hipMemcpy(h.data(), pr->log.p, sizeof(LogEnt) * h.size(),
                     hipMemcpyDeviceToHost));
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
  TRY(write_ctl(pr, c));
  TRY(enqueue_tile(pr, pr->X.p, 0));
  StageArgs a = stage_args(pr, 0);
  a.red = nullptr;
  hipLaunchKernelGGL(k_stage, dim3(pr->nblk), dim3(256), 0, pr->stream, a,
                     pr->desc);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(pr->stream));
  std::vector<double2> k(pr->NP);
  HIPCHK(hipMemcpy(k.data(), pr->K.p, sizeof(double2) * pr->NP,
                   hipMemcpyDeviceToHost));
  const int N = pr->sys->N;
  for (int p = 0; p < pr->P; ++p)
    for (int i = 0; i < N; ++i) {
      dxdt[(size_t)p * 2 * N + 2 * i] = k[(size_t)i * pr->P + p].x;
      dxdt[(size_t)p * 2 * N + 2 * i + 1] = k[(size_t)i * pr->P + p].y;
    }
  for (auto &x : c) x.done = 1;
  return write_ctl(pr, c);
}

extern "C" int mpci_prop_fixed(mpci_prop *pr, double t0, double dt, int steps,
                               const double *F, int track, double *pop) {
  if (!pr || !F || steps < 0) return fail(MPCI_ERR_ARG, "bad argument");
  if (track < 0 || track >= pr->sys->N)
    return fail(MPCI_ERR_ARG, "track index out of range");
  if (steps == 0) return MPCI_OK;
  TRY(pr->Ftab.alloc((size_t)pr->P * steps));
  HIPCHK(hipMemcpy(pr->Ftab.p, F, sizeof(double) * pr->P * steps,
                   hipMemcpyHostToDevice));
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
    x.max_attempts = steps;
  }
  TRY(write_ctl(pr, c));
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

static int run_controlled(mpci_prop *pr, std::vector<Ctl> &c,
                          std::vector<std::vector<LogEnt>> &acc) {
  TRY(write_ctl(pr, c));
  for (;;) {
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
      est = std::max(est, e);
    }
    if (!any) break;
    const int n = std::max(1, std::min(est, pr->logcap));
    TRY(run_attempts(pr, n));
    TRY(read_ctl(pr, c));
    TRY(drain_log(pr, c, acc));
  }
  return MPCI_OK;
}

extern "C" int mpci_prop_adaptive(mpci_prop *pr, const double *t0,
                                  const double *t1, const double *h0,
                                  double atol, double rtol, int max_attempts,
                                  int track, mpci_stats *stats, double *log,
                                  int log_cap) {
  if (!pr || !t0 || !t1 || !h0) return fail(MPCI_ERR_ARG, "NULL argument");
  if (!(atol >= 0.0) || !(rtol >= 0.0) || atol + rtol <= 0.0 ||
      max_attempts < 1)
    return fail(MPCI_ERR_ARG, "need atol, rtol >= 0 (not both 0) and "
                              "max_attempts >= 1");
  if (track < 0 || track >= pr->sys->N)
    return fail(MPCI_ERR_ARG, "track index out of range");
  pr->track = track;
  std::vector<Ctl> c(pr->P);
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
  std::vector<std::vector<LogEnt>> acc(pr->P);
  TRY(run_controlled(pr, c, acc));
  int rc = MPCI_OK;
  for (int p = 0; p < pr->P; ++p) {
    if (stats) {
      stats[p].accepted = c[p].accepted;
      stats[p].rejected = c[p].rejected;
      stats[p].attempts = c[p].attempts;
      stats[p].status = c[p].status ? MPCI_ERR_LIMIT : MPCI_OK;
      stats[p].t = c[p].t;
      stats[p].h = c[p].h;
    }
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

extern "C" int mpci_prop_schedule(mpci_prop *pr, const double *t0, int n,
                                  const double *hs) {
  if (!pr || !t0 || (n > 0 && !hs) || n < 0)
    return fail(MPCI_ERR_ARG, "bad argument");
  if (n == 0) return MPCI_OK;
  TRY(pr->hs.alloc((size_t)pr->P * n));
  HIPCHK(hipMemcpy(pr->hs.p, hs, sizeof(double) * pr->P * n,
                   hipMemcpyHostToDevice));
  drop_graph(pr);  // hs moved
  std::vector<Ctl> c(pr->P);
  std::memset(c.data(), 0, sizeof(Ctl) * pr->P);
  for (int p = 0; p < pr->P; ++p) {
    Ctl &x = c[p];
    x.mode = kModeSchedule;
    x.fmode = kFieldPulse;
    x.t = t0[p];
    x.hc = hs[(size_t)p * n];
    x.h = x.hc;
    x.nsteps = n;
    x.max_attempts = n;
  }
  std::vector<std::vector<LogEnt>> acc(pr->P);
  return run_controlled(pr, c, acc);
}

extern "C" int mpci_prop_populations(mpci_prop *pr, int nobs, const double *W,
                                     double *out) {
  if (!pr || nobs < 1 || !W || !out) return fail(MPCI_ERR_ARG, "bad argument");
  const int N = pr->sys->N;
  DevBuf<double> dW, dout;
  TRY(dW.alloc((size_t)nobs * N));
  TRY(dout.alloc((size_t)nobs * pr->P));
  HIPCHK(hipMemcpy(dW.p, W, sizeof(double) * nobs * N, hipMemcpyHostToDevice));
  hipLaunchKernelGGL(k_weighted_pop, dim3(pr->P, nobs), dim3(256), 0,
                     pr->stream, pr->X.p, dW.p, N, pr->P, dout.p);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(pr->stream));
  HIPCHK(hipMemcpy(out, dout.p, sizeof(double) * nobs * pr->P,
                   hipMemcpyDeviceToHost));
  return MPCI_OK;
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
                                     long long *bytes_per_launch) {
  if (!pr) return fail(MPCI_ERR_ARG, "NULL argument");
  if (total_ms) *total_ms = pr->kernel_ms;
  if (launches) *launches = pr->kernel_launches;
  if (bytes_per_launch) *bytes_per_launch = 8LL * pr->sys->S;
  return MPCI_OK;
}
