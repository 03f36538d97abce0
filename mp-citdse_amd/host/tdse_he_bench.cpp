// Stand-alone driver of the single-pulse hot path over the C ABI, for
// hardware-counter runs that need a plain process (no Python runtime).
//
//   tdse_he_bench [attempts] [copy_mib] [batch_pulses] [he|h]
//
// Builds the seeded synthetic He system (4 blocks of 4000 states, velocity
// gauge; with `h`, the H-shaped system of 21 blocks of 998 states that the
// multi-instance bench leg uses), propagates one sin^2 pulse (w = 0.7, E0 = 0.03, 10 cycles) from the
// ground state with the adaptive RKF78 (atol = rtol = 1e-10) for `attempts`
// attempts, launched individually (no graph), then copies copy_mib MiB
// device to device three times.  The copy reads and writes exactly
// copy_mib MiB per launch, which scales byte counters (FETCH_SIZE,
// WRITE_SIZE) measured for the coupling kernel in the same run.
//
// With batch_pulses = P > 1 it then propagates P pulses on the same system
// in the batched form (FP64 MFMA coupling, autotuned tile) for the same
// number of attempts: w ~ U[0.5, 1], E0 log-uniform in [0.005, 0.05],
// 4-20 cycles, drawn from the seeded generator as bench.py draws them.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../../include/mpci_tdse.h"
#include "synth.h"

static int check(int rc, const char *what) {
  if (rc != MPCI_OK) {
    std::fprintf(stderr, "%s failed: %d %s\n", what, rc, mpci_last_error());
    std::exit(1);
  }
  return rc;
}

int main(int argc, char **argv) {
  const int attempts = argc > 1 ? std::atoi(argv[1]) : 20;
  const long long copy_mib = argc > 2 ? std::atoll(argv[2]) : 1024;
  const int batch = argc > 3 ? std::atoi(argv[3]) : 0;
  const bool hshape = argc > 4 && argv[4][0] == 'h' && argv[4][1] == 0;
  // the seeds and energy ranges of mpci/synth.py as bench.py uses them
  const int nb = hshape ? 21 : 4;
  const std::vector<int> s(nb, hshape ? 998 : 4000);
  const unsigned long long seed = hshape ? 20261203ULL : 20261202ULL;
  const double e_lo = hshape ? -0.5 : -2.9;
  std::vector<std::vector<double>> E(nb), D(nb - 1);
  std::vector<double *> Ep(nb), Dp(nb - 1);
  for (int L = 0; L < nb; ++L) {
    E[L].resize(s[L]);
    Ep[L] = E[L].data();
  }
  for (int L = 0; L + 1 < nb; ++L) {
    D[L].resize((size_t)s[L] * s[L + 1]);
    Dp[L] = D[L].data();
  }
  if (mpci_synth_system(seed, nb, s.data(), e_lo, 5.0, Ep.data(), Dp.data()))
    return 1;
  std::vector<const double *> Ec(Ep.begin(), Ep.end()), Dc(Dp.begin(), Dp.end());
  mpci_system *sys = nullptr;
  check(mpci_system_create(nb, s.data(), 'v', Ec.data(), Dc.data(), &sys),
        "mpci_system_create");
  mpci_prop *prop = nullptr;
  check(mpci_prop_create(sys, 1, &prop), "mpci_prop_create");
  check(mpci_prop_set_graph(prop, 0), "mpci_prop_set_graph");
  mpci_pulse pl;
  const double w = 0.7, e0 = 0.03;
  check(mpci_pulse_sine(e0 * e0, w, 0.0, 10, 'v', &pl), "mpci_pulse_sine");
  check(mpci_prop_set_pulse(prop, 0, &pl), "mpci_prop_set_pulse");
  const int N = nb * s[0];
  std::vector<double> x(2 * (size_t)N, 0.0);
  x[0] = 1.0;
  check(mpci_prop_set_state(prop, 0, x.data()), "mpci_prop_set_state");
  const double t0 = 0.0, t1 = mpci_period(w, 10), h0 = 0.05;
  mpci_stats st;
  auto a = std::chrono::steady_clock::now();
  const int rc = mpci_prop_adaptive(prop, &t0, &t1, &h0, 1e-10, 1e-10,
                                    attempts, 0, &st, nullptr, 0);
  if (rc != MPCI_OK && rc != MPCI_ERR_LIMIT) check(rc, "mpci_prop_adaptive");
  const double sec =
      std::chrono::duration<double>(std::chrono::steady_clock::now() - a)
          .count();
  std::printf("attempts %d accepted %d t %.6f in %.3f s\n", st.attempts,
              st.accepted, st.t, sec);
  if (copy_mib > 0) {
    const size_t bytes = (size_t)copy_mib << 20;
    void *src = nullptr, *dst = nullptr;
    if (hipMalloc(&src, bytes) != hipSuccess ||
        hipMalloc(&dst, bytes) != hipSuccess)
      return 1;
    if (hipMemset(src, 0, bytes) != hipSuccess) return 1;
    for (int r = 0; r < 3; ++r)
      if (hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice) != hipSuccess)
        return 1;
    if (hipDeviceSynchronize() != hipSuccess) return 1;
    std::printf("copied %lld MiB device to device 3 times\n", copy_mib);
    (void)hipFree(src);
    (void)hipFree(dst);
  }
  mpci_prop_destroy(prop);
  if (batch > 1) {
    const int P = batch;
    mpci_prop *bp = nullptr;
    check(mpci_prop_create(sys, P, &bp), "mpci_prop_create (batch)");
    check(mpci_prop_set_graph(bp, 0), "mpci_prop_set_graph");
    // the pulse batch of mpci/synth.py pulse_batch (seed + 3, stream 1000)
    std::vector<double> u(3 * (size_t)P);
    mpci_synth_uniform(20261202ULL + 3, 1000, 3LL * P, 0.0, 1.0, u.data());
    std::vector<double> bt0(P, 0.0), bt1(P), bh0(P, 0.05);
    for (int p = 0; p < P; ++p) {
      const double wp = 0.5 + 0.5 * u[p];
      const double ep = std::exp(std::log(0.005) +
                                 (std::log(0.05) - std::log(0.005)) * u[P + p]);
      int cyc = (int)(u[2 * P + p] * 17);
      cyc = 4 + (cyc < 16 ? cyc : 16);
      mpci_pulse pp;
      check(mpci_pulse_sine(ep * ep, wp, 0.0, cyc, 'v', &pp),
            "mpci_pulse_sine");
      check(mpci_prop_set_pulse(bp, p, &pp), "mpci_prop_set_pulse");
      check(mpci_prop_set_state(bp, p, x.data()), "mpci_prop_set_state");
      bt1[p] = mpci_period(wp, cyc);
    }
    std::vector<mpci_stats> bst(P);
    a = std::chrono::steady_clock::now();
    const int brc = mpci_prop_adaptive(bp, bt0.data(), bt1.data(), bh0.data(),
                                       1e-10, 1e-10, attempts, 0, bst.data(),
                                       nullptr, 0);
    if (brc != MPCI_OK && brc != MPCI_ERR_LIMIT)
      check(brc, "mpci_prop_adaptive (batch)");
    const double bsec =
        std::chrono::duration<double>(std::chrono::steady_clock::now() - a)
            .count();
    long long acc = 0;
    for (const auto &b : bst) acc += b.accepted;
    std::printf("batch P=%d: %lld accepted pulse-steps in %.3f s\n", P, acc,
                bsec);
    mpci_prop_destroy(bp);
  }
  mpci_system_destroy(sys);
  return 0;
}
