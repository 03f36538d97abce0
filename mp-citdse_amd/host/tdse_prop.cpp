// Output side of the tdse drop-in: tdse::propV / tdse::propL with the files,
// precision and cadence of src/tdse/tdse.cpp:146-310, stepping on the GPU.
//
// The reference writes one population line and one field line per step
// and a coefficient file at t0, after every step st with st % (steps/10)
// == 0, and at the end.  Here the steps between two coefficient files run
// as one device-resident segment (mpci_prop_fixed); the host evaluates the
// frozen field F(t + dt) with t accumulated by t += dt exactly as the
// reference does and hands the table to the device.
#include <cmath>
#include <complex>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "tdse_host.hpp"

namespace {

struct Engine {
  mpci_system *sys = nullptr;
  mpci_prop *prop = nullptr;
  ~Engine() {
    if (prop) mpci_prop_destroy(prop);
    if (sys) mpci_system_destroy(sys);
  }
};

int report(int rc, const char *what) {
  if (rc != MPCI_OK)
    std::cerr << "tdse: " << what << " failed: " << mpci_last_error() << "\n";
  return rc;
}

int make_engine(Engine &eng, char gauge, int L_max,
                std::vector<int> &state_sz, stvupt &eig, stvupt &dipoles,
                const std::vector<double> &ct) {
  std::vector<const double *> E, D;
  for (int L = 0; L <= L_max; ++L) E.push_back(eig[L]->data());
  for (int L = 0; L < L_max; ++L) D.push_back(dipoles[L]->data());
  int rc = mpci_system_create(L_max + 1, state_sz.data(), gauge, E.data(),
                              D.data(), &eng.sys);
  if (rc != MPCI_OK) return report(rc, "mpci_system_create");
  rc = mpci_prop_create(eng.sys, 1, &eng.prop);
  if (rc != MPCI_OK) return report(rc, "mpci_prop_create");
  return report(mpci_prop_set_state(eng.prop, 0, ct.data()),
                "mpci_prop_set_state");
}

// <pot>_obs.dat (extension): pes's scalars of the device state at every
// coefficient-file time, evaluated on the GPU (mpci_prop_observables).
class ObsLog {
 public:
  ObsLog(const std::string &output, const tdrd::Extensions &ext) : ext_(ext) {
    if (!ext.observables) return;
    f_.open(output + "_obs.dat", std::ios::out);
    f_ << "#time (a.u.), ground state pop, yield, excited population, "
          "norm\n";
  }
  int record(mpci_prop *prop, double t) {
    if (!ext_.observables) return MPCI_OK;
    double v[MPCI_OBS_ROWS];
    const int rc = mpci_prop_observables(prop, ext_.obs_electrons,
                                         ext_.obs_threshold, v);
    if (rc != MPCI_OK) return report(rc, "mpci_prop_observables");
    f_ << std::setprecision(16) << t;
    for (double x : v) f_ << " " << x;
    f_ << "\n";
    return MPCI_OK;
  }

 private:
  const tdrd::Extensions &ext_;
  std::fstream f_;
};

double popOf(const std::vector<double> &ct, int k) {
  return std::norm(std::complex<double>(ct[2 * k], ct[2 * k + 1]));
}

int prop_fixed(char gauge, std::string output, int L_max, double t, double dt,
               int steps, int pop_n, int pop_l, const mpci_pulse &pars,
               int ct_sz, std::vector<int> &offs, std::vector<int> &state_sz,
               stvupt &eig, stvupt &dipoles, std::vector<double> &ct,
               const tdrd::Extensions &ext) {
  const int print = steps / 10;
  tdse::writeCt(output, t, ct, ct_sz);
  if (ext.full_precision) tdse::writeCt(output, t, ct, ct_sz, 17);

  Engine eng;
  int rc = make_engine(eng, gauge, L_max, state_sz, eig, dipoles, ct);
  if (rc != MPCI_OK) return rc;
  ObsLog obs(output, ext);
  if ((rc = obs.record(eng.prop, t)) != MPCI_OK) return rc;

  std::fstream field_fl, f_pop;
  field_fl.open(output + "_field.dat", std::ios::out);
  field_fl << "#time (a.u.), field (A(t) or E(t)) (a.u.)\n";
  f_pop.open(output + "_pop" + std::to_string(pop_n + 1) +
                 std::to_string(pop_l) + ".dat",
             std::ios::out);
  const int track = offs[pop_l] + pop_n;
  f_pop << "#time (a.u.), population\n";
  f_pop << std::setprecision(16) << t << " " << popOf(ct, track) << "\n";

  int st = 0;
  std::vector<double> F, pop;
  while (st < steps) {
    // the segment ends at the next step whose index is a multiple of
    // `print` (a coefficient file follows it), or at the last step; with
    // fewer than 10 steps (print == 0, where the reference divides by
    // zero) every step ends a segment
    int end = (print > 0) ? ((st + print - 1) / print) * print : st;
    if (end >= steps) end = steps - 1;
    const int n = end - st + 1;
    F.resize(n);
    pop.resize(n);
    double tt = t;
    for (int k = 0; k < n; ++k) {
      F[k] = mpci_pulse_eval(&pars, tt + dt);
      tt += dt;
    }
    rc = mpci_prop_fixed(eng.prop, t, dt, n, F.data(), track, pop.data());
    if (rc != MPCI_OK) return report(rc, "mpci_prop_fixed");
    for (int k = 0; k < n; ++k) {
      field_fl << t + dt << " " << F[k] << "\n";
      t += dt;
      f_pop << std::setprecision(16) << t << " " << pop[k] << "\n";
    }
    st = end + 1;
    if (print == 0 || end % print == 0) {
      rc = mpci_prop_get_state(eng.prop, 0, ct.data());
      if (rc != MPCI_OK) return report(rc, "mpci_prop_get_state");
      std::cout << F[n - 1] << "\n";
      tdse::writeCt(output, t, ct, ct_sz);
      if (ext.full_precision) tdse::writeCt(output, t, ct, ct_sz, 17);
      if ((rc = obs.record(eng.prop, t)) != MPCI_OK) return rc;
    }
  }
  rc = mpci_prop_get_state(eng.prop, 0, ct.data());
  if (rc != MPCI_OK) return report(rc, "mpci_prop_get_state");
  tdse::writeCt(output, t, ct, ct_sz);
  if (ext.full_precision) tdse::writeCt(output, t, ct, ct_sz, 17);
  if ((rc = obs.record(eng.prop, t)) != MPCI_OK) return rc;
  field_fl.close();
  f_pop.close();
  return 0;
}

}  // namespace

namespace tdse {

void writeCt(const std::string &output, double t,
             const std::vector<double> &ct, int ct_sz, int precision) {
  std::fstream f_ct;
  const std::string tag = precision > 0 ? "_ct17_" : "_ct_";
  f_ct.open(output + tag + std::to_string(t) + ".dat", std::ios::out);
  if (precision > 0) f_ct << std::setprecision(precision);
  f_ct << "#Index, Re(c(t)), Im(c(t))\n";
  for (auto k = 0; k < ct_sz * 2; k += 2)
    f_ct << k / 2 << "  " << ct[k] << "  " << ct[k + 1] << "\n";
  f_ct.close();
}

int propV(std::string output, int L_max, double t, double dt, int steps,
          int pop_n, int pop_l, const mpci_pulse &pars, int ct_sz,
          std::vector<int> &offs, std::vector<int> &state_sz, stvupt &eig,
          stvupt &dipoles, std::vector<double> &ct,
          const tdrd::Extensions &ext) {
  return prop_fixed('v', output, L_max, t, dt, steps, pop_n, pop_l, pars,
                    ct_sz, offs, state_sz, eig, dipoles, ct, ext);
}

int propL(std::string output, int L_max, double t, double dt, int steps,
          int pop_n, int pop_l, const mpci_pulse &pars, int ct_sz,
          std::vector<int> &offs, std::vector<int> &state_sz, stvupt &eig,
          stvupt &dipoles, std::vector<double> &ct,
          const tdrd::Extensions &ext) {
  return prop_fixed('l', output, L_max, t, dt, steps, pop_n, pop_l, pars,
                    ct_sz, offs, state_sz, eig, dipoles, ct, ext);
}

int propAdaptive(std::string output, char gauge, int L_max, double t,
                 double t_end, double h0, int pop_n, int pop_l,
                 const mpci_pulse &pars, int ct_sz, std::vector<int> &offs,
                 std::vector<int> &state_sz, stvupt &eig, stvupt &dipoles,
                 std::vector<double> &ct, const tdrd::Extensions &ext) {
  writeCt(output, t, ct, ct_sz);
  if (ext.full_precision) writeCt(output, t, ct, ct_sz, 17);
  Engine eng;
  int rc = make_engine(eng, gauge, L_max, state_sz, eig, dipoles, ct);
  if (rc != MPCI_OK) return rc;
  rc = mpci_prop_set_pulse(eng.prop, 0, &pars);
  if (rc != MPCI_OK) return report(rc, "mpci_prop_set_pulse");
  ObsLog obs(output, ext);
  if ((rc = obs.record(eng.prop, t)) != MPCI_OK) return rc;

  std::fstream field_fl, f_pop;
  field_fl.open(output + "_field.dat", std::ios::out);
  field_fl << "#time (a.u.), field (A(t) or E(t)) (a.u.)\n";
  f_pop.open(output + "_pop" + std::to_string(pop_n + 1) +
                 std::to_string(pop_l) + ".dat",
             std::ios::out);
  const int track = offs[pop_l] + pop_n;
  f_pop << "#time (a.u.), population\n";
  f_pop << std::setprecision(16) << t << " " << popOf(ct, track) << "\n";

  // ten segments ending at t0 + k (t_end - t0) / 10: a coefficient file at
  // each segment end, one population and field line per accepted step
  const double t0 = t;
  double h = h0;
  const int log_cap = 1 << 20;
  std::vector<double> log((size_t)log_cap * 3);
  for (int k = 1; k <= 10 && t < t_end; ++k) {
    const double t1 = (k == 10) ? t_end : t0 + k * (t_end - t0) / 10.0;
    if (!(t1 > t)) continue;
    mpci_stats st{};
    rc = mpci_prop_adaptive(eng.prop, &t, &t1, &h, ext.atol, ext.rtol,
                            100000000, track, &st, log.data(), log_cap);
    if (rc != MPCI_OK) return report(rc, "mpci_prop_adaptive");
    for (int i = 0; i < st.accepted && i < log_cap; ++i) {
      const double ti = log[3 * i], pi = log[3 * i + 2];
      field_fl << ti << " " << mpci_pulse_eval(&pars, ti) << "\n";
      f_pop << std::setprecision(16) << ti << " " << pi << "\n";
    }
    t = st.t;
    h = st.h;
    rc = mpci_prop_get_state(eng.prop, 0, ct.data());
    if (rc != MPCI_OK) return report(rc, "mpci_prop_get_state");
    std::cout << mpci_pulse_eval(&pars, t) << "\n";
    writeCt(output, t, ct, ct_sz);
    if (ext.full_precision) writeCt(output, t, ct, ct_sz, 17);
    if ((rc = obs.record(eng.prop, t)) != MPCI_OK) return rc;
  }
  field_fl.close();
  f_pop.close();
  return 0;
}

}  // namespace tdse
