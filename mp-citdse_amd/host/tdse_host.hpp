// Host-side mirror of MP-CITDSE's tdse interface (namespaces tdrd and tdse,
// src/tdse/include/td_read.hpp:38-109 and tdse.hpp:37-102) on top of the
// MI355X engine's C ABI (include/mpci_tdse.h).  Same names, argument meaning
// and output files; compute goes through the HIP engine only.
#pragma once

#include <functional>
#include <memory>
#include <string>
#include <vector>

#include "../../include/mpci_tdse.h"

using stvupt = std::vector<std::unique_ptr<std::vector<double>>>;

namespace tdrd {

// Propagator options that the reference does not have; read from optional
// Propagator_Settings keys so existing YAML files keep their meaning.
struct Extensions {
  std::string mode = "fixed";  // "fixed" (reference) or "adaptive"
  double atol = 1e-10;
  double rtol = 1e-10;
  bool full_precision = false;  // extra <pot>_ct17_<t>.dat files
  // <pot>_obs.dat: pes's ground population, yield, excited population and
  // norm (src/pes/pes.cpp:138-146, 199-213) at every coefficient-file time
  bool observables = false;
  int obs_electrons = 2;       // which of pes's definitions (1 or 2)
  double obs_threshold = 0.0;  // two electrons: dat/<pot>0.h5:/En[0]
};

// td_read.cpp:3-48 (echoes the settings to stdout exactly as the reference).
int readConfig(std::string file, std::string &pot, std::string set_base,
               std::string option, int &L_max, char &gauge,
               std::vector<int> &state_sz, double &timestep,
               std::string &shape, double &w, double &Io, double &cepd,
               int &cycles, Extensions *ext = nullptr);
// td_read.cpp:50-86
int readEnergies(std::string pot, std::string setname, int L_max, int &ct_sz,
                 std::vector<int> &state_sz, std::vector<int> &offs,
                 stvupt &eig);
// td_read.cpp:88-120
int readDipoles(std::string pot, std::string setname, char gauge, int L_max,
                std::vector<int> &state_sz, stvupt &dipoles);
// td_read.cpp:122-144
int readInitCt(std::string file, int ct_sz, std::vector<double> &ct);
// Start time from a restart file name (tdse_main.cpp:153-157).
double timeFromCtName(const std::string &path);

}  // namespace tdrd

namespace tdse {

using fieldFcn = std::function<double(const mpci_pulse &, double)>;

// tdse.cpp:146-227 / 229-310: the same loop, files and cadence, with the
// RHS and the RKF78 step on the GPU.  Returns 0, or a negative code after
// printing the engine's message.
int propV(std::string output, int L_max, double t, double dt, int steps,
          int pop_n, int pop_l, const mpci_pulse &pars, int ct_sz,
          std::vector<int> &offs, std::vector<int> &state_sz, stvupt &eig,
          stvupt &dipoles, std::vector<double> &ct,
          const tdrd::Extensions &ext = tdrd::Extensions());
int propL(std::string output, int L_max, double t, double dt, int steps,
          int pop_n, int pop_l, const mpci_pulse &pars, int ct_sz,
          std::vector<int> &offs, std::vector<int> &state_sz, stvupt &eig,
          stvupt &dipoles, std::vector<double> &ct,
          const tdrd::Extensions &ext = tdrd::Extensions());

// Adaptive extension: integrates to t_end with the embedded-error control.
int propAdaptive(std::string output, char gauge, int L_max, double t,
                 double t_end, double h0, int pop_n, int pop_l,
                 const mpci_pulse &pars, int ct_sz, std::vector<int> &offs,
                 std::vector<int> &state_sz, stvupt &eig, stvupt &dipoles,
                 std::vector<double> &ct, const tdrd::Extensions &ext);

// ct file writer (tdse.cpp:152-158): default ostream precision.
void writeCt(const std::string &output, double t,
             const std::vector<double> &ct, int ct_sz, int precision = 0);

}  // namespace tdse
