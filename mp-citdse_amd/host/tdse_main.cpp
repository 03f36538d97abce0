// `tdse` drop-in: the command-line and file contract of MP-CITDSE's tdse
// program (src/tdse/tdse_main.cpp:9-201) in front of the MI355X engine.
//
// Kept as the contract: the option string "he:f:o:s:n:l:", the help text
// and the user-facing messages, the settings.yaml copy with its overwrite
// prompt, the dat/ input names, the cfg-<L>.inp copies of a two-electron
// run, the "Start time" echo of a restart, steps = int((tau - t) / dt), and
// a carrier-envelope phase of 0 handed to the pulse set-up.
//
// Own behaviour where the reference aborts or misbehaves: a missing -e, a
// negative -l (the reference tests -n there, tdse_main.cpp:61), a tracked
// state outside the basis, an unknown gauge, a missing setting, HDF5 file
// or dataset, and a restart into a directory that already holds a
// settings.yaml copy all end with a message and a non-zero exit.
//
// Extensions, on optional Propagator_Settings keys that the reference's
// inputs do not have: mode: "adaptive" with tol | atol / rtol;
// full_precision: true (17-digit <pot>_ct17_<t>.dat files beside the
// ct files); observables: true (a <pot>_obs.dat side file with pes's ground
// population, norm, yield and excited population at every ct-file time).
#include <unistd.h>

#include <filesystem>
#include <iostream>
#include <optional>
#include <string>
#include <vector>

#include "h5io.hpp"
#include "tdse_host.hpp"

namespace fs = std::filesystem;

namespace {

// The reference returns -1 from main for help and usage errors.
constexpr int kUsageExit = -1;

constexpr char kHelp[] =
    "Program for propagating either the 1- or 2-e^- tdse\n"
    "-e <1 or 2> the number of electrons in the tdse\n"
    "-f <path> yaml input file with the input settings\n"
    "-o <path> output directory for the coefficient files\n"
    "-s <path> file with the start coefficients\n"
    "-n <int> primary quantum number of state to track\n"
    "-l <int> angular quantum number of state to track\n";

// One invocation, as the command line states it.
struct Request {
  int electrons = 0;       // -e
  std::string settings;    // -f
  std::string out_dir;     // -o
  std::string start_file;  // -s; empty: start in the lowest state
  int track_n = 0;         // -n, stored 0-based within its block
  int track_l = 0;         // -l
};

// Reads the options.  Parsing stops at the first option it does not know,
// as the reference's loop does.  Returns true when the run goes on.
bool parse_request(int argc, char *argv[], Request &rq) {
  for (int opt; (opt = getopt(argc, argv, "he:f:o:s:n:l:")) != -1;) {
    if (opt == 'h') {
      std::cout << kHelp;
      return false;
    }
    if (opt != 'e' && opt != 'f' && opt != 'o' && opt != 's' && opt != 'n' &&
        opt != 'l')
      break;
    const std::string value = optarg;
    switch (opt) {
    case 'e':
      rq.electrons = std::stoi(value);
      if (rq.electrons != 1 && rq.electrons != 2) {
        std::cout << "Invalid number of electrons, use 1 or 2\n";
        return false;
      }
      break;
    case 'f':
      rq.settings = value;
      break;
    case 'o':
      rq.out_dir = value;
      break;
    case 's':
      rq.start_file = value;
      break;
    case 'n':
      rq.track_n = std::stoi(value) - 1;
      if (rq.track_n < 0) {
        std::cout << "Invalid primary quantum number -n, use int > 0\n";
        return false;
      }
      break;
    case 'l':
      rq.track_l = std::stoi(value);
      if (rq.track_l < 0) {
        std::cout << "Invalid angular momentum number -l, use int > 0\n";
        return false;
      }
      break;
    }
  }
  if (rq.electrons == 0) {
    std::cout << "Invalid number of electrons, use 1 or 2\n";
    return false;
  }
  return true;
}

// The output directory and its settings.yaml copy.  A fresh start over an
// existing copy asks before replacing it; the answer also decides how the
// cfg-<L>.inp files are copied later.  Empty when the user declines.
std::optional<fs::copy_options> prepare_output(const Request &rq) {
  if (!fs::exists(fs::path(rq.out_dir))) fs::create_directory(rq.out_dir);
  const std::string copy = rq.out_dir + "/settings.yaml";
  fs::copy_options mode = fs::copy_options::none;
  if (rq.start_file.empty() && fs::exists(fs::path(copy))) {
    std::cout << "File " << copy << " exists. Overwrite? [y/n]\n";
    char answer = 'n';
    std::cin >> answer;
    if (answer == 'n') {
      std::cout << rq.out_dir << " contians old data, please choose a "
                << "different output directory!\n";
      return std::nullopt;
    }
    if (answer == 'y') mode = fs::copy_options::update_existing;
  }
  fs::copy_file(rq.settings, copy, mode);
  return mode;
}

// Where a run's inputs live, relative to the working directory: the
// one-electron files of h1e/d1e or the CI files of w2e.
struct Inputs {
  const char *lmax_section, *lmax_key;  // the key that holds L_max
  std::string energy_prefix, energy_set;
  std::string dipole_prefix, dipole_set;
};

Inputs inputs_for(int electrons) {
  if (electrons == 1) return {"Basis_Settings", "l_max", "", "", "", ""};
  return {"Global_Settings", "L_max", "", "", "", ""};
}

void name_inputs(Inputs &in, int electrons, const std::string &pot) {
  const std::string dat = "dat/" + pot;
  if (electrons == 1) {
    in.energy_prefix = in.dipole_prefix = dat;
    in.energy_set = "En";
    in.dipole_set = "d_if";
  } else {
    in.energy_prefix = dat + "CI";
    in.dipole_prefix = dat + "CI_";
    in.energy_set = "En_CI";
    in.dipole_set = "CI_dip";
  }
}

// The configuration lists a two-electron run is built from travel with
// its output.
void copy_configurations(const Request &rq, int L_max,
                         fs::copy_options mode) {
  for (int L = 0; L <= L_max; ++L) {
    const std::string name = "cfg-" + std::to_string(L) + ".inp";
    fs::copy_file("dat/" + name, rq.out_dir + "/" + name, mode);
  }
}

// Everything readConfig returns, in one place.
struct Settings {
  std::string pot, shape;
  int L_max = 0, cycles = 0;
  char gauge = 0;
  std::vector<int> sizes;
  double dt = 0, w_eV = 0, Io_W = 0, cepd = 0;
  tdrd::Extensions ext;
};

// The pulse of the settings with the phase set to 0, as the reference
// hands it over (tdse_main.cpp:165-193).  False after a message when the
// shape is unknown.
bool make_pulse(const Settings &st, double tau, mpci_pulse &pulse) {
  double Io_au = 0, w_au = 0;
  mpci_to_au(st.Io_W, st.w_eV, &Io_au, &w_au);
  int rc;
  if (st.shape == "sine") {
    rc = mpci_pulse_sine(Io_au, w_au, 0.0, st.cycles, st.gauge, &pulse);
  } else if (st.shape == "gaussian") {
    rc = mpci_pulse_gauss(Io_au, w_au, tau, st.cycles, 0.0, st.gauge, &pulse);
  } else {
    std::cout << "Invalid envelope shape, use \"sine\" or \"gaussian\"\n";
    return false;
  }
  if (rc != MPCI_OK)
    throw std::runtime_error(std::string("pulse set-up: ") +
                             mpci_last_error());
  return true;
}

int run(int argc, char *argv[]) {
  Request rq;
  if (!parse_request(argc, argv, rq)) return kUsageExit;
  const std::optional<fs::copy_options> copy_mode = prepare_output(rq);
  if (!copy_mode) return kUsageExit;

  Inputs in = inputs_for(rq.electrons);
  Settings st;
  tdrd::readConfig(rq.settings, st.pot, in.lmax_section, in.lmax_key,
                   st.L_max, st.gauge, st.sizes, st.dt, st.shape, st.w_eV,
                   st.Io_W, st.cepd, st.cycles, &st.ext);
  name_inputs(in, rq.electrons, st.pot);
  if (rq.electrons == 2) copy_configurations(rq, st.L_max, *copy_mode);

  // blocks 0..L_max of energies, pairs 0..L_max-1 of dipoles
  stvupt energies, dipoles;
  for (int L = 0; L <= st.L_max; ++L) {
    energies.emplace_back(new std::vector<double>);
    if (L < st.L_max) dipoles.emplace_back(new std::vector<double>);
  }
  int ct_sz = 0;
  std::vector<int> offsets;
  tdrd::readEnergies(in.energy_prefix, in.energy_set, st.L_max, ct_sz,
                     st.sizes, offsets, energies);
  tdrd::readDipoles(in.dipole_prefix, in.dipole_set, st.gauge, st.L_max,
                    st.sizes, dipoles);

  // the lowest state of block 0, or the coefficients of a restart file
  std::vector<double> ct(2 * (size_t)ct_sz, 0.0);
  ct[0] = 1.0;
  double t = 0.0;
  if (!rq.start_file.empty()) {
    t = tdrd::timeFromCtName(rq.start_file);
    std::cout << "Start time (a.u.): " << t << "\n";
    tdrd::readInitCt(rq.start_file, ct_sz, ct);
  }

  // pulse length in atomic units: conv::En_ev_au_ (src/include/au.hpp:30)
  const double tau = mpci_period(st.w_eV * (1. / 27.211396181), st.cycles);
  const int steps = static_cast<int>((tau - t) / st.dt);

  if (rq.track_l > st.L_max || rq.track_n >= st.sizes[rq.track_l]) {
    std::cout << "tracked state (-n, -l) is outside the basis\n";
    return kUsageExit;
  }
  if (st.gauge != 'v' && st.gauge != 'l') {
    std::cout << "Invalid gauge, use 'v' or 'l'\n";
    return kUsageExit;
  }
  mpci_pulse pulse{};
  if (!make_pulse(st, tau, pulse)) return kUsageExit;

  if (st.ext.observables) {
    // pes's ionisation threshold: the lowest one-electron l = 0 energy
    // (src/pes/pes.cpp:168-176)
    st.ext.obs_electrons = rq.electrons;
    if (rq.electrons == 2)
      st.ext.obs_threshold =
          mpci::h5_read_prefix_1d("dat/" + st.pot + "0.h5", "En", 1)[0];
  }

  const std::string prefix = rq.out_dir + "/" + st.pot;
  int rc;
  if (st.ext.mode == "adaptive")
    rc = tdse::propAdaptive(prefix, st.gauge, st.L_max, t, tau, st.dt,
                            rq.track_n, rq.track_l, pulse, ct_sz, offsets,
                            st.sizes, energies, dipoles, ct, st.ext);
  else if (st.gauge == 'v')
    rc = tdse::propV(prefix, st.L_max, t, st.dt, steps, rq.track_n,
                     rq.track_l, pulse, ct_sz, offsets, st.sizes, energies,
                     dipoles, ct, st.ext);
  else
    rc = tdse::propL(prefix, st.L_max, t, st.dt, steps, rq.track_n,
                     rq.track_l, pulse, ct_sz, offsets, st.sizes, energies,
                     dipoles, ct, st.ext);
  return rc == 0 ? 0 : 1;
}

}  // namespace

int main(int argc, char *argv[]) {
  try {
    return run(argc, argv);
  } catch (const std::exception &e) {
    std::cerr << "tdse: " << e.what() << "\n";
    return 1;
  }
}
