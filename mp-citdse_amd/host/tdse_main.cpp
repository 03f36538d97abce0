// `tdse` drop-in: the command line of src/tdse/tdse_main.cpp:9-201
// (`tdse -e 1|2 -f <yaml> -o <dir> [-s <ct file>] [-n N -l L]`), the same
// dat/ file prefixes, settings copy and prompt, restart handling, step count
// and shape x gauge dispatch, with the propagation on the MI355X engine.
//
// Differences, all for inputs on which the reference aborts: a missing
// setting, HDF5 file or dataset, an existing settings.yaml copy on restart,
// or fewer than 10 steps (a division by zero at tdse.cpp:206) give a
// message and a non-zero exit instead of a crash; fewer than 10 steps then
// write a coefficient file after every step.
// Extension (optional keys, absent from the reference's inputs):
// Propagator_Settings.mode: "adaptive" with tol | atol / rtol, and
// Propagator_Settings.full_precision: true (17-digit <pot>_ct17_<t>.dat).
#include <unistd.h>

#include <cmath>
#include <cstdlib>
#include <filesystem>
#include <iostream>
#include <string>
#include <vector>

#include "tdse_host.hpp"

namespace {

int run(int argc, char *argv[]) {
  std::string opt_file;
  std::string out_dir;
  int L_max, cycles, e_num = 0;
  int pop_n = 0, pop_l = 0;
  std::string pot, init_ct = "ground";
  std::string base, option;
  std::string en_pot, dip_pot;
  std::string en_set, dip_set;
  std::string o_file_prefix;
  char gauge;
  std::string shape;
  double dt, w, Io, cepd;
  std::vector<int> state_sz;

  for (;;) {
    switch (getopt(argc, argv, "he:f:o:s:n:l:")) {
    case 'h':
      std::cout << "Program for propagating either the 1- or 2-e^- tdse\n"
                << "-e <1 or 2> the number of electrons in the tdse\n"
                << "-f <path> yaml input file with the input settings\n"
                << "-o <path> output directory for the coefficient files\n"
                << "-s <path> file with the start coefficients\n"
                << "-n <int> primary quantum number of state to track\n"
                << "-l <int> angular quantum number of state to track\n";
      return -1;
    case 'e':
      e_num = std::stoi(optarg);
      if (e_num != 1 && e_num != 2) {
        std::cout << "Invalid number of electrons, use 1 or 2\n";
        return -1;
      }
      continue;
    case 'f':
      opt_file = optarg;
      continue;
    case 'o':
      out_dir = optarg;
      continue;
    case 's':
      init_ct = optarg;
      continue;
    case 'n':
      pop_n = std::stoi(optarg);
      if (pop_n < 1) {
        std::cout << "Invalid primary quantum number -n, use int > 0\n";
        return -1;
      }
      --pop_n;
      continue;
    case 'l':
      pop_l = std::stoi(optarg);
      if (pop_n < 0) {
        std::cout << "Invalid angular momentum number -l, use int > 0\n";
        return -1;
      }
      continue;
    }
    break;
  }

  const std::filesystem::path out_path{out_dir};
  if (!std::filesystem::exists(out_path))
    std::filesystem::create_directory(out_path);

  auto set_file = out_dir + "/settings.yaml";
  const std::filesystem::path file_path{set_file};
  std::filesystem::copy_options cpopt = std::filesystem::copy_options::none;
  if (std::filesystem::exists(file_path) && init_ct.compare("ground") == 0) {
    char ovw = 'n';
    std::cout << "File " << set_file << " exists. Overwrite? [y/n]\n";
    std::cin >> ovw;
    switch (ovw) {
    case 'y':
      cpopt = std::filesystem::copy_options::update_existing;
      break;
    case 'n':
      std::cout << out_dir
                << " contians old data, please choose a different output"
                << " directory!\n";
      return -1;
    }
  }

  std::filesystem::copy_file(opt_file, set_file, cpopt);

  switch (e_num) {
  case 1:
    base = "Basis_Settings";
    option = "l_max";
    break;
  case 2:
    base = "Global_Settings";
    option = "L_max";
    break;
  default:
    std::cout << "Invalid number of electrons, use 1 or 2\n";
    return -1;
  }

  tdrd::Extensions ext;
  tdrd::readConfig(opt_file, pot, base, option, L_max, gauge, state_sz, dt,
                   shape, w, Io, cepd, cycles, &ext);

  switch (e_num) {
  case 1:
    en_pot = "dat/" + pot;
    en_set = "En";
    dip_pot = "dat/" + pot;
    dip_set = "d_if";
    break;
  case 2:
    en_pot = "dat/" + pot + "CI";
    en_set = "En_CI";
    dip_pot = "dat/" + pot + "CI_";
    dip_set = "CI_dip";
    for (int L = 0; L <= L_max; ++L)
      std::filesystem::copy_file("dat/cfg-" + std::to_string(L) + ".inp",
                                 out_dir + "/cfg-" + std::to_string(L) +
                                     ".inp",
                                 cpopt);
    break;
  }

  o_file_prefix = out_dir + "/" + pot;

  int ct_sz = 0;
  std::vector<int> offs;
  stvupt eig;
  stvupt dipoles;
  eig.push_back(std::make_unique<std::vector<double>>());
  for (auto i = 0; i < L_max; ++i) {
    eig.push_back(std::make_unique<std::vector<double>>());
    dipoles.push_back(std::make_unique<std::vector<double>>());
  }

  tdrd::readEnergies(en_pot, en_set, L_max, ct_sz, state_sz, offs, eig);
  tdrd::readDipoles(dip_pot, dip_set, gauge, L_max, state_sz, dipoles);

  std::vector<double> ct(ct_sz * 2);
  ct[0] = 1.0;
  double t = 0.0;

  if (init_ct.compare("ground") != 0) {
    t = tdrd::timeFromCtName(init_ct);
    std::cout << "Start time (a.u.): " << t << "\n";
    tdrd::readInitCt(init_ct, ct_sz, ct);
  }

  // conv::En_ev_au_ (src/include/au.hpp:30-31)
  double tau = mpci_period(w * (1. / 27.211396181), cycles);
  int steps = (tau - t) / dt;

  double auIo, auw;
  mpci_to_au(Io, w, &auIo, &auw);
  mpci_pulse pars;

  if (pop_l < 0 || pop_l > L_max || pop_n >= state_sz[pop_l]) {
    std::cout << "tracked state (-n, -l) is outside the basis\n";
    return -1;
  }
  if (gauge != 'v' && gauge != 'l') {
    std::cout << "Invalid gauge, use 'v' or 'l'\n";
    return -1;
  }
  int rc;
  if (shape.compare("sine") == 0) {
    rc = mpci_pulse_sine(auIo, auw, 0.0, cycles, gauge, &pars);
  } else if (shape.compare("gaussian") == 0) {
    rc = mpci_pulse_gauss(auIo, auw, tau, cycles, 0.0, gauge, &pars);
  } else {
    std::cout << "Invalid envelope shape, use \"sine\" or \"gaussian\"\n";
    return -1;
  }
  if (rc != MPCI_OK) {
    std::cerr << "tdse: " << mpci_last_error() << "\n";
    return 1;
  }

  if (ext.mode == "adaptive")
    rc = tdse::propAdaptive(o_file_prefix, gauge, L_max, t, tau, dt, pop_n,
                            pop_l, pars, ct_sz, offs, state_sz, eig, dipoles,
                            ct, ext);
  else if (gauge == 'v')
    rc = tdse::propV(o_file_prefix, L_max, t, dt, steps, pop_n, pop_l, pars,
                     ct_sz, offs, state_sz, eig, dipoles, ct, ext);
  else
    rc = tdse::propL(o_file_prefix, L_max, t, dt, steps, pop_n, pop_l, pars,
                     ct_sz, offs, state_sz, eig, dipoles, ct, ext);
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
