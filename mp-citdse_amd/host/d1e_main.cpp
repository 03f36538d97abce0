// d1e drop-in: one-electron dipole blocks on the GPU, reading and writing
// the files of MP-CITDSE's d1e program (src/d1e/dmx1e_main.cpp:6-33,
// src/d1e/dmx1e.cpp:3-139):
//
//   d1e -f <yaml>
//
// It reads the radial functions sampled on the outer quadrature,
// dat/<pot>_w1e<l>.h5:/Pr_o (and /Pr_p, their derivatives, for the velocity
// gauge), with the nodes and weights /Qr_o, /Qw_o of dat/<pot>_w1e0.h5, and
// writes dat/<pot><l><l+1><g>.h5:/d_if for l < l_max -- the dipole input of
// `tdse -e 1`.  The arithmetic runs in lib/libmpci_ci.so (include/mpci_ci.h).
#include <unistd.h>

#include <iostream>
#include <string>
#include <vector>

#include "../../include/mpci_ci.h"
#include "h5io.hpp"
#include "yaml_lite.hpp"

namespace {

// dmx1e::readConfig (src/d1e/dmx1e.cpp:3-22): the same keys and echo.
void read_config(const std::string &file, std::string &pot, int &qsz,
                 char &gauge, int &l_max) {
  mpci::YamlDoc settings = mpci::YamlDoc::load_file(file);
  pot = settings.str("Global_Settings", "potential");
  std::cout << "Core Potential:                              " << pot
            << std::endl;
  qsz = settings.as_int("Global_Settings", "Outer_quadrature_size");
  std::cout << "No. of outer quadrature points:              " << qsz
            << std::endl;
  gauge = settings.as_char("Global_Settings", "gauge");
  std::cout << "Gauge type ('l' length/'v' velocity):        " << gauge
            << std::endl;
  l_max = settings.as_int("Basis_Settings", "l_max");
  std::cout << "Maximum one electron angular momentum:       " << l_max
            << std::endl;
}

}  // namespace

int main(int argc, char *argv[]) {
  std::string opt_file;
  for (;;) {
    switch (getopt(argc, argv, "hf:")) {
    case 'h':
      std::cout << "Program for calculating the 1e dipole matrix\n"
                << "elements <nl|d_(v/l)|n'l'>\n"
                << "-f <path> yaml input file with the input settings\n";
      return -1;
    case 'f':
      opt_file = optarg;
      continue;
    }
    break;
  }
  try {
    std::string pot;
    int qsz = 0, l_max = 0;
    char gauge = 'v';
    read_config(opt_file, pot, qsz, gauge, l_max);
    if (qsz < 1) throw std::runtime_error("Outer_quadrature_size must be >= 1");
    if (gauge != 'v' && gauge != 'l')
      throw std::runtime_error("gauge must be 'v' or 'l'");
    const std::string prefix = "dat/" + pot;
    // dmx1e::genDipole (dmx1e.cpp:24-139): functions of every l, the nodes
    // and weights of the l = 0 file
    std::vector<std::vector<double>> wfn(l_max + 1);
    wfn[0] = mpci::h5_read_all(prefix + "_w1e0.h5", "Pr_o");
    const long lc_sz = (long)wfn[0].size();
    if (lc_sz % qsz != 0)
      throw std::runtime_error("Pr_o does not hold whole functions of "
                               "Outer_quadrature_size points");
    const int nen = (int)(lc_sz / qsz);
    // whole datasets, whatever their rank, as the reference reads them
    auto read_n = [&](const std::string &file, const char *name, long n) {
      std::vector<double> v = mpci::h5_read_all(file, name);
      if ((long)v.size() != n)
        throw std::runtime_error(std::string(name) + " in " + file +
                                 " holds " + std::to_string(v.size()) +
                                 " values, expected " + std::to_string(n));
      return v;
    };
    std::vector<double> qx = read_n(prefix + "_w1e0.h5", "Qr_o", qsz);
    std::vector<double> qw = read_n(prefix + "_w1e0.h5", "Qw_o", qsz);
    for (int l = 1; l <= l_max; ++l)
      wfn[l] = read_n(prefix + "_w1e" + std::to_string(l) + ".h5", "Pr_o",
                      lc_sz);
    std::vector<double> D((size_t)nen * nen), wfnp;
    for (int l = 0; l < l_max; ++l) {
      if (gauge == 'v')
        wfnp = read_n(prefix + "_w1e" + std::to_string(l + 1) + ".h5", "Pr_p",
                      lc_sz);
      if (mpci_d1e_dipoles(nen, qsz, l, gauge, qx.data(), qw.data(),
                           wfn[l].data(), wfn[l + 1].data(),
                           gauge == 'v' ? wfnp.data() : nullptr,
                           D.data()) != 0)
        throw std::runtime_error(std::string("1e dipoles: ") +
                                 mpci_ci_last_error());
      mpci::h5_write_2d(prefix + std::to_string(l) + std::to_string(l + 1) +
                            gauge + ".h5",
                        "d_if", D.data(), nen, nen);
    }
  } catch (const std::exception &ex) {
    std::cerr << "d1e: " << ex.what() << "\n";
    return 1;
  }
  return 0;
}
