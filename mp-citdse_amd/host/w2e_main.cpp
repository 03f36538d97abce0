// w2e drop-in: forms the CI two-electron basis on the GPU, reading and
// writing the files of MP-CITDSE's w2e program (src/w2e/w2e_main.cpp:6-40,
// src/w2e/w2e.cpp:3-125):
//
//   w2e -f <yaml>
//
// Per L = 0..L_max it reads dat/<pot>2_<L>En.h5:/e_2e (product-basis
// energies) and dat/<pot>V12_<L>.h5:/V_12 (packed upper triangle of the
// electron-electron interaction), diagonalises H = diag(e_2e) + V_12 and
// writes dat/<pot>CI<L>.h5 (/En_CI eigenvalues, /CI_vecs eigenvectors); per
// L < L_max it reads dat/<pot>2_<L><L+1><g>.h5:/d_if and writes
// dat/<pot>CI_<L><L+1><g>.h5:/CI_dip = V_L^T d_if V_{L+1}.  The arithmetic
// runs in lib/libmpci_ci.so (include/mpci_ci.h).
#include <unistd.h>

#include <iostream>
#include <string>
#include <vector>

#include "../../include/mpci_ci.h"
#include "h5io.hpp"
#include "yaml_lite.hpp"

namespace {

// w2e::readConfig (src/w2e/w2e.cpp:3-15): the same keys and stdout echo.
void read_config(const std::string &file, std::string &pot, char &gauge,
                 int &L_max) {
  mpci::YamlDoc settings = mpci::YamlDoc::load_file(file);
  std::cout << "Global Settings:" << std::endl;
  pot = settings.str("Global_Settings", "potential");
  std::cout << "  Core Potential:                       " << pot << std::endl;
  L_max = settings.as_int("Global_Settings", "L_max");
  std::cout << "  max l:                                " << L_max << std::endl;
  gauge = settings.as_char("Global_Settings", "gauge");
  std::cout << "  Gauge type ('l' length/'v' velocity): " << gauge
            << std::endl;
}

void check(int rc, const char *what) {
  if (rc != 0)
    throw std::runtime_error(std::string(what) + ": " + mpci_ci_last_error());
}

}  // namespace

int main(int argc, char *argv[]) {
  std::string opt_file;
  for (;;) {
    switch (getopt(argc, argv, "hf:")) {
    case 'h':
      std::cout << "Program for forming the CI correlated 2e^- basis\n"
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
    char gauge = 'v';
    int L_max = 0;
    read_config(opt_file, pot, gauge, L_max);
    const std::string prefix = "dat/" + pot;

    // w2e::formCIh0 (w2e.cpp:17-82)
    std::vector<std::vector<double>> vecs(L_max + 1);
    std::vector<int> sz(L_max + 1);
    for (int L = 0; L <= L_max; ++L) {
      const std::string l = std::to_string(L);
      std::vector<double> ens =
          mpci::h5_read_all(prefix + "2_" + l + "En.h5", "e_2e");
      const int n = (int)ens.size();
      std::vector<double> v12 =
          mpci::h5_read_all(prefix + "V12_" + l + ".h5", "V_12");
      if ((long long)v12.size() < (long long)n * (n + 1) / 2)
        throw std::runtime_error("V_12 of L=" + l + " holds fewer than "
                                 "n(n+1)/2 values");
      std::vector<double> eig(n);
      vecs[L].resize((size_t)n * n);
      check(mpci_ci_basis(n, ens.data(), v12.data(), eig.data(),
                          vecs[L].data()),
            "CI basis");
      sz[L] = n;
      mpci::h5_write_all(prefix + "CI" + l + ".h5",
                         {{"En_CI", eig.data(), {n}},
                          {"CI_vecs", vecs[L].data(), {n, n}}});
      if (L == 0)
        std::cout << "Ground state energy (a.u.): " << eig[0] << "\n";
    }

    // w2e::formCIDipoles (w2e.cpp:84-125)
    for (int L = 0; L < L_max; ++L) {
      const std::string ll =
          std::to_string(L) + std::to_string(L + 1) + gauge;
      std::vector<long> dims;
      std::vector<double> d =
          mpci::h5_read_all(prefix + "2_" + ll + ".h5", "d_if", &dims);
      if (dims.size() != 2 || dims[0] != sz[L] || dims[1] != sz[L + 1])
        throw std::runtime_error("d_if of " + ll + " does not match the "
                                 "basis sizes");
      std::vector<double> out(d.size());
      check(mpci_ci_dipoles(sz[L], sz[L + 1], d.data(), vecs[L].data(),
                            vecs[L + 1].data(), out.data()),
            "CI dipoles");
      mpci::h5_write_2d(prefix + "CI_" + ll + ".h5", "CI_dip", out.data(),
                        sz[L], sz[L + 1]);
    }
  } catch (const std::exception &ex) {
    std::cerr << "w2e: " << ex.what() << "\n";
    return 1;
  }
  return 0;
}
