// tdse_gen: writes seeded synthetic inputs in the HDF5 layout the tdse
// program reads (the files the structure stage produces):
//   -e 1: dat/<pot><l>.h5:/En and dat/<pot><l><l+1><g>.h5:/d_if
//         (h1e.cpp:22-58, dmx1e.cpp:94-130)
//   -e 2: dat/<pot>CI<L>.h5:/En_CI, dat/<pot>CI_<L><L+1><g>.h5:/CI_dip
//         (w2e.cpp:67-71, 115-120), dat/<pot>0.h5:/En (the 1e threshold
//         pes reads, pes.cpp:168-176) and placeholder dat/cfg-<L>.inp files
//         (tdse copies them, tdse_main.cpp:122-126).
// The values come from host/synth.c, so `mpci.synth.system_arrays` with
// the same sizes and seed reproduces them bit for bit.
//
// usage: tdse_gen -e 1|2 -p <pot> -s n0,n1,... -g v|l [-d <dir>]
//                 [-r <seed>] [-t <threshold>]
#include <unistd.h>

#include <cstdint>
#include <cstdlib>
#include <filesystem>
#include <fstream>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "h5io.hpp"
#include "synth.h"

int main(int argc, char *argv[]) {
  int e_num = 0;
  std::string pot, dir = "dat", sizes;
  char gauge = 'v';
  uint64_t seed = 20261202;
  double threshold = -2.0;
  for (;;) {
    int c = getopt(argc, argv, "e:p:s:g:d:r:t:");
    if (c == -1) break;
    switch (c) {
    case 'e': e_num = std::atoi(optarg); break;
    case 'p': pot = optarg; break;
    case 's': sizes = optarg; break;
    case 'g': gauge = optarg[0]; break;
    case 'd': dir = optarg; break;
    case 'r': seed = std::strtoull(optarg, nullptr, 10); break;
    case 't': threshold = std::atof(optarg); break;
    default:
      std::cerr << "usage: tdse_gen -e 1|2 -p pot -s n0,n1,... -g v|l "
                   "[-d dir] [-r seed] [-t threshold]\n";
      return 2;
    }
  }
  std::vector<int> s;
  std::stringstream ss(sizes);
  std::string item;
  while (std::getline(ss, item, ',')) s.push_back(std::atoi(item.c_str()));
  if ((e_num != 1 && e_num != 2) || pot.empty() || s.empty() ||
      (gauge != 'v' && gauge != 'l')) {
    std::cerr << "tdse_gen: need -e 1|2, -p, -s and -g v|l\n";
    return 2;
  }
  for (int n : s)
    if (n < 1) {
      std::cerr << "tdse_gen: block sizes must be >= 1\n";
      return 2;
    }
  const int nb = (int)s.size();
  std::vector<std::vector<double>> E(nb), D(nb > 1 ? nb - 1 : 0);
  std::vector<double *> Ep(nb), Dp(nb > 1 ? nb - 1 : 1, nullptr);
  for (int L = 0; L < nb; ++L) {
    E[L].resize(s[L]);
    Ep[L] = E[L].data();
  }
  for (int L = 0; L + 1 < nb; ++L) {
    D[L].resize((size_t)s[L] * s[L + 1]);
    Dp[L] = D[L].data();
  }
  if (mpci_synth_system(seed, nb, s.data(), -2.9, 5.0, Ep.data(),
                        Dp.data()) != 0) {
    std::cerr << "tdse_gen: synthesis failed\n";
    return 1;
  }
  try {
    std::filesystem::create_directories(dir);
    const std::string base = dir + "/" + pot;
    for (int L = 0; L < nb; ++L) {
      const std::string l = std::to_string(L);
      if (e_num == 1)
        mpci::h5_write_1d(base + l + ".h5", "En", E[L].data(), s[L]);
      else
        mpci::h5_write_1d(base + "CI" + l + ".h5", "En_CI", E[L].data(), s[L]);
    }
    for (int L = 0; L + 1 < nb; ++L) {
      const std::string ll = std::to_string(L) + std::to_string(L + 1) + gauge;
      if (e_num == 1)
        mpci::h5_write_2d(base + ll + ".h5", "d_if", D[L].data(), s[L],
                          s[L + 1]);
      else
        mpci::h5_write_2d(base + "CI_" + ll + ".h5", "CI_dip", D[L].data(),
                          s[L], s[L + 1]);
    }
    if (e_num == 2) {
      mpci::h5_write_1d(base + "0.h5", "En", &threshold, 1);
      for (int L = 0; L < nb; ++L) {
        std::ofstream f(dir + "/cfg-" + std::to_string(L) + ".inp");
        f << "# synthetic configuration list placeholder for L=" << L << "\n";
      }
    }
  } catch (const std::exception &ex) {
    std::cerr << "tdse_gen: " << ex.what() << "\n";
    return 1;
  }
  return 0;
}
