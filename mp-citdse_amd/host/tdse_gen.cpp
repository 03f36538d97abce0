// tdse_gen: writes seeded synthetic inputs in the HDF5 layout the tdse
// program reads (the files the structure stage produces):
//   -e 1: dat/<pot><l>.h5:/En and dat/<pot><l><l+1><g>.h5:/d_if
//         (h1e.cpp:22-58, dmx1e.cpp:94-130)
//   -e 2: dat/<pot>CI<L>.h5:/En_CI, dat/<pot>CI_<L><L+1><g>.h5:/CI_dip
//         (w2e.cpp:67-71, 115-120), dat/<pot>0.h5:/En (the 1e threshold
//         pes reads, pes.cpp:168-176) and placeholder dat/cfg-<L>.inp files
//         (tdse copies them, tdse_main.cpp:122-126).
//   -w:   with -e 2, instead of the CI files, the product-basis inputs the
//         w2e stage reads (w2e.cpp:25-42, 93-102): dat/<pot>2_<L>En.h5:/e_2e,
//         dat/<pot>V12_<L>.h5:/V_12 (packed upper triangle) and
//         dat/<pot>2_<L><L+1><g>.h5:/d_if, plus the threshold and cfg files;
//         with -e 1, the radial functions the d1e stage reads
//         (dmx1e.cpp:34-62): dat/<pot>_w1e<l>.h5 with /Pr_o and /Pr_p
//         (n x q samples and derivatives, every -s size equal) and /Qr_o,
//         /Qw_o (q quadrature nodes and weights, -q q), plus
//         dat/<pot><l>.h5:/En energies for the later tdse run.
// The values come from host/synth.c, so `mpci.synth.system_arrays` (and
// `mpci.synth.w2e_arrays` for -w) with the same sizes and seed reproduces
// them bit for bit.
//
// usage: tdse_gen -e 1|2 -p <pot> -s n0,n1,... -g v|l [-d <dir>]
//                 [-r <seed>] [-t <threshold>] [-w] [-q <quadrature size>]
#include <unistd.h>

#include <algorithm>
#include <cmath>
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

namespace {

// Seeded product-basis inputs of the w2e stage: e_2e ~ U[-2.9, 5] sorted,
// V_12 entries ~ N(0, (0.2 / sqrt(n))^2) as the packed upper triangle, d_if
// entries ~ N(0, 1 / n_{L+1}).  Streams 5000 + L, 6000 + L and 7000 + L
// of the generator in host/synth.c.
int mpci_synth_w2e(uint64_t seed, int nb, const int *s,
                   std::vector<std::vector<double>> *ens,
                   std::vector<std::vector<double>> *v12,
                   std::vector<std::vector<double>> *dip) {
  ens->assign(nb, {});
  v12->assign(nb, {});
  dip->assign(nb > 1 ? nb - 1 : 0, {});
  for (int L = 0; L < nb; ++L) {
    const long long n = s[L];
    (*ens)[L].resize(n);
    mpci_synth_uniform(seed, 5000 + L, n, -2.9, 5.0, (*ens)[L].data());
    std::sort((*ens)[L].begin(), (*ens)[L].end());
    (*v12)[L].resize(n * (n + 1) / 2);
    mpci_synth_normal(seed, 6000 + L, n * (n + 1) / 2, 0.2 / std::sqrt(n),
                      (*v12)[L].data());
  }
  for (int L = 0; L + 1 < nb; ++L) {
    (*dip)[L].resize((size_t)s[L] * s[L + 1]);
    mpci_synth_normal(seed, 7000 + L, (long long)s[L] * s[L + 1],
                      1.0 / std::sqrt((double)s[L + 1]), (*dip)[L].data());
  }
  return 0;
}

// Seeded d1e inputs: quadrature nodes x ~ U[0.05, 60] sorted and weights
// w ~ U[0, 0.2] (streams 9000, 9001); per l, function samples and
// derivative samples ~ N(0, 0.1^2) (streams 8000 + l, 8500 + l).
void mpci_synth_w1e_quadrature(uint64_t seed, int qsz, double *qx,
                               double *qw) {
  mpci_synth_uniform(seed, 9000, qsz, 0.05, 60.0, qx);
  std::sort(qx, qx + qsz);
  mpci_synth_uniform(seed, 9001, qsz, 0.0, 0.2, qw);
}

void mpci_synth_w1e_functions(uint64_t seed, int l, int nen, int qsz,
                              double *po, double *pp) {
  mpci_synth_normal(seed, 8000 + l, (long long)nen * qsz, 0.1, po);
  mpci_synth_normal(seed, 8500 + l, (long long)nen * qsz, 0.1, pp);
}

}  // namespace

int main(int argc, char *argv[]) {
  int e_num = 0;
  std::string pot, dir = "dat", sizes;
  char gauge = 'v';
  uint64_t seed = 20261202;
  double threshold = -2.0;
  bool w2e_inputs = false;
  int qsz = 64;
  for (;;) {
    int c = getopt(argc, argv, "e:p:s:g:d:r:t:wq:");
    if (c == -1) break;
    switch (c) {
    case 'w': w2e_inputs = true; break;
    case 'q': qsz = std::atoi(optarg); break;
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
  if (w2e_inputs && e_num == 1) {
    for (int n : s)
      if (n != s[0]) {
        std::cerr << "tdse_gen: -e 1 -w needs equal sizes (one basis size "
                     "for every l)\n";
        return 2;
      }
    if (qsz < 1) {
      std::cerr << "tdse_gen: -q must be >= 1\n";
      return 2;
    }
    try {
      std::filesystem::create_directories(dir);
      const std::string base = dir + "/" + pot;
      const int nen = s[0];
      std::vector<double> qx(qsz), qw(qsz);
      mpci_synth_w1e_quadrature(seed, qsz, qx.data(), qw.data());
      std::vector<double> po((size_t)nen * qsz), pp((size_t)nen * qsz),
          en(nen);
      for (int l = 0; l < nb; ++l) {
        mpci_synth_w1e_functions(seed, l, nen, qsz, po.data(), pp.data());
        std::vector<mpci::H5Out> sets = {{"Pr_o", po.data(), {nen, qsz}},
                                         {"Pr_p", pp.data(), {nen, qsz}}};
        if (l == 0) {
          sets.push_back({"Qr_o", qx.data(), {qsz}});
          sets.push_back({"Qw_o", qw.data(), {qsz}});
        }
        mpci::h5_write_all(base + "_w1e" + std::to_string(l) + ".h5", sets);
        mpci_synth_uniform(seed, 9500 + l, nen, -0.5, 5.0, en.data());
        std::sort(en.begin(), en.end());
        mpci::h5_write_1d(base + std::to_string(l) + ".h5", "En", en.data(),
                          nen);
      }
    } catch (const std::exception &ex) {
      std::cerr << "tdse_gen: " << ex.what() << "\n";
      return 1;
    }
    return 0;
  }
  if (w2e_inputs) {
    try {
      std::filesystem::create_directories(dir);
      const std::string base = dir + "/" + pot;
      std::vector<std::vector<double>> ens, v12, dip;
      if (mpci_synth_w2e(seed, nb, s.data(), &ens, &v12, &dip) != 0) {
        std::cerr << "tdse_gen: synthesis failed\n";
        return 1;
      }
      for (int L = 0; L < nb; ++L) {
        const std::string l = std::to_string(L);
        mpci::h5_write_1d(base + "2_" + l + "En.h5", "e_2e", ens[L].data(),
                          s[L]);
        mpci::h5_write_1d(base + "V12_" + l + ".h5", "V_12", v12[L].data(),
                          (long)v12[L].size());
      }
      for (int L = 0; L + 1 < nb; ++L) {
        const std::string ll =
            std::to_string(L) + std::to_string(L + 1) + gauge;
        mpci::h5_write_2d(base + "2_" + ll + ".h5", "d_if", dip[L].data(),
                          s[L], s[L + 1]);
      }
      mpci::h5_write_1d(base + "0.h5", "En", &threshold, 1);
      for (int L = 0; L < nb; ++L) {
        std::ofstream f(dir + "/cfg-" + std::to_string(L) + ".inp");
        f << "# synthetic configuration list placeholder for L=" << L << "\n";
      }
    } catch (const std::exception &ex) {
      std::cerr << "tdse_gen: " << ex.what() << "\n";
      return 1;
    }
    return 0;
  }
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
