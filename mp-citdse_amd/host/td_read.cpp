// Input side of the tdse drop-in: tdrd::* with the semantics of
// src/tdse/td_read.cpp (settings echo, leading-block HDF5 reads, restart
// coefficients) on the YAML subset reader and the HDF5 C API.
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <numeric>
#include <sstream>
#include <stdexcept>

#include "h5io.hpp"
#include "tdse_host.hpp"
#include "yaml_lite.hpp"

namespace tdrd {

int readConfig(std::string file, std::string &pot, std::string set_base,
               std::string option, int &L_max, char &gauge,
               std::vector<int> &state_sz, double &timestep,
               std::string &shape, double &w, double &Io, double &cepd,
               int &cycles, Extensions *ext) {
  mpci::YamlDoc settings = mpci::YamlDoc::load_file(file);
  std::cout << "Global Settings:" << std::endl;
  pot = settings.str("Global_Settings", "potential");
  std::cout << "  Core Potential:                       " << pot << std::endl;
  L_max = settings.as_int(set_base, option);
  std::cout << "  max l/L:                              " << L_max << std::endl;
  gauge = settings.as_char("Global_Settings", "gauge");
  std::cout << "  Gauge type ('l' length/'v' velocity): " << gauge << std::endl;

  std::cout << "Propagator Settings:" << std::endl;
  std::vector<int> loc_sz =
      settings.as_int_list("Propagator_Settings", "states_in_l");
  if (static_cast<int>(loc_sz.size()) != L_max + 1)
    throw std::runtime_error("Propagator_Settings.states_in_l must have "
                             "L_max + 1 entries");
  std::cout << "  states per l/L:          ";
  for (auto i = 0; i < static_cast<int>(loc_sz.size()); ++i) {
    std::cout << loc_sz[i] << " ";
    state_sz.push_back(loc_sz[i]);
  }
  std::cout << std::endl;
  timestep = settings.as_double("Propagator_Settings", "dt");
  std::cout << "  timestep dt:                          " << timestep
            << std::endl;

  std::cout << "Field Parameters:" << std::endl;
  shape = settings.str("Field_Parameters", "shape");
  std::cout << "  envelope shape:                       " << shape << std::endl;
  w = settings.as_double("Field_Parameters", "w");
  std::cout << "  photon energy:                        " << w << std::endl;
  Io = settings.as_double("Field_Parameters", "Io");
  std::cout << "  Peak intensity:                       " << Io << std::endl;
  cepd = settings.as_double("Field_Parameters", "cepd");
  std::cout << "  phase shift:                          " << cepd << std::endl;
  cycles = settings.as_int("Field_Parameters", "cycles");
  std::cout << "  number of cycles:                     " << cycles
            << std::endl;

  if (ext) {
    const char *ps = "Propagator_Settings";
    if (settings.has(ps, "mode")) ext->mode = settings.str(ps, "mode");
    if (settings.has(ps, "tol")) {
      ext->atol = ext->rtol = settings.as_double(ps, "tol");
    }
    if (settings.has(ps, "atol")) ext->atol = settings.as_double(ps, "atol");
    if (settings.has(ps, "rtol")) ext->rtol = settings.as_double(ps, "rtol");
    if (settings.has(ps, "full_precision"))
      ext->full_precision = settings.as_bool(ps, "full_precision");
    if (ext->mode != "fixed" && ext->mode != "adaptive")
      throw std::runtime_error("Propagator_Settings.mode must be \"fixed\" "
                               "or \"adaptive\"");
    if (ext->mode == "adaptive")
      std::cout << "  adaptive RKF78, atol/rtol:            " << ext->atol
                << " " << ext->rtol << std::endl;
  }
  return 0;
}

int readEnergies(std::string pot, std::string setname, int L_max, int &ct_sz,
                 std::vector<int> &state_sz, std::vector<int> &offs,
                 stvupt &eig) {
  ct_sz = std::accumulate(state_sz.begin(), state_sz.end(), 0);
  int sum = 0;
  offs.push_back(sum);
  for (auto L = 0; L <= L_max; ++L) {
    const int L_sz = state_sz[L];
    sum += L_sz;
    offs.push_back(sum);
    std::string filename = pot + std::to_string(L) + ".h5";
    *eig[L] = mpci::h5_read_prefix_1d(filename, setname, L_sz);
  }
  return 0;
}

int readDipoles(std::string pot, std::string setname, char gauge, int L_max,
                std::vector<int> &state_sz, stvupt &dipoles) {
  for (auto L = 0; L < L_max; ++L) {
    std::string filename =
        pot + std::to_string(L) + std::to_string(L + 1) + gauge + ".h5";
    *dipoles[L] = mpci::h5_read_block_2d(filename, setname, state_sz[L],
                                         state_sz[L + 1]);
  }
  return 0;
}

int readInitCt(std::string file, int ct_sz, std::vector<double> &ct) {
  std::ifstream fl(file);
  if (!fl) throw std::runtime_error("cannot open start coefficients " + file);
  std::string temp;
  int i = 0;
  while (i < ct_sz * 2) {
    if (!std::getline(fl, temp))
      throw std::runtime_error(file + " holds fewer than ct_sz coefficients");
    size_t a = 0;
    while (a < temp.size() && temp[a] == ' ') ++a;  // regex "^ +"
    temp = temp.substr(a);
    if (!temp.empty() && temp[0] == '#') continue;
    std::istringstream iss(temp);
    int idx;
    double real, imag;
    if (!(iss >> idx >> real >> imag))
      throw std::runtime_error("bad coefficient line in " + file + ": " +
                               temp);
    ct[i] = real;
    ct[i + 1] = imag;
    i += 2;
  }
  return 0;
}

double timeFromCtName(const std::string &path) {
  std::string fln = path;
  size_t slash = fln.find_last_of('/');
  if (slash != std::string::npos) fln = fln.substr(slash + 1);
  size_t t_start = fln.find_first_of("0123456789");
  if (t_start == std::string::npos)
    throw std::runtime_error("no start time in file name " + fln);
  return std::strtod(fln.c_str() + t_start, nullptr);
}

}  // namespace tdrd
