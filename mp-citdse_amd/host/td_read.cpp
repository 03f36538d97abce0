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
  const mpci::YamlDoc doc = mpci::YamlDoc::load_file(file);
  const char *glob = "Global_Settings", *prop = "Propagator_Settings",
             *field = "Field_Parameters";
  // The stdout echo is part of the drop-in contract: each value is printed
  // as soon as it is read, under its section heading, with the label
  // padded to the value column of td_read.cpp:10-45.
  auto section = [](const char *title) { std::cout << title << std::endl; };
  auto show = [](const char *label, const auto &value) {
    std::cout << "  " << label << value << std::endl;
  };

  section("Global Settings:");
  show("Core Potential:                       ",
       pot = doc.str(glob, "potential"));
  show("max l/L:                              ",
       L_max = doc.as_int(set_base, option));
  show("Gauge type ('l' length/'v' velocity): ",
       gauge = doc.as_char(glob, "gauge"));

  section("Propagator Settings:");
  const std::vector<int> sizes = doc.as_int_list(prop, "states_in_l");
  if (static_cast<int>(sizes.size()) != L_max + 1)
    throw std::runtime_error("Propagator_Settings.states_in_l must have "
                             "L_max + 1 entries");
  state_sz.insert(state_sz.end(), sizes.begin(), sizes.end());
  std::string size_list;
  for (const int n : sizes) size_list += std::to_string(n) + ' ';
  show("states per l/L:          ", size_list);
  show("timestep dt:                          ",
       timestep = doc.as_double(prop, "dt"));

  section("Field Parameters:");
  show("envelope shape:                       ",
       shape = doc.str(field, "shape"));
  show("photon energy:                        ", w = doc.as_double(field, "w"));
  show("Peak intensity:                       ",
       Io = doc.as_double(field, "Io"));
  show("phase shift:                          ",
       cepd = doc.as_double(field, "cepd"));
  show("number of cycles:                     ",
       cycles = doc.as_int(field, "cycles"));

  if (ext) {
    // optional extension keys, absent from the reference's inputs
    if (doc.has(prop, "mode")) ext->mode = doc.str(prop, "mode");
    if (doc.has(prop, "tol")) ext->atol = ext->rtol = doc.as_double(prop, "tol");
    if (doc.has(prop, "atol")) ext->atol = doc.as_double(prop, "atol");
    if (doc.has(prop, "rtol")) ext->rtol = doc.as_double(prop, "rtol");
    if (doc.has(prop, "full_precision"))
      ext->full_precision = doc.as_bool(prop, "full_precision");
    if (doc.has(prop, "observables"))
      ext->observables = doc.as_bool(prop, "observables");
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

// The time is the number that starts at the first digit of the file name
// (tdse_main.cpp:153-157), which for the reference's <pot>_ct_<t>.dat files
// is <t>.  The full-precision side files <pot>_ct17_<t>.dat carry a digit in
// their tag, so for them the number is read after the tag instead.
double timeFromCtName(const std::string &path) {
  static const std::string kFullTag = "_ct17_";
  const size_t slash = path.find_last_of('/');
  const std::string name =
      slash == std::string::npos ? path : path.substr(slash + 1);
  const size_t tag = name.rfind(kFullTag);
  const size_t from = tag != std::string::npos
                          ? tag + kFullTag.size()
                          : name.find_first_of("0123456789");
  if (from == std::string::npos || from >= name.size())
    throw std::runtime_error("no start time in file name " + name);
  const char *begin = name.c_str() + from;
  char *end = nullptr;
  const double t = std::strtod(begin, &end);
  if (end == begin)
    throw std::runtime_error("no start time in file name " + name);
  return t;
}

}  // namespace tdrd
