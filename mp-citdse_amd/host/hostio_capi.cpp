// extern "C" access to the drop-in's host I/O (YAML subset, HDF5 leading
// blocks, coefficient files, restart time parse) for callers without C++
// and for the CPU tests.  Returns 0 on success, -1 with a message otherwise.
#include <cstring>
#include <exception>
#include <string>
#include <vector>

#include "h5io.hpp"
#include "tdse_host.hpp"
#include "yaml_lite.hpp"

namespace {
thread_local std::string g_io_err;

template <class F>
int guarded(F &&f) {
  try {
    f();
    return 0;
  } catch (const std::exception &e) {
    g_io_err = e.what();
    return -1;
  }
}
}  // namespace

extern "C" {

const char *mpci_io_last_error(void) { return g_io_err.c_str(); }

int mpci_io_yaml_str(const char *path, const char *section, const char *key,
                     char *out, int cap) {
  return guarded([&] {
    std::string v = mpci::YamlDoc::load_file(path).str(section, key);
    if ((int)v.size() + 1 > cap) throw std::runtime_error("buffer too small");
    std::memcpy(out, v.c_str(), v.size() + 1);
  });
}

int mpci_io_yaml_int_list(const char *path, const char *section,
                          const char *key, int *out, int cap, int *n) {
  return guarded([&] {
    std::vector<int> v =
        mpci::YamlDoc::load_file(path).as_int_list(section, key);
    if ((int)v.size() > cap) throw std::runtime_error("buffer too small");
    for (size_t i = 0; i < v.size(); ++i) out[i] = v[i];
    *n = (int)v.size();
  });
}

int mpci_io_yaml_double(const char *path, const char *section,
                        const char *key, double *out) {
  return guarded(
      [&] { *out = mpci::YamlDoc::load_file(path).as_double(section, key); });
}

int mpci_io_read_1d(const char *file, const char *dataset, long n,
                    double *out) {
  return guarded([&] {
    auto v = mpci::h5_read_prefix_1d(file, dataset, n);
    std::memcpy(out, v.data(), sizeof(double) * v.size());
  });
}

int mpci_io_read_2d(const char *file, const char *dataset, long rows,
                    long cols, double *out) {
  return guarded([&] {
    auto v = mpci::h5_read_block_2d(file, dataset, rows, cols);
    std::memcpy(out, v.data(), sizeof(double) * v.size());
  });
}

int mpci_io_write_1d(const char *file, const char *dataset,
                     const double *data, long n) {
  return guarded([&] { mpci::h5_write_1d(file, dataset, data, n); });
}

int mpci_io_write_2d(const char *file, const char *dataset,
                     const double *data, long rows, long cols) {
  return guarded([&] { mpci::h5_write_2d(file, dataset, data, rows, cols); });
}

int mpci_io_write_ct(const char *prefix, double t, const double *ct,
                     int ct_sz, int precision) {
  return guarded([&] {
    std::vector<double> v(ct, ct + 2 * (size_t)ct_sz);
    tdse::writeCt(prefix, t, v, ct_sz, precision);
  });
}

int mpci_io_read_ct(const char *path, int ct_sz, double *out) {
  return guarded([&] {
    std::vector<double> v(2 * (size_t)ct_sz);
    tdrd::readInitCt(path, ct_sz, v);
    std::memcpy(out, v.data(), sizeof(double) * v.size());
  });
}

int mpci_io_time_from_name(const char *path, double *t) {
  return guarded([&] { *t = tdrd::timeFromCtName(path); });
}

}  // extern "C"
