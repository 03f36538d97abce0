// HDF5 access for the tdse inputs, on the HDF5 C API.
//
// Reads mirror tdrd::readEnergies / tdrd::readDipoles
// (src/tdse/td_read.cpp:50-120): the leading `n` entries of a 1-D dataset
// (/En or /En_CI) and the top-left rows x cols block of a 2-D row-major
// dataset (/d_if or /CI_dip) through a hyperslab selection.  Writes produce
// the layouts of the structure programs (h1e.cpp:22-58, dmx1e.cpp:94-130,
// w2e.cpp:67-120): 1-D and 2-D FP64 datasets.
#pragma once

#include <stdexcept>
#include <string>
#include <vector>

namespace mpci {

class H5Error : public std::runtime_error {
 public:
  using std::runtime_error::runtime_error;
};

// Leading n doubles of a 1-D dataset.
std::vector<double> h5_read_prefix_1d(const std::string &file,
                                      const std::string &dataset, long n);
// Top-left rows x cols block of a 2-D dataset, row-major.
std::vector<double> h5_read_block_2d(const std::string &file,
                                     const std::string &dataset, long rows,
                                     long cols);
// Extent of a dataset (1 or 2 dims).
std::vector<long> h5_dims(const std::string &file, const std::string &dataset);

// Read a whole 1-D or 2-D dataset (row-major), with its extent.
std::vector<double> h5_read_all(const std::string &file,
                                const std::string &dataset,
                                std::vector<long> *dims = nullptr);

void h5_write_1d(const std::string &file, const std::string &dataset,
                 const double *data, long n);
// Several datasets into one new file (w2e.cpp:67-71 writes En_CI and
// CI_vecs side by side).
struct H5Out {
  std::string name;
  const double *data;
  std::vector<long> dims;
};
void h5_write_all(const std::string &file, const std::vector<H5Out> &sets);
void h5_write_2d(const std::string &file, const std::string &dataset,
                 const double *data, long rows, long cols);

}  // namespace mpci
