#include "h5io.hpp"

#include <hdf5.h>

namespace mpci {

namespace {

// RAII wrapper around an hid_t with its close function.
struct H5Id {
  hid_t id;
  herr_t (*close)(hid_t);
  H5Id(hid_t i, herr_t (*c)(hid_t)) : id(i), close(c) {}
  ~H5Id() {
    if (id >= 0) close(id);
  }
  H5Id(const H5Id &) = delete;
  H5Id &operator=(const H5Id &) = delete;
};

void quiet() {
  static bool done = false;
  if (!done) {
    H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
    done = true;
  }
}

hid_t must(hid_t id, const std::string &what) {
  if (id < 0) throw H5Error(what);
  return id;
}

void must_ok(herr_t e, const std::string &what) {
  if (e < 0) throw H5Error(what);
}

std::vector<double> read_block(const std::string &file,
                               const std::string &dataset, int rank,
                               const hsize_t *count) {
  quiet();
  H5Id f(must(H5Fopen(file.c_str(), H5F_ACC_RDONLY, H5P_DEFAULT),
              "cannot open HDF5 file " + file),
         H5Fclose);
  H5Id d(must(H5Dopen2(f.id, dataset.c_str(), H5P_DEFAULT),
              "no dataset " + dataset + " in " + file),
         H5Dclose);
  H5Id space(must(H5Dget_space(d.id), "dataspace of " + dataset), H5Sclose);
  const int frank = H5Sget_simple_extent_ndims(space.id);
  if (frank != rank)
    throw H5Error(dataset + " in " + file + " has rank " +
                  std::to_string(frank) + ", expected " +
                  std::to_string(rank));
  hsize_t dims[2] = {0, 0};
  H5Sget_simple_extent_dims(space.id, dims, nullptr);
  for (int r = 0; r < rank; ++r)
    if (count[r] > dims[r])
      throw H5Error(dataset + " in " + file + " is smaller than the " +
                    "requested states_in_l block");
  hsize_t offset[2] = {0, 0}, stride[2] = {1, 1}, block[2] = {1, 1};
  must_ok(H5Sselect_hyperslab(space.id, H5S_SELECT_SET, offset, stride, count,
                              block),
          "hyperslab on " + dataset);
  H5Id mem(must(H5Screate_simple(rank, count, nullptr), "memspace"),
           H5Sclose);
  size_t n = 1;
  for (int r = 0; r < rank; ++r) n *= count[r];
  std::vector<double> out(n);
  must_ok(H5Dread(d.id, H5T_NATIVE_DOUBLE, mem.id, space.id, H5P_DEFAULT,
                  out.data()),
          "read of " + dataset + " in " + file);
  return out;
}

void write_into(hid_t f, const std::string &dataset, const double *data,
                int rank, const hsize_t *dims) {
  H5Id space(must(H5Screate_simple(rank, dims, nullptr), "dataspace"),
             H5Sclose);
  H5Id d(must(H5Dcreate2(f, dataset.c_str(), H5T_IEEE_F64LE, space.id,
                         H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT),
              "cannot create dataset " + dataset),
         H5Dclose);
  must_ok(H5Dwrite(d.id, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT,
                   data),
          "write of " + dataset);
}

void write_block(const std::string &file, const std::string &dataset,
                 const double *data, int rank, const hsize_t *dims) {
  quiet();
  H5Id f(must(H5Fcreate(file.c_str(), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT),
              "cannot create HDF5 file " + file),
         H5Fclose);
  write_into(f.id, dataset, data, rank, dims);
}

}  // namespace

std::vector<double> h5_read_prefix_1d(const std::string &file,
                                      const std::string &dataset, long n) {
  hsize_t c[1] = {(hsize_t)n};
  return read_block(file, dataset, 1, c);
}

std::vector<double> h5_read_block_2d(const std::string &file,
                                     const std::string &dataset, long rows,
                                     long cols) {
  hsize_t c[2] = {(hsize_t)rows, (hsize_t)cols};
  return read_block(file, dataset, 2, c);
}

std::vector<long> h5_dims(const std::string &file,
                          const std::string &dataset) {
  quiet();
  H5Id f(must(H5Fopen(file.c_str(), H5F_ACC_RDONLY, H5P_DEFAULT),
              "cannot open HDF5 file " + file),
         H5Fclose);
  H5Id d(must(H5Dopen2(f.id, dataset.c_str(), H5P_DEFAULT),
              "no dataset " + dataset + " in " + file),
         H5Dclose);
  H5Id space(must(H5Dget_space(d.id), "dataspace"), H5Sclose);
  int rank = H5Sget_simple_extent_ndims(space.id);
  hsize_t dims[8];
  H5Sget_simple_extent_dims(space.id, dims, nullptr);
  std::vector<long> out;
  for (int r = 0; r < rank; ++r) out.push_back((long)dims[r]);
  return out;
}

void h5_write_1d(const std::string &file, const std::string &dataset,
                 const double *data, long n) {
  hsize_t d[1] = {(hsize_t)n};
  write_block(file, dataset, data, 1, d);
}

void h5_write_2d(const std::string &file, const std::string &dataset,
                 const double *data, long rows, long cols) {
  hsize_t d[2] = {(hsize_t)rows, (hsize_t)cols};
  write_block(file, dataset, data, 2, d);
}

std::vector<double> h5_read_all(const std::string &file,
                                const std::string &dataset,
                                std::vector<long> *dims) {
  const std::vector<long> d = h5_dims(file, dataset);
  if (dims) *dims = d;
  if (d.size() == 1) return h5_read_prefix_1d(file, dataset, d[0]);
  if (d.size() == 2) return h5_read_block_2d(file, dataset, d[0], d[1]);
  throw H5Error("dataset " + dataset + " in " + file + " is not 1-D or 2-D");
}

void h5_write_all(const std::string &file, const std::vector<H5Out> &sets) {
  quiet();
  H5Id f(must(H5Fcreate(file.c_str(), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT),
              "cannot create HDF5 file " + file),
         H5Fclose);
  for (const H5Out &o : sets) {
    std::vector<hsize_t> d(o.dims.begin(), o.dims.end());
    write_into(f.id, o.name, o.data, (int)d.size(), d.data());
  }
}

}  // namespace mpci
