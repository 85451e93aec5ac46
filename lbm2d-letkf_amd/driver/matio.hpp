// Batched matrix files of the reference's LETKF diagnostics
// (src/util/cu_matrix.hpp:97-117, src/util/letkf_solver.h:244-345):
//   <dir>/<label>.<float|double>_<rows>_<cols>_<batch>.<t>.h5, dataset "data",
//   dims {rows, cols, batch}, the buffer being column-major per matrix with the batch
//   outermost (element (r, c, b) at r + rows*(c + cols*b)), written as it lies.
// Built with HDF5 (LBMLETKF_HDF5) the files are HDF5 as the reference's; without
// it the same buffer goes to a raw file of the same name ending in .bin.
#pragma once

#include <stdexcept>
#include <string>
#include <vector>

#ifdef LBMLETKF_HDF5
#include <hdf5.h>
#include <hdf5_hl.h>
#endif

#include "datio.hpp"

namespace lbmletkf_io {

template <class T> const char* type_name();
template <> inline const char* type_name<float>() { return "float"; }
template <> inline const char* type_name<double>() { return "double"; }

inline bool have_hdf5() {
#ifdef LBMLETKF_HDF5
    return true;
#else
    return false;
#endif
}

template <class T>
std::string matrix_file(const std::string& dir, const std::string& label, size_t rows, size_t cols, size_t batch,
                        long long t) {
    return dir + "/" + label + "." + type_name<T>() + "_" + std::to_string(rows) + "_" + std::to_string(cols) + "_" +
           std::to_string(batch) + "." + std::to_string(t) + (have_hdf5() ? ".h5" : ".bin");
}

template <class T>
void save_matrix(const std::string& dir, const std::string& label, size_t rows, size_t cols, size_t batch,
                 long long t, const std::vector<T>& a) {
    if (a.size() != rows * cols * batch) throw std::runtime_error("save_matrix: size mismatch for " + label);
    const std::string fname = matrix_file<T>(dir, label, rows, cols, batch, t);
#ifdef LBMLETKF_HDF5
    const hid_t fid = H5Fcreate(fname.c_str(), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
    if (fid < 0) throw std::runtime_error("could not open file: " + fname);
    const hsize_t dims[3] = {rows, cols, batch};
    const herr_t e = H5LTmake_dataset(fid, "data", 3, dims,
                                      sizeof(T) == 4 ? H5T_NATIVE_FLOAT : H5T_NATIVE_DOUBLE, a.data());
    H5Fclose(fid);
    if (e < 0) throw std::runtime_error("could not write h5 data: " + fname);
#else
    std::FILE* fp = std::fopen(fname.c_str(), "wb");
    if (!fp) throw std::runtime_error("could not open file: " + fname);
    const size_t w = std::fwrite(a.data(), sizeof(T), a.size(), fp);
    std::fclose(fp);
    if (w != a.size()) throw std::runtime_error("short write: " + fname);
#endif
}

}  // namespace lbmletkf_io
