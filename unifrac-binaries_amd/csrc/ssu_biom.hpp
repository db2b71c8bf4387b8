// BIOM 2.x (HDF5) table reader: /observation/matrix/{indptr,indices,data},
// /observation/ids and /sample/ids (the datasets src/biom.hpp:64-72 names).
// libhdf5 is opened at run time, so the in-memory paths never depend on it.
#pragma once

#include <cstdint>
#include <memory>
#include <string>

#include "ssu_table.hpp"

namespace ssu {

bool load_biom(const char* filename, Table& out, std::string* err);

// number of entries of /sample/ids (src/biom.hpp load_n_samples)
bool biom_n_samples(const char* filename, uint64_t& n, std::string* err);

// Distance-matrix file, format "BDSM" version "2020.12" (src/api.cpp:1787-1830
// without PCoA / statistics): "format", "version", "order" (sample ids) and,
// when save_dist, the n x n "matrix" in IEEE little-endian FP64 or FP32.
bool write_matrix_hdf5(const char* filename, uint32_t n, const char* const* ids, const void* matrix, bool fp64,
                       bool save_dist, std::string* err);

// An HDF5 output file written dataset by dataset, with the dataset encodings of
// src/api.cpp:1167-1262: one fixed-length Fortran string, arrays of
// variable-length C strings, 1-D and 2-D IEEE little-endian reals, 1-D
// little-endian uint32.  The file closes when the writer is destroyed.
class H5Writer {
public:
    static std::unique_ptr<H5Writer> create(const char* filename, std::string* err);
    ~H5Writer();
    bool string(const char* name, const std::string& v);
    bool strings(const char* name, uint64_t n, const char* const* v);
    bool reals(const char* name, uint64_t n, const void* v, bool fp64);
    bool matrix(const char* name, uint64_t rows, uint64_t cols, const void* v, bool fp64);
    bool u32s(const char* name, uint64_t n, const uint32_t* v);

private:
    H5Writer() = default;
    int64_t file_ = -1;
};

}  // namespace ssu
