// BIOM 2.x (HDF5) table reader: /observation/matrix/{indptr,indices,data},
// /observation/ids and /sample/ids (the datasets src/biom.hpp:64-72 names).
// libhdf5 is opened at run time, so the in-memory paths never depend on it.
#pragma once

#include <cstdint>
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

}  // namespace ssu
