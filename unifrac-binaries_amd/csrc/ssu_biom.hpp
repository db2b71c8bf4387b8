// BIOM 2.x (HDF5) table reader: /observation/matrix/{indptr,indices,data},
// /observation/ids and /sample/ids (the datasets src/biom.hpp:64-72 names).
// libhdf5 is opened at run time, so the in-memory paths never depend on it.
#pragma once

#include <string>

#include "ssu_table.hpp"

namespace ssu {

bool load_biom(const char* filename, Table& out, std::string* err);

}  // namespace ssu
