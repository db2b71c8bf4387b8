// Grouping column of a tab-separated metadata file, for PERMANOVA.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace ssu {

// src/tsv.cpp:120-178 (indexed_tsv::read_grouping): the first line is the
// header, the first field of each row the sample id.  Each distinct value of
// `column` becomes a group number in order of first appearance among the rows
// whose id is in `ids`; rows with other ids are skipped.  Fails when the
// column is absent or a sample of `ids` has no row (the reference leaves the
// first sample unchecked; every sample is checked here).
bool read_grouping(const std::string& filename, const std::vector<std::string>& ids, const std::string& column,
                   std::vector<uint32_t>& grouping, uint32_t& n_groups, std::string* err);

// src/api.cpp:924-945: "a,b,c" -> {"a", "b", "c"}
std::vector<std::string> split_commas(const char* list);

}  // namespace ssu
