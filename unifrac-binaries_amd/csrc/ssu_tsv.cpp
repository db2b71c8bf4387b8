#include "ssu_tsv.hpp"

#include <fstream>
#include <map>
#include <sstream>
#include <unordered_map>

namespace ssu {

namespace {

std::vector<std::string> split(const std::string& line, char sep) {
    std::vector<std::string> out;
    size_t b = 0;
    while (true) {
        const size_t e = line.find(sep, b);
        if (e == std::string::npos) {
            out.push_back(line.substr(b));
            return out;
        }
        out.push_back(line.substr(b, e - b));
        b = e + 1;
    }
}

}  // namespace

bool read_grouping(const std::string& filename, const std::vector<std::string>& ids, const std::string& column,
                   std::vector<uint32_t>& grouping, uint32_t& n_groups, std::string* err) {
    std::ifstream in(filename);
    if (!in) { if (err) *err = "cannot open grouping file " + filename; return false; }
    std::string line;
    if (!std::getline(in, line)) { if (err) *err = "empty grouping file " + filename; return false; }
    const std::vector<std::string> header = split(line, '\t');
    size_t col = 0;
    for (size_t i = 1; i < header.size(); ++i)
        if (header[i] == column) { col = i; break; }
    if (col == 0) { if (err) *err = "column " + column + " not found in " + filename; return false; }
    std::unordered_map<std::string, uint32_t> index;
    for (size_t i = 0; i < ids.size(); ++i) index[ids[i]] = (uint32_t)i;
    const uint32_t unset = 0xffffffffu;
    grouping.assign(ids.size(), unset);
    std::map<std::string, uint32_t> values;
    uint32_t next = 0;
    while (std::getline(in, line)) {
        const std::vector<std::string> row = split(line, '\t');
        const auto it = index.find(row[0]);
        if (it == index.end()) continue;
        if (col >= row.size()) { if (err) *err = "row of " + row[0] + " lacks column " + column; return false; }
        const auto v = values.find(row[col]);
        uint32_t g;
        if (v != values.end()) g = v->second;
        else values[row[col]] = g = next++;
        grouping[it->second] = g;
    }
    for (size_t i = 0; i < ids.size(); ++i)
        if (grouping[i] == unset) { if (err) *err = "sample " + ids[i] + " missing from " + filename; return false; }
    n_groups = next;
    return true;
}

std::vector<std::string> split_commas(const char* list) {
    if (!list) return {};
    return split(list, ',');
}

}  // namespace ssu
