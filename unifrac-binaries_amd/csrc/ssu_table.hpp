// Observation-major CSR feature table (the reference's support_biom_t layout,
// src/api.hpp:145-154) with the per-sample totals of src/biom_inmem.cpp:513-526.
#pragma once

#include <cstdint>
#include <string>
#include <unordered_map>
#include <vector>

namespace ssu {

struct Table {
    uint32_t n_obs = 0, n_samples = 0;
    std::vector<std::string> obs_ids, sample_ids;
    // CSR by observation: row i holds (sample index, value) pairs
    const uint32_t* indptr = nullptr;
    const uint32_t* indices = nullptr;
    const double* data = nullptr;
    std::vector<double> totals;                            // per-sample sum of values
    // Found by the pass that forms the totals: every sample index < n_samples
    // (entries outside are left out of the totals), every value finite and >= 0,
    // every value an integer
    bool index_ok = true, values_nonneg = true, values_integral = true;
    std::unordered_map<std::string, uint32_t> obs_index;   // id -> row

    // Borrow caller arrays (src/biom_inmem.cpp:342-384 borrows them too).
    // index_obs = false skips obs_ids and obs_index (callers that match rows to tips themselves).
    static Table borrow(const char* const* obs_ids, const char* const* sample_ids, const uint32_t* indptr,
                        const uint32_t* indices, const double* data, uint32_t n_obs, uint32_t n_samples,
                        bool index_obs = true);
    // Own the arrays (file readers, filtered copies).
    static Table own(std::vector<std::string> obs_ids, std::vector<std::string> sample_ids,
                     std::vector<uint32_t> indptr, std::vector<uint32_t> indices, std::vector<double> data);
    // Two dense samples -> table, keeping values > 0 (src/biom_inmem.cpp:183-213).
    static Table dense_pair(const char* const* obs_ids, const double* s1, const double* s2, uint32_t n_obs);

    // Drop samples whose total is below min_total, then observations left
    // without entries (src/biom_inmem.cpp:83-128, 284-338).
    Table filtered(double min_total) const;

    Table() = default;
    Table(const Table&) = delete;             // the CSR pointers may point into own storage
    Table& operator=(const Table&) = delete;
    Table(Table&&) = default;
    Table& operator=(Table&&) = default;

    uint64_t nnz() const { return n_obs ? indptr[n_obs] - indptr[0] : 0; }

private:
    std::vector<uint32_t> own_indptr_, own_indices_;
    std::vector<double> own_data_;
    void finish(bool index_obs = true);
};

}  // namespace ssu
