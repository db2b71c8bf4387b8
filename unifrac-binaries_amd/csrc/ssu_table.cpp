#include "ssu_table.hpp"

#include <algorithm>
#include <cmath>
#include <thread>

namespace ssu {

namespace {
// What one read of the entries establishes besides the totals
struct EntryScan {
    bool index_ok = true;   // every sample index < n_samples
    bool nonneg = true;     // every value finite and >= 0
    bool integral = true;   // every value an integer
};

// Sample totals over observation ranges on several threads, when that gives the
// sequential sum bit for bit: every value an integer and the sum of all their
// magnitudes below 2^53, so every partial sum, in any order, is an integer
// below 2^53 and each FP64 add is exact (count tables).  The same read checks
// the sample indices and the values.  false: the totals are not so (the checks
// are filled either way).
bool integer_totals(const uint32_t* indptr, const uint32_t* indices, const double* data, uint32_t n_obs,
                    uint32_t n_samples, double* tot, EntryScan& scan) {
    const uint64_t nnz = n_obs ? (uint64_t)indptr[n_obs] - indptr[0] : 0;
    const unsigned threads = std::max(1u, std::min(8u, std::thread::hardware_concurrency()));
    if (nnz < (1u << 20) || threads < 2) {
        // in observation order on this thread (the reference's order)
        std::fill(tot, tot + n_samples, 0.0);
        for (uint32_t i = 0; i < n_obs; ++i)
            for (uint32_t j = indptr[i]; j < indptr[i + 1]; ++j) {
                const uint32_t s = indices[j];
                const double v = data[j];
                scan.nonneg &= v >= 0.0 && v <= 1.7976931348623157e308;
                scan.integral &= v == std::nearbyint(v);
                if (s < n_samples) tot[s] += v;
                else scan.index_ok = false;
            }
        return true;
    }
    // observation ranges of about equal entry counts
    std::vector<uint32_t> cut(threads + 1, n_obs);
    cut[0] = 0;
    for (unsigned t = 1, i = 0; t < threads; ++t) {
        const uint64_t goal = indptr[0] + nnz * t / threads;
        while (i < n_obs && indptr[i] < goal) ++i;
        cut[t] = i;
    }
    std::vector<std::vector<double>> part(threads);
    std::vector<uint8_t> integral(threads, 1), index_ok(threads, 1), nonneg(threads, 1), whole_vals(threads, 1);
    std::vector<double> mag(threads, 0.0);
    std::vector<std::thread> pool;
    auto run = [&](unsigned t) {
        std::vector<double>& p = part[t];
        p.assign((size_t)n_samples + 1, 0.0);   // slot n_samples takes out-of-range indices
        bool ok = true, ix = true, nn = true, wv = true;
        double a = 0.0;   // sum of magnitudes (exact while below 2^53)
        for (uint32_t i = cut[t]; i < cut[t + 1]; ++i)
            for (uint32_t j = indptr[i]; j < indptr[i + 1]; ++j) {
                const uint32_t s = indices[j];
                const double v = data[j];
                const bool w = v == std::nearbyint(v);
                wv &= w;
                ok &= w && std::fabs(v) < 0x1p52;
                nn &= v >= 0.0 && v <= 1.7976931348623157e308;
                ix &= s < n_samples;
                a += std::fabs(v);
                p[s < n_samples ? s : n_samples] += v;
            }
        integral[t] = ok;
        index_ok[t] = ix;
        nonneg[t] = nn;
        whole_vals[t] = wv;
        mag[t] = a;
    };
    for (unsigned t = 1; t < threads; ++t) pool.emplace_back(run, t);
    run(0);
    for (auto& th : pool) th.join();
    double a = 0.0;
    bool whole = true;
    for (unsigned t = 0; t < threads; ++t) {
        scan.index_ok &= index_ok[t] != 0;
        scan.nonneg &= nonneg[t] != 0;
        scan.integral &= whole_vals[t] != 0;
        whole = whole && integral[t] && mag[t] < 0x1p52;
        a += mag[t];
    }
    if (!whole || !(a < 0x1p52)) return false;
    for (uint32_t s = 0; s < n_samples; ++s) {
        double x = 0.0;
        for (unsigned t = 0; t < threads; ++t) x += part[t][s];
        tot[s] = x;
    }
    return true;
}
}  // namespace

void Table::finish(bool index_obs) {
    // src/biom_inmem.cpp:513-526: totals accumulate in observation order (for
    // integer counts any order gives the same sums)
    totals.assign(n_samples, 0.0);
    double* tot = totals.data();
    EntryScan scan;
    if (!integer_totals(indptr, indices, data, n_obs, n_samples, tot, scan)) {
        std::fill(totals.begin(), totals.end(), 0.0);
        for (uint32_t i = 0; i < n_obs; ++i)
            for (uint32_t j = indptr[i]; j < indptr[i + 1]; ++j)
                if (indices[j] < n_samples) tot[indices[j]] += data[j];
    }
    index_ok = scan.index_ok;
    values_nonneg = scan.nonneg;
    values_integral = scan.integral;
    obs_index.clear();
    if (!index_obs) return;
    obs_index.reserve(n_obs * 2 + 1);
    for (uint32_t i = 0; i < n_obs; ++i) obs_index[obs_ids[i]] = i;
}

Table Table::borrow(const char* const* oids, const char* const* sids, const uint32_t* ip, const uint32_t* ix,
                    const double* d, uint32_t n_obs, uint32_t n_samples, bool index_obs) {
    Table t;
    t.n_obs = n_obs;
    t.n_samples = n_samples;
    if (index_obs) {
        t.obs_ids.reserve(n_obs);
        for (uint32_t i = 0; i < n_obs; ++i) t.obs_ids.emplace_back(oids[i]);
    }
    t.sample_ids.reserve(n_samples);
    for (uint32_t i = 0; i < n_samples; ++i) t.sample_ids.emplace_back(sids[i]);
    t.indptr = ip;
    t.indices = ix;
    t.data = d;
    t.finish(index_obs);
    return t;
}

Table Table::own(std::vector<std::string> oids, std::vector<std::string> sids, std::vector<uint32_t> ip,
                 std::vector<uint32_t> ix, std::vector<double> d) {
    Table t;
    t.n_obs = (uint32_t)oids.size();
    t.n_samples = (uint32_t)sids.size();
    t.obs_ids = std::move(oids);
    t.sample_ids = std::move(sids);
    t.own_indptr_ = std::move(ip);
    t.own_indices_ = std::move(ix);
    t.own_data_ = std::move(d);
    if (t.own_indptr_.empty()) t.own_indptr_.push_back(0);
    t.indptr = t.own_indptr_.data();
    t.indices = t.own_indices_.data();
    t.data = t.own_data_.data();
    t.finish();
    return t;
}

Table Table::dense_pair(const char* const* oids, const double* s1, const double* s2, uint32_t n_obs) {
    std::vector<std::string> o, s{"Sample1", "Sample2"};
    std::vector<uint32_t> ip{0}, ix;
    std::vector<double> d;
    const double* cols[2] = {s1, s2};
    for (uint32_t i = 0; i < n_obs; ++i) {
        o.emplace_back(oids[i]);
        for (uint32_t j = 0; j < 2; ++j)
            if (cols[j][i] > 0.0) { ix.push_back(j); d.push_back(cols[j][i]); }
        ip.push_back((uint32_t)ix.size());
    }
    return own(std::move(o), std::move(s), std::move(ip), std::move(ix), std::move(d));
}

Table Table::filtered(double min_total) const {
    std::vector<int64_t> new_index(n_samples, -1);
    std::vector<std::string> sids;
    for (uint32_t s = 0; s < n_samples; ++s)
        if (totals[s] >= min_total) { new_index[s] = (int64_t)sids.size(); sids.push_back(sample_ids[s]); }
    std::vector<std::string> oids;
    std::vector<uint32_t> ip{0}, ix;
    std::vector<double> d;
    if (!sids.empty()) {
        for (uint32_t i = 0; i < n_obs; ++i) {
            const size_t before = ix.size();
            for (uint32_t j = indptr[i]; j < indptr[i + 1]; ++j) {
                const int64_t ni = indices[j] < n_samples ? new_index[indices[j]] : -1;
                if (ni >= 0) { ix.push_back((uint32_t)ni); d.push_back(data[j]); }
            }
            if (ix.size() > before) { oids.push_back(obs_ids[i]); ip.push_back((uint32_t)ix.size()); }
        }
    }
    if (oids.empty()) sids.clear();  // degenerate: nothing left (src/biom_inmem.cpp:332-337)
    return own(std::move(oids), std::move(sids), std::move(ip), std::move(ix), std::move(d));
}

}  // namespace ssu
