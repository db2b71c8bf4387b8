#include "ssu_tune.hpp"

#include <algorithm>
#include <cmath>

namespace ssu {

namespace {

// Marginal seconds per unit on one MI355X, fitted to whole-plan timings at
// three thresholds each on the 16,384-sample (100k tips) and 8,192-sample
// (50k tips) workloads (total = fixed + terms * term + postings * posting); the
// modelled optimum lands within the measured one for every kind (weighted
// 0.244 vs 0.25, FP64 weighted 0.229 vs 0.25, generalized 0.738 vs 0.8,
// Gram 0.075 vs 0.06 - 0.1).  The per-term cost includes the list building
// and row lists, whose work follows the postings.
struct Rates {
    double term;      // one (pair, node) term of the tiled kernel; Gram: one (pair, column) unit
    double posting;   // one posting visited by the row walk
    double onesided = 0.0;   // VAW: one (posting slot, total) one-sided term
};

Rates rates_of(SplitKind k) {
    switch (k) {
    case SplitKind::weighted32: return {4.2e-14, 7.0e-13};
    case SplitKind::weighted64: return {7.1e-14, 1.36e-12};
    case SplitKind::generalized32: return {6.0e-13, 1.1e-12};
    case SplitKind::gram: return {3.4e-15, 6.1e-13};
    // FP64 generalized (an FP64 power and division per tile term, two powers per
    // posting): fitted to the 8,192-sample workload at theta 0.5 / 0.707 / 0.9
    // (within 2% of each phase's measured time)
    case SplitKind::generalized64: return {7.7e-12, 7.5e-12};
    // VAW: the dense kernel's term over the dense nodes' rows; a posting (three
    // FP64 reciprocal square roots); a one-sided term at one total.  Fitted to the
    // 8,192-sample workload at theta 0.25 / 0.5 / 1 (weighted and presence terms
    // within 15% of each other; the weighted rates are used for both)
    case SplitKind::vaw32: return {4.4e-13, 1.72e-12, 4.13e-13};
    case SplitKind::vaw64: return {2.26e-12, 2.84e-12, 6.2e-13};
    }
    return {1.0, 1.0};
}

// expected tiled-kernel terms of a node held by c of n samples, placed at
// random: the pair blocks visit a node either of their sample panels holds
// (the kernels' counts agree within 1% on the workloads above); a Gram
// column costs every pair
double tile_terms(SplitKind k, uint32_t n, uint32_t c) {
    const double pairs = 0.5 * (double)n * n;
    if (k == SplitKind::gram || k == SplitKind::vaw32 || k == SplitKind::vaw64) return pairs;
    const double q = 1.0 - (double)c / n;   // a given sample lacks the node
    const double q2 = q * q, q4 = q2 * q2, q8 = q4 * q4, q16 = q8 * q8, q32 = q16 * q16;
    return pairs * (1.0 - q32 * q32 * q32);   // 1 - q^96
}

}  // namespace

SplitChoice choose_split(SplitKind kind, uint32_t n, const std::vector<uint32_t>& occupancy, uint64_t max_slots,
                         uint32_t cols) {
    SplitChoice best;
    if (n < 2 || occupancy.empty()) return best;
    const Rates r = rates_of(kind);
    // nodes by occupancy (0: held by no sample, no cost either way); an
    // occupancy above n (never, for counts of n samples) counts as n
    std::vector<uint64_t> hist((size_t)n + 1, 0);
    for (uint32_t x : occupancy) ++hist[std::min(x, n)];
    // cost with every node on the tiles; then move nodes to the posting lists in
    // increasing occupancy (thr = c + 1 moves every node held by at most c samples)
    double tiles = 0.0;
    for (uint32_t c = 1; c <= n; ++c)
        if (hist[c]) tiles += (double)hist[c] * tile_terms(kind, n, c) * r.term;
    best.model_ms_none = best.model_ms = tiles * 1e3;
    double cost = tiles, posts = 0.0;
    uint64_t slots = 0;
    for (uint32_t c = 0; c <= n; ++c) {
        if (!hist[c]) continue;
        if (c > 0) {
            const double m = (double)hist[c], x = c;
            cost -= m * tile_terms(kind, n, c) * r.term;
            posts += m * (0.5 * x * (x - 1.0) * r.posting + x * (double)cols * r.onesided);
            slots += hist[c] * c;
        }
        if (slots >= max_slots) break;
        if (cost + posts < best.model_ms * 1e-3) {
            best.model_ms = (cost + posts) * 1e3;
            best.thr = c + 1;
        }
    }
    return best;
}

}  // namespace ssu
