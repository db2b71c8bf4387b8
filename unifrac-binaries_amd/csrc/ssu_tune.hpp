// Occupancy-split thresholds chosen per workload.
//
// A node held by c of the n samples costs the tiled kernels a fixed amount of
// pair work (every block whose sample panels hold it), and the posting-list
// kernels c (c - 1) / 2 posting visits plus a per-posting build cost.  The
// split sends nodes held by fewer than `thr` samples to the posting lists;
// choose_split evaluates the modelled time of every candidate threshold over
// the plan's actual occupancies and returns the cheapest.  The rates are
// MI355X measurements of these kernels; the model is deterministic, so one
// workload always gets the same split (and so the same bits).
#pragma once

#include <cstdint>
#include <vector>

namespace ssu {

enum class SplitKind : int {
    weighted32 = 0,     // FP32 weighted: 32 x 64 blocks over the nodes either panel holds
    weighted64 = 1,     // FP64 weighted: 64 x 64 blocks over the nodes both panels hold
    generalized32 = 2,  // FP32 generalized: 32 x 64 blocks, both panels' nodes, two sums
    gram = 3,           // unweighted: int8 Gram columns (every pair, no masks)
    generalized64 = 4,  // FP64 generalized: as generalized32 with FP64 operands and terms
    vaw32 = 5,          // FP32 VAW: dense pair tiles over every pair, one-sided sums per (posting, total)
    vaw64 = 6,          // FP64 VAW: the same with FP64 operands and terms
};

struct SplitChoice {
    uint32_t thr = 0;          // nodes held by fewer samples go to the posting lists (0: no split)
    double model_ms = 0.0;     // modelled time of the chosen split
    double model_ms_none = 0;  // modelled time without a split
};

// occupancy: structural occupancy of each processed node (0: held by no
// sample).  max_slots bounds the postings (sum of min(occupancy, thr - 1)).
// cols (VAW): the number of totals each posting's one-sided term is summed at.
SplitChoice choose_split(SplitKind kind, uint32_t n, const std::vector<uint32_t>& occupancy, uint64_t max_slots,
                         uint32_t cols = 0);

}  // namespace ssu
