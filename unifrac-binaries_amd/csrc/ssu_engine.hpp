// Device pipeline for one (table, tree, method) problem.  No HIP types here:
// the C API layer includes this header from plain C++.
#pragma once

#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "ssu_table.hpp"
#include "ssu_tree.hpp"

namespace ssu {

enum class Method : int {
    unweighted = 0,
    unweighted_unnormalized = 1,
    weighted_normalized = 2,
    weighted_unnormalized = 3,
    generalized = 4,
};

struct ComputeSpec {
    Method method = Method::unweighted;
    bool fp64 = false;             // "*_fp64" method names
    bool vaw = false;              // variance adjusted
    double alpha = 1.0;            // generalized
    bool bypass_tips = false;
    bool normalize = true;         // normalize_sample_counts
    int plain_kernel = 0;     // 0: MFMA Gram (unweighted) / shared-node kernel (weighted); 1: dense pair tiles
    // per-sample sums S in segments that depend on the node count alone, so that a
    // sample's S does not depend on the other samples of the table (refinement sub-tables)
    bool fixed_segments = false;
};

// The spec a table's own values run with.  FP32 VAW terms hold the node counts
// and the sample totals in FP32, and divide by sqrt(mi (m - mi)) with m - mi
// small near the root: exact when every value is an integer and any two sample
// totals sum below 2^24, off by up to 2.6e-5 relative otherwise (fractional
// counts, measured).  Such tables compute in FP64, and the caller rounds the
// FP64 matrix to FP32.  Other specs are returned as given.
ComputeSpec spec_for_table(const Table& t, ComputeSpec s);

// Phases of one execution, timed with HIP events on the plan's stream during
// execute(); PH_SPARSE is the posting-list share of an occupancy split (zero
// without one), PH_MAIN the main kernel alone.
enum Phase : int { PH_EMBED = 0, PH_PREP, PH_MAIN, PH_COMBINE, PH_TOTAL, PH_SPARSE, PH_COUNT };
constexpr int EV_SPLIT = PH_COUNT, EV_ROWS = PH_COUNT + 1;   // extra events: posting lists start, rows end

struct PlanImpl;

// A table's CSR (entries indptr[0] .. indptr[n_obs]) copied to device memory
// on a thread of its own, so that the copy overlaps the host work that comes
// before a plan (the tree, the sample totals); Plan::create takes it over.
class TableUpload {
public:
    static std::shared_ptr<TableUpload> start(const uint32_t* indptr, const uint32_t* indices, const double* data,
                                              uint32_t n_obs);
    ~TableUpload();
    struct Impl;
    std::unique_ptr<Impl> impl;
};

class Plan {
public:
    ~Plan();
    // Uploads the table and postorder arrays; returns nullptr and sets err on failure.
    // `pre`: the table's CSR already on its way to the device (TableUpload::start on
    // the same arrays), taken over by the plan instead of uploading it again.
    static std::unique_ptr<Plan> create(const Table& table, const PostorderNodes& nodes, const ComputeSpec& spec,
                                        std::string* err, const std::shared_ptr<TableUpload>& pre = nullptr);
    // Full n x n distance matrix into device memory (FP64 if spec.fp64, FP32 otherwise).
    bool execute(std::string* err);
    // The same with the table's CSR values, the per-sample totals and the result
    // in caller-owned device buffers (nullptr: the plan's own).  quiet: records no
    // events and does not synchronise, so the work can be captured into a graph;
    // every buffer the plan needs must already exist (run it once non-quiet first).
    // FP32 weighted and generalized plans mark the pairs whose distance is small
    // enough for the FP32 rounding of the proportions to exceed 1e-6 relative
    // (refines()); a non-quiet call recomputes them in FP64 before it returns.
    // d_flags (n + 1 words, optional) receives the marks instead of the plan's
    // own buffer; after a quiet call, refine() with the same buffers finishes
    // the job once the work has run.
    // d_sums (n doubles, optional) receives the per-sample sums S in place of the
    // plan's own buffer (the marks of the unnormalised weighted distance are tested
    // against them again by refine()).
    bool execute_into(const double* d_data, const double* d_totals, void* d_out, bool quiet, std::string* err,
                      uint32_t* d_flags = nullptr, double* d_sums = nullptr);
    // FP64 recomputation of the pairs marked in d_flags (nullptr: the plan's
    // own marks) on the sub-table of the samples involved, written into d_out
    // (nullptr: the plan's own result).  Synchronous; nothing to do when no
    // pair is marked.
    bool refine(const double* d_data, const double* d_totals, void* d_out, const uint32_t* d_flags, double* d_sums,
                std::string* err);
    bool refines() const;
    uint64_t refined_pairs() const;   // pairs the last refine() recomputed
    // Later executions compute only the pairs of stripes [start, stop) (stripe s
    // holds the pairs (k, (k + s + 1) mod n), src/unifrac_task_impl.hpp:367):
    // the tiles, pair blocks and posting-list columns that hold one.  Entries of
    // other pairs in the result are left unspecified.  start = stop = 0: every
    // pair again.  false (err set) when stop exceeds (n + 1) / 2.
    bool set_stripes(uint32_t start, uint32_t stop, std::string* err);
    // Launch on a caller's HIP stream (nullptr: the plan's own stream).
    void set_stream(void* stream);
    void* stream() const;
    bool copy_matrix(void* host, std::string* err) const;                 // n*n elements
    // execute() and copy_matrix() in one, the copy of each finished band of rows
    // overlapping the device work on the rows below (shared-node kernels)
    bool execute_to_host(void* host, std::string* err);
    bool copy_rows(uint32_t row0, uint32_t nrows, void* host, std::string* err) const;   // nrows*n elements
    bool copy_stripes(uint32_t start, uint32_t stop, double* const* host, std::string* err) const;
    // Faith PD over all non-root nodes (needs only the embedding).
    bool faith(double* host, std::string* err);
    const float* timings() const;  // ms per Phase, last execute()
    uint64_t sparse_terms() const;   // (pair, node) terms walked from posting lists in the last execute()
    uint32_t n_samples() const;
    uint64_t processed_nodes() const;  // nodes entering the pair accumulation
    // (pair, node) terms the last execute's main kernel evaluated, padding and
    // the full diagonal tiles included (shared-node kernel: visited nodes only)
    uint64_t kernel_terms() const;
    bool out_fp64() const;
    // occupancy split in use: nodes held by fewer than this fraction of the
    // samples go to the posting lists (0: none)
    double split_theta() const;
    const char* main_kernel() const;
    const void* device_output() const;

private:
    Plan();
    std::unique_ptr<PlanImpl> p_;
};

// true iff a HIP device is usable; err explains why not
bool device_available(std::string* err);

// Device memory blocks through the process-wide cache that plans allocate from
// (got: the block's size, at least bytes); pool_release_all returns every
// cached block to the runtime.
void* pool_alloc(size_t bytes, size_t* got, std::string* err);
void pool_free(void* p, size_t bytes);
void pool_release_all();

// bytes of device memory into pageable host memory through a process-wide
// pinned staging area: chunks copied over PCIe on `stream` while earlier
// chunks are spread into `host` by a pool of threads
bool copy_to_host(void* host, const void* dev, size_t bytes, void* stream, std::string* err);
// The same for byte ranges [off, off + len) of dev into host at the same
// offsets, in order; a range's copies wait on the device (on `stream`) for the
// events in `ready`, so later ranges can still be computing when this starts.
// prefault: while the first copy waits, touch the destination's pages (a
// freshly allocated result) so that the host copies do not take the faults.
// Bytes [off, off + len) of the device buffer, or, with rows > 1, a rectangle of
// `rows` rows of len bytes each, pitch bytes apart (the same layout on both sides)
struct CopyRange {
    size_t off, len;
    std::vector<void*> ready;   // hipEvent_t
    size_t rows = 1, pitch = 0;
};
bool copy_ranges_to_host(char* host, const char* dev, const std::vector<CopyRange>& ranges, void* stream,
                         std::string* err, bool prefault = false);
// the reverse, host to device; returns once the host buffer has been read
bool copy_to_device(void* dev, const void* host, size_t bytes, void* stream, std::string* err);

}  // namespace ssu
