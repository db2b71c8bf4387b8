// Rarefaction replicates on the device: each replicate subsamples every kept
// sample of one table to the same depth (src/biom_subsampled.cpp:155-262:
// with replacement = multinomial draws in proportion to the counts; without =
// a uniform sample of `depth` of the sample's items), then computes one or more
// distance matrices from it.  All replicates and methods run on one stream; the
// sequence is captured once into a HIP graph and replayed by later executions.
#pragma once

#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "ssu_engine.hpp"

namespace ssu {

struct RarefyImpl;

class RarefyBatch {
public:
    ~RarefyBatch();
    // Samples whose total is below depth are dropped first (src/biom_subsampled.cpp:290).
    // With no specs the batch only draws the tables (copy_counts).
    static std::unique_ptr<RarefyBatch> create(const Table& table, const PhyloTree& tree,
                                               const std::vector<ComputeSpec>& specs, uint32_t depth,
                                               bool with_replacement, uint64_t seed, uint32_t reps,
                                               std::string* err);
    bool execute(std::string* err);       // all replicates x methods; graph replay after the first call
    uint32_t n_samples() const;           // kept samples
    uint32_t n_obs() const;
    uint64_t nnz() const;
    uint32_t reps() const;
    const Table& kept() const;            // the kept-sample table (its CSR order is the counts' order)
    const std::vector<uint32_t>& sample_index() const;   // original index of each kept sample
    const std::vector<uint32_t>& obs_index() const;      // original observation row of each kept row
    bool copy_matrix(size_t spec, uint32_t rep, void* host, std::string* err) const;
    bool copy_counts(uint32_t rep, double* host, std::string* err) const;
    bool out_fp64(size_t spec) const;
    float last_ms() const;                // device time of the last execute()
    bool graph_used() const;              // the last execute() replayed the captured graph
    uint64_t pair_nodes() const;          // sum over replicates and methods of n(n-1)/2 * processed nodes
    uint64_t refined_pairs() const;       // pairs the last execute() recomputed in FP64 (all replicates)

private:
    friend class RarefyGroup;
    RarefyBatch();
    std::unique_ptr<RarefyImpl> p_;
};

struct RarefyGroupImpl;

// Several batches (tables of uneven sample counts) in one HIP graph launch:
// every batch's captured graph is a child node of one parent graph, with no
// dependence between them, so the device runs the tables side by side.  The
// batches stay owned by the caller and must outlive the group; each keeps its
// own outputs, as after its own execute().
class RarefyGroup {
public:
    ~RarefyGroup();
    // batches not yet captured run their first (eager) execute here
    static std::unique_ptr<RarefyGroup> create(const std::vector<RarefyBatch*>& batches, std::string* err);
    bool execute(std::string* err);   // one launch of the parent graph, then each batch's refinement
    float last_ms() const;            // device time of the last execute(), launch to last kernel

private:
    RarefyGroup();
    std::unique_ptr<RarefyGroupImpl> p_;
};

}  // namespace ssu
