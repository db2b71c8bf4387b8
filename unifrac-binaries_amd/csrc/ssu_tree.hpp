// Phylogeny held as balanced parentheses, and the postorder node list the
// device kernels consume.
//
// Behaviour follows the reference BPTree (src/tree.hpp, src/tree.cpp):
// newick parsing with its quoting / single-descendant / length-as-float rules,
// shear to a tip set, collapse of single-child nodes.  Index caches are built
// in O(n) with explicit stacks instead of the reference's backward scans.
#pragma once

#include <cstdint>
#include <string>
#include <unordered_map>
#include <unordered_set>
#include <vector>

namespace ssu {

class PhyloTree {
public:
    std::vector<uint8_t> open;        // 1 = '(' , 0 = ')'
    std::vector<double> length;       // branch length, stored at the open position
    std::vector<std::string> name;    // node name, stored at the open position
    std::vector<uint32_t> partner;    // matching parenthesis
    std::vector<int32_t> parent_open; // open position of the enclosing node, -1 for the root

    PhyloTree() = default;
    // src/tree.cpp:30-70 (BPTree(newick))
    static PhyloTree from_newick(const char* newick);
    // src/tree.cpp:72-99 (BPTree(bool*, double*, char**, n))
    static PhyloTree from_arrays(const bool* structure, const double* lengths, const char* const* names, int n);
    static PhyloTree from_vectors(std::vector<uint8_t> open, std::vector<double> length, std::vector<std::string> name);

    uint32_t nparens() const { return (uint32_t)open.size(); }
    bool valid() const { return valid_; }
    bool is_leaf(uint32_t i) const { return open[i] && i + 1 < open.size() && !open[i + 1]; }  // src/tree.cpp:264-266
    uint32_t close_of(uint32_t i) const { return open[i] ? partner[i] : i; }
    uint32_t open_of(uint32_t i) const { return open[i] ? i : partner[i]; }

    // names of all tips (src/tree.cpp:114-124)
    std::unordered_set<std::string> tip_names() const;
    size_t n_tips() const;

    // src/tree.cpp:145-163
    PhyloTree shear(const std::unordered_set<std::string>& keep) const;
    // src/tree.cpp:165-191
    PhyloTree collapse() const;

private:
    bool valid_ = false;
    void index();
    PhyloTree masked(const std::vector<uint8_t>& keep, const std::vector<double>& lengths) const;
};

// Non-root nodes in postorder (src/unifrac_cmp.cpp:100-131: k = 0 .. nparens/2 - 2).
struct PostorderNodes {
    uint32_t m = 0;                   // number of non-root nodes
    std::vector<int32_t> parent;      // postorder index of the parent; m for the root
    std::vector<double> length;
    std::vector<uint8_t> is_leaf;
    std::vector<int32_t> obs_row;     // table row for a leaf, -1 otherwise
    std::vector<int32_t> height;      // 0 for leaves, 1 + max(child) otherwise
};

PostorderNodes postorder_nodes(const PhyloTree& t, const std::unordered_map<std::string, uint32_t>& obs_index);

// Observation id -> table row over the caller's strings: 64 open-addressing
// tables picked by the top bits of a 64-bit hash, each slot holding the hash
// (strings are compared only when hashes agree), built on several threads
// for large tables (each table's ids inserted in row order by one thread); for
// repeated ids the last row wins, as in Table::obs_index.
class NameIndex {
public:
    NameIndex(const char* const* ids, uint32_t n);
    int32_t find(const char* name) const;   // -1: absent
    uint32_t distinct() const { return distinct_; }
private:
    static constexpr uint32_t PARTS = 64;
    struct Slot { uint64_t hash; int32_t row; };   // row -1: empty
    std::vector<Slot> slot_;
    uint32_t base_[PARTS] = {}, mask_[PARTS] = {};
    const char* const* ids_;
    uint32_t distinct_ = 0;
};

// The tree given as parentheses arrays, sheared to the leaves whose names are
// table rows, collapsed and walked in postorder: the same nodes as
// postorder_nodes(PhyloTree::from_arrays(...).shear(ids).collapse(), rows),
// in two passes over the parentheses without building the intermediate trees.
// matched: distinct table ids that name a leaf.  false: unbalanced parentheses.
bool synced_postorder(const bool* structure, const double* lengths, const char* const* names, uint32_t n,
                      const NameIndex& rows, PostorderNodes& out, uint32_t& matched);

}  // namespace ssu
