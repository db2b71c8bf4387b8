// C ABI of libssu.so (include/ssu.h): the reference's src/api.hpp entry points
// over the gfx950 device pipeline.
#include "ssu.h"

#include <fcntl.h>
#include <sys/mman.h>
#include <unistd.h>

#include <charconv>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <memory>
#include <random>
#include <sstream>
#include <string>
#include <thread>
#include <vector>

#include "ssu_biom.hpp"
#include "ssu_engine.hpp"
#include "ssu_rarefy.hpp"
#include "ssu_stats.hpp"
#include "ssu_tsv.hpp"
#include "ssu_tune.hpp"
#include "ssu_results.hpp"
#include "ssu_table.hpp"
#include "ssu_tree.hpp"

using ssu::ComputeSpec;
using ssu::Method;
using ssu::PhyloTree;
using ssu::Plan;
using ssu::Table;

namespace ssu {
namespace {
thread_local std::string g_last_error;
}
void set_error(const std::string& e) {
    g_last_error = e;
    if (!e.empty()) fprintf(stderr, "ssu: %s\n", e.c_str());
}
const std::string& last_error() { return g_last_error; }
}  // namespace ssu

namespace {

std::mt19937_64 g_rng(5489u);

// Wall-clock phases of the calling thread's last one-off matrix call, in ms:
// inputs (table and tree structures), tree (shear, collapse, postorder), plan
// (uploads, schedules, occupancy splits), execute (device work, refinement
// included), copy (result to the host), total.
enum CallPhase { CP_INPUT = 0, CP_TREE, CP_PLAN, CP_EXECUTE, CP_COPY, CP_TOTAL, CP_COUNT };
thread_local double g_call_ms[CP_COUNT];
using Clock = std::chrono::steady_clock;
double ms_since(Clock::time_point t0) {
    return std::chrono::duration<double, std::milli>(Clock::now() - t0).count();
}
using ssu::destroy_full;
using ssu::dup_ids;
using ssu::init_full;
using ssu::set_error;

// src/api.cpp:55-88 (SET_METHOD): bare names run in FP32, "_fp64" in FP64
bool parse_method(const char* name, Method& m, bool& fp64) {
    if (!name) return false;
    std::string s(name);
    fp64 = false;
    auto strip = [&](const char* suffix) {
        const size_t l = strlen(suffix);
        if (s.size() > l && s.compare(s.size() - l, l, suffix) == 0) { s.resize(s.size() - l); return true; }
        return false;
    };
    if (strip("_fp64")) fp64 = true;
    else strip("_fp32");
    if (s == "unweighted") m = Method::unweighted;
    else if (s == "weighted_normalized") m = Method::weighted_normalized;
    else if (s == "weighted_unnormalized") m = Method::weighted_unnormalized;
    else if (s == "unweighted_unnormalized") m = Method::unweighted_unnormalized;
    else if (s == "generalized") m = Method::generalized;
    else return false;
    return true;
}


// src/api.cpp:94-100 (VALIDATE_TREE_TABLE)
ComputeStatus validate(const Table& t, const PhyloTree& tree) {
    if (t.n_samples == 0 || t.n_obs == 0) return table_empty;
    if (!tree.valid()) return tree_missing;
    const auto tips = tree.tip_names();
    for (const auto& id : t.obs_ids)
        if (!tips.count(id)) return table_and_tree_do_not_overlap;
    return okay;
}

// src/api.cpp:90-92 (SYNC_TREE_TABLE) followed by the postorder walk
ssu::PostorderNodes synced_nodes(const Table& t, const PhyloTree& tree) {
    std::unordered_set<std::string> keep(t.obs_ids.begin(), t.obs_ids.end());
    PhyloTree s = tree.shear(keep).collapse();
    return ssu::postorder_nodes(s, t.obs_index);
}

ComputeSpec make_spec(Method m, bool fp64, bool vaw, double alpha, bool bypass, bool normalize) {
    ComputeSpec s;
    s.method = m;
    s.fp64 = fp64;
    s.vaw = vaw;
    s.alpha = alpha;
    s.bypass_tips = bypass;
    s.normalize = normalize;
    return s;
}

// Full n x n matrix of the given element type in host memory.
template <class TOut>
ComputeStatus run_full(const Table& t, const ssu::PostorderNodes& nodes, const ComputeSpec& spec, TOut* out,
                       const std::shared_ptr<ssu::TableUpload>& pre = nullptr) {
    std::string err;
    auto t0 = Clock::now();
    std::unique_ptr<Plan> plan = Plan::create(t, nodes, ssu::spec_for_table(t, spec), &err, pre);
    if (!plan) { set_error(err); return output_error; }
    g_call_ms[CP_PLAN] = ms_since(t0);
    t0 = Clock::now();
    const size_t nn = (size_t)t.n_samples * t.n_samples;
    if (plan->out_fp64() == (sizeof(TOut) == sizeof(double))) {
        // device work and the copy to the host together, overlapped where the path allows
        if (!plan->execute_to_host(out, &err)) { set_error(err); return output_error; }
        g_call_ms[CP_EXECUTE] = ms_since(t0);
        g_call_ms[CP_COPY] = 0.0;
        return okay;
    }
    if (!plan->execute(&err)) { set_error(err); return output_error; }
    g_call_ms[CP_EXECUTE] = ms_since(t0);
    t0 = Clock::now();
    struct CopyTime {
        Clock::time_point t0;
        ~CopyTime() { g_call_ms[CP_COPY] = ms_since(t0); }
    } copy_time{t0};
    if (plan->out_fp64()) {
        std::vector<double> tmp(nn);
        if (!plan->copy_matrix(tmp.data(), &err)) { set_error(err); return output_error; }
        for (size_t i = 0; i < nn; ++i) out[i] = (TOut)tmp[i];
    } else {
        std::vector<float> tmp(nn);
        if (!plan->copy_matrix(tmp.data(), &err)) { set_error(err); return output_error; }
        for (size_t i = 0; i < nn; ++i) out[i] = (TOut)tmp[i];
    }
    return okay;
}

template <class TOut>
ComputeStatus run_full(const Table& t, const PhyloTree& tree, const ComputeSpec& spec, TOut* out) {
    const auto t0 = Clock::now();
    ssu::PostorderNodes nodes = synced_nodes(t, tree);
    g_call_ms[CP_TREE] = ms_since(t0);
    return run_full<TOut>(t, nodes, spec, out);
}

// A rarefied table: the kept samples and rows of the batch with one
// replicate's counts, zero entries and rows left empty dropped
// (src/biom_subsampled.cpp:298-321, steal_nonzero).
Table rarefied_table(const Table& k, const std::vector<double>& counts) {
    const uint32_t base = k.n_obs ? k.indptr[0] : 0;
    std::vector<std::string> oids;
    std::vector<uint32_t> ip{0}, ix;
    std::vector<double> d;
    for (uint32_t i = 0; i < k.n_obs; ++i) {
        const size_t before = ix.size();
        for (uint32_t j = k.indptr[i]; j < k.indptr[i + 1]; ++j)
            if (counts[j - base] > 0) { ix.push_back(k.indices[j]); d.push_back(counts[j - base]); }
        if (ix.size() > before) { oids.push_back(k.obs_ids[i]); ip.push_back((uint32_t)ix.size()); }
    }
    return Table::own(std::move(oids), std::vector<std::string>(k.sample_ids), std::move(ip), std::move(ix),
                      std::move(d));
}

// Rarefaction (src/biom_subsampled.cpp:277-321) drawn on the device: samples
// below the depth are dropped, the rest drawn down to `depth` counts.  The
// generator is this library's counter-based one, keyed from the
// ssu_set_random_seed stream, so only the statistics match the reference's.
ComputeStatus rarefied_tables(const Table& t, const PhyloTree& tree, uint32_t depth, bool with_replacement,
                              uint32_t reps, std::vector<Table>& out) {
    std::string err;
    auto b = ssu::RarefyBatch::create(t, tree, {}, depth, with_replacement, g_rng(), reps, &err);
    if (!b) {
        if (err.find("reaches the rarefaction depth") != std::string::npos) return table_empty;
        set_error(err);
        return output_error;
    }
    if (!b->execute(&err)) { set_error(err); return output_error; }
    std::vector<double> counts(b->nnz());
    for (uint32_t r = 0; r < reps; ++r) {
        if (!b->copy_counts(r, counts.data(), &err)) { set_error(err); return output_error; }
        out.push_back(rarefied_table(b->kept(), counts));
        if (out.back().n_obs == 0) return table_empty;
    }
    return okay;
}

// src/api.cpp:638-698 (one_off_matrix_T) + :700-714 (optional subsampling)
template <class TReal, class TMat>
ComputeStatus matrix_common(const Table& t0, const PhyloTree& tree, const char* method_name, bool vaw, double alpha,
                            bool bypass, bool normalize, unsigned subsample_depth, bool with_replacement,
                            const char* mmap_dir, TMat** result) {
    Method m;
    bool fp64;
    if (!parse_method(method_name, m, fp64)) return unknown_method;
    const Table* tp = &t0;
    Table sub;
    if (subsample_depth > 0) {
        std::vector<Table> subs;
        const ComputeStatus st = rarefied_tables(t0, tree, subsample_depth, with_replacement, 1, subs);
        if (st != okay) return st;
        sub = std::move(subs[0]);
        if (sub.n_samples == 0 || sub.n_obs == 0) return table_empty;
        tp = &sub;
    }
    const Table& t = *tp;
    if (*result == nullptr) {
        if (!init_full<TReal, TMat>(*result, t.sample_ids, mmap_dir)) {
            set_error("result allocation failed");
            return output_error;
        }
    }
    return run_full<TReal>(t, tree, make_spec(m, fp64, vaw, alpha, bypass, normalize), (*result)->matrix);
}

template <class TReal, class TMat>
ComputeStatus matrix_inmem(const support_biom_t* td, const support_bptree_t* tr, const char* method_name, bool vaw,
                           double alpha, bool bypass, bool normalize, unsigned depth, bool with_repl,
                           const char* mmap_dir, TMat** result, bool want_fp64) {
    // src/api.cpp:775-856: method must match the result precision
    Method m;
    bool fp64;
    if (!parse_method(method_name, m, fp64)) return unknown_method;
    if (fp64 != want_fp64) return invalid_method;
    if (!td || !tr || td->n_samples <= 0 || td->n_obs <= 0) return table_empty;
    for (double& x : g_call_ms) x = 0.0;
    const auto t0 = Clock::now();
    if (depth > 0) {   // rarefied first: the sampler shears the tree to the kept table itself
        Table t = Table::borrow(td->obs_ids, td->sample_ids, td->indptr, td->indices, td->data, (uint32_t)td->n_obs,
                                (uint32_t)td->n_samples);
        PhyloTree tree = PhyloTree::from_arrays(tr->structure, tr->lengths, tr->names, tr->n_parens);
        ComputeStatus v = validate(t, tree);
        if (v != okay) return v;
        g_call_ms[CP_INPUT] = ms_since(t0);
        v = matrix_common<TReal, TMat>(t, tree, method_name, vaw, alpha, bypass, normalize, depth, with_repl,
                                       mmap_dir, result);
        g_call_ms[CP_TOTAL] = ms_since(t0);
        return v;
    }
    // the table's CSR on its way to the device from the start; the tree sheared,
    // collapsed and walked in postorder straight from its arrays (src/api.cpp:90-100
    // SYNC_TREE_TABLE, VALIDATE_TREE_TABLE) on a second thread, while this one
    // borrows the table (sample totals) without an id map
    std::shared_ptr<ssu::TableUpload> pre;
    if (ssu::device_available(nullptr))
        pre = ssu::TableUpload::start(td->indptr, td->indices, td->data, (uint32_t)td->n_obs);
    ssu::PostorderNodes nodes;
    uint32_t matched = 0, distinct = 0;
    bool tree_ok = false;
    double tree_ms = 0.0;
    std::thread tree_thread([&] {
        const auto t1 = Clock::now();
        ssu::NameIndex idx(td->obs_ids, (uint32_t)td->n_obs);
        distinct = idx.distinct();
        const double idx_ms = ms_since(t1);
        tree_ok = tr->n_parens > 0 && ssu::synced_postorder(tr->structure, tr->lengths, tr->names,
                                                            (uint32_t)tr->n_parens, idx, nodes, matched);
        tree_ms = ms_since(t1);
        static const bool timing = getenv("SSU_TIMING") && atoi(getenv("SSU_TIMING")) != 0;
        if (timing)   // SSU_TIMING=1: the tree thread's two steps on stderr
            fprintf(stderr, "[ssu tree] name index %7.2f ms, shear and postorder %7.2f ms\n", idx_ms,
                    tree_ms - idx_ms);
    });
    Table t = Table::borrow(td->obs_ids, td->sample_ids, td->indptr, td->indices, td->data, (uint32_t)td->n_obs,
                            (uint32_t)td->n_samples, false);
    g_call_ms[CP_INPUT] = ms_since(t0);
    tree_thread.join();
    g_call_ms[CP_TREE] = tree_ms;   // (overlapped with the inputs)
    if (!tree_ok) return tree_missing;
    if (matched < distinct) return table_and_tree_do_not_overlap;
    if (*result == nullptr && !init_full<TReal, TMat>(*result, t.sample_ids, mmap_dir)) {
        set_error("result allocation failed");
        return output_error;
    }
    const ComputeStatus v =
        run_full<TReal>(t, nodes, make_spec(m, fp64, vaw, alpha, bypass, normalize), (*result)->matrix, pre);
    g_call_ms[CP_TOTAL] = ms_since(t0);
    return v;
}

bool file_exists(const char* f) {
    std::ifstream in(f);
    return in.good();
}

std::string read_file(const char* f) {
    std::ifstream in(f);
    std::stringstream ss;
    ss << in.rdbuf();
    return ss.str();
}

ComputeStatus load_table(const char* biom_filename, Table& t) {
    if (!biom_filename || !file_exists(biom_filename)) return table_missing;
    std::string err;
    if (!ssu::load_biom(biom_filename, t, &err)) {
        set_error(err);
        return table_missing;
    }
    return okay;
}

template <class TReal, class TMat>
ComputeStatus matrix_files(const char* biom_filename, const PhyloTree* tree_in, const char* tree_filename,
                           const char* method_name, bool vaw, double alpha, bool bypass, bool normalize, unsigned depth,
                           bool with_repl, const char* mmap_dir, TMat** result) {
    if (!tree_in) {
        if (!biom_filename || !file_exists(biom_filename)) return table_missing;
        if (!tree_filename || !file_exists(tree_filename)) return tree_missing;
    } else if (!biom_filename || !file_exists(biom_filename)) {
        return table_missing;
    }
    PhyloTree parsed;
    if (!tree_in) parsed = PhyloTree::from_newick(read_file(tree_filename).c_str());
    const PhyloTree& tree = tree_in ? *tree_in : parsed;
    Table t;
    ComputeStatus st = load_table(biom_filename, t);
    if (st != okay) return st;
    st = validate(t, tree);
    if (st != okay) return st;
    return matrix_common<TReal, TMat>(t, tree, method_name, vaw, alpha, bypass, normalize, depth, with_repl, mmap_dir,
                                      result);
}

// Condensed upper triangle (src/api.cpp:492-525, one_off_inmem_cpp), computed from the full matrix.
ComputeStatus condensed_common(const Table& t, const PhyloTree& tree, const char* method_name, bool vaw, double alpha,
                               bool bypass, bool normalize, mat_t** result) {
    Method m;
    bool fp64;
    if (!parse_method(method_name, m, fp64)) return unknown_method;
    const uint64_t n = t.n_samples;
    std::vector<double> full(n * n);
    ComputeStatus st = run_full<double>(t, tree, make_spec(m, fp64, vaw, alpha, bypass, normalize), full.data());
    if (st != okay) return st;
    mat_t* r = (mat_t*)malloc(sizeof(mat_t));
    r->n_samples = (unsigned)n;
    r->cf_size = (unsigned)(n * (n - 1) / 2);
    r->is_upper_triangle = true;
    r->sample_ids = dup_ids(t.sample_ids);
    r->condensed_form = (double*)malloc(sizeof(double) * (r->cf_size ? r->cf_size : 1));
    uint64_t c = 0;
    for (uint64_t i = 0; i < n; ++i)
        for (uint64_t j = i + 1; j < n; ++j) r->condensed_form[c++] = full[i * n + j];
    *result = r;
    return okay;
}

ComputeStatus condensed_files(const char* biom_filename, const PhyloTree* tree_in, const char* tree_filename,
                              const char* method_name, bool vaw, double alpha, bool bypass, bool normalize,
                              mat_t** result) {
    if (!biom_filename || !file_exists(biom_filename)) return table_missing;
    if (!tree_in && (!tree_filename || !file_exists(tree_filename))) return tree_missing;
    PhyloTree parsed;
    if (!tree_in) parsed = PhyloTree::from_newick(read_file(tree_filename).c_str());
    const PhyloTree& tree = tree_in ? *tree_in : parsed;
    Table t;
    ComputeStatus st = load_table(biom_filename, t);
    if (st != okay) return st;
    st = validate(t, tree);
    if (st != okay) return st;
    return condensed_common(t, tree, method_name, vaw, alpha, bypass, normalize, result);
}

// src/api.cpp:723-773 (one_dense_pair): element [0][1] of the 2-sample matrix
ComputeStatus dense_pair(unsigned n_obs, const char** obs_ids, const double* s1, const double* s2,
                         const PhyloTree& tree, const char* method_name, bool vaw, double alpha, bool bypass,
                         bool normalize, double* result) {
    Method m;
    bool fp64;
    if (!parse_method(method_name, m, fp64)) return unknown_method;
    Table t = Table::dense_pair(obs_ids, s1, s2, n_obs);
    ComputeStatus st = validate(t, tree);
    if (st != okay) return st;
    double full[4];
    st = run_full<double>(t, tree, make_spec(m, fp64, vaw, alpha, bypass, normalize), full);
    if (st == okay) *result = full[1];
    return st;
}

ComputeStatus faith_common(const Table& t0, const PhyloTree& tree, r_vec** result) {
    // src/api.cpp:569-604: drop zero-count samples (total < 1), shear, Faith PD
    if (t0.n_samples == 0 || t0.n_obs == 0) return table_empty;
    Table nz = t0.filtered(1.0);
    const bool degenerate = nz.n_samples == 0 || nz.n_obs == 0;
    if (degenerate) fprintf(stderr, "WARNING: All samples had zero counts. Forcing zero result.\n");
    else if (nz.n_samples != t0.n_samples || nz.n_obs != t0.n_obs)
        fprintf(stderr, "WARNING: Some samples had zero counts and were filtered out.\n");
    const Table& t = degenerate ? t0 : nz;
    r_vec* r = (r_vec*)malloc(sizeof(r_vec));
    r->n_samples = t.n_samples;
    r->values = (double*)calloc(t.n_samples ? t.n_samples : 1, sizeof(double));
    r->sample_ids = dup_ids(t.sample_ids);
    if (!degenerate) {
        std::string err;
        ssu::PostorderNodes nodes = synced_nodes(t, tree);
        std::unique_ptr<Plan> plan = Plan::create(t, nodes, ComputeSpec{}, &err);
        if (!plan || !plan->faith(r->values, &err)) {
            set_error(err);
            r_vec* rr = r;
            destroy_results_vec(&rr);
            return output_error;
        }
    }
    *result = r;
    return okay;
}

}  // namespace

// ===================================================================== trees

void ssu_set_random_seed(unsigned int new_seed) { g_rng.seed(new_seed); }

IOStatus read_bptree_opaque(const char* tree_filename, opaque_bptree_t** tree_data) {
    if (!tree_data) return unexpected_end;
    if (!tree_filename || !file_exists(tree_filename)) return open_error;
    *tree_data = (opaque_bptree_t*)new PhyloTree(PhyloTree::from_newick(read_file(tree_filename).c_str()));
    return read_okay;
}

void load_bptree_opaque(const char* newick, opaque_bptree_t** tree_data) {
    *tree_data = (opaque_bptree_t*)new PhyloTree(PhyloTree::from_newick(newick));
}

void convert_bptree_opaque(const support_bptree_t* in, opaque_bptree_t** tree_data) {
    *tree_data = (opaque_bptree_t*)new PhyloTree(
        PhyloTree::from_arrays(in->structure, in->lengths, in->names, in->n_parens));
}

void destroy_bptree_opaque(opaque_bptree_t** tree_data) {
    if (tree_data) {
        delete (PhyloTree*)(*tree_data);
        *tree_data = nullptr;
    }
}

// src/api.cpp:166-172: the number of distinct tip names
int get_bptree_opaque_els(opaque_bptree_t* tree_data) {
    return tree_data ? (int)((PhyloTree*)tree_data)->n_tips() : 0;
}

// ================================================================ destructors

void destroy_mat(mat_t** result) {
    if (!result || !*result) return;
    for (unsigned i = 0; i < (*result)->n_samples; ++i) free((*result)->sample_ids[i]);
    free((*result)->sample_ids);
    free((*result)->condensed_form);
    free(*result);
    *result = nullptr;
}

void destroy_mat_full_fp64(mat_full_fp64_t** result) { destroy_full<double, mat_full_fp64_t>(result); }
void destroy_mat_full_fp32(mat_full_fp32_t** result) { destroy_full<float, mat_full_fp32_t>(result); }

void destroy_partial_mat(partial_mat_t** result) {
    if (!result || !*result) return;
    partial_mat_t* r = *result;
    for (uint32_t i = 0; i < r->n_samples; ++i) free(r->sample_ids[i]);
    free(r->sample_ids);
    for (uint32_t i = 0; i < r->stripe_stop - r->stripe_start; ++i) free(r->stripes[i]);
    free(r->stripes);
    free(r);
    *result = nullptr;
}

void destroy_partial_dyn_mat(partial_dyn_mat_t** result) {
    if (!result || !*result) return;
    partial_dyn_mat_t* r = *result;
    for (uint32_t i = 0; i < r->n_samples; ++i) free(r->sample_ids[i]);
    free(r->sample_ids);
    if (r->stripes)
        for (uint32_t i = 0; i < r->stripe_stop - r->stripe_start; ++i) free(r->stripes[i]);
    free(r->stripes);
    free(r->offsets);
    free(r->filename);
    free(r);
    *result = nullptr;
}

void destroy_results_vec(r_vec** result) {
    if (!result || !*result) return;
    for (unsigned i = 0; i < (*result)->n_samples; ++i) free((*result)->sample_ids[i]);
    free((*result)->sample_ids);
    free((*result)->values);
    free(*result);
    *result = nullptr;
}

// =================================================================== one_off

ComputeStatus one_off_v3(const char* biom_filename, const char* tree_filename, const char* unifrac_method,
                         bool variance_adjust, double alpha, bool bypass_tips, bool normalize_sample_counts,
                         unsigned int, mat_t** result) {
    return condensed_files(biom_filename, nullptr, tree_filename, unifrac_method, variance_adjust, alpha, bypass_tips,
                           normalize_sample_counts, result);
}

ComputeStatus one_off_wtree_v3(const char* biom_filename, const opaque_bptree_t* tree_data, const char* unifrac_method,
                               bool variance_adjust, double alpha, bool bypass_tips, bool normalize_sample_counts,
                               unsigned int, mat_t** result) {
    if (!tree_data) return tree_missing;
    return condensed_files(biom_filename, (const PhyloTree*)tree_data, nullptr, unifrac_method, variance_adjust, alpha,
                           bypass_tips, normalize_sample_counts, result);
}

ComputeStatus one_off(const char* biom_filename, const char* tree_filename, const char* unifrac_method,
                      bool variance_adjust, double alpha, bool bypass_tips, unsigned int n_substeps, mat_t** result) {
    return one_off_v3(biom_filename, tree_filename, unifrac_method, variance_adjust, alpha, bypass_tips, true,
                      n_substeps, result);
}

ComputeStatus one_off_wtree(const char* biom_filename, const opaque_bptree_t* tree_data, const char* unifrac_method,
                            bool variance_adjust, double alpha, bool bypass_tips, unsigned int n_substeps,
                            mat_t** result) {
    return one_off_wtree_v3(biom_filename, tree_data, unifrac_method, variance_adjust, alpha, bypass_tips, true,
                            n_substeps, result);
}

// ============================================================ in-memory matrix

ComputeStatus one_off_matrix_inmem_v3(const support_biom_t* table_data, const support_bptree_t* tree_data,
                                      const char* unifrac_method, bool variance_adjust, double alpha, bool bypass_tips,
                                      bool normalize_sample_counts, unsigned int, unsigned int subsample_depth,
                                      bool subsample_with_replacement, const char* mmap_dir,
                                      mat_full_fp64_t** result) {
    return matrix_inmem<double, mat_full_fp64_t>(table_data, tree_data, unifrac_method, variance_adjust, alpha,
                                                 bypass_tips, normalize_sample_counts, subsample_depth,
                                                 subsample_with_replacement, mmap_dir, result, true);
}

ComputeStatus one_off_matrix_inmem_v2(const support_biom_t* table_data, const support_bptree_t* tree_data,
                                      const char* unifrac_method, bool variance_adjust, double alpha, bool bypass_tips,
                                      unsigned int n_substeps, unsigned int subsample_depth,
                                      bool subsample_with_replacement, const char* mmap_dir,
                                      mat_full_fp64_t** result) {
    return one_off_matrix_inmem_v3(table_data, tree_data, unifrac_method, variance_adjust, alpha, bypass_tips, true,
                                   n_substeps, subsample_depth, subsample_with_replacement, mmap_dir, result);
}

ComputeStatus one_off_inmem(const support_biom_t* table_data, const support_bptree_t* tree_data,
                            const char* unifrac_method, bool variance_adjust, double alpha, bool bypass_tips,
                            unsigned int n_substeps, mat_full_fp64_t** result) {
    return one_off_matrix_inmem_v3(table_data, tree_data, unifrac_method, variance_adjust, alpha, bypass_tips, true,
                                   n_substeps, 0, true, nullptr, result);
}

ComputeStatus one_off_matrix_inmem_fp32_v3(const support_biom_t* table_data, const support_bptree_t* tree_data,
                                           const char* unifrac_method, bool variance_adjust, double alpha,
                                           bool bypass_tips, bool normalize_sample_counts, unsigned int,
                                           unsigned int subsample_depth, bool subsample_with_replacement,
                                           const char* mmap_dir, mat_full_fp32_t** result) {
    return matrix_inmem<float, mat_full_fp32_t>(table_data, tree_data, unifrac_method, variance_adjust, alpha,
                                                bypass_tips, normalize_sample_counts, subsample_depth,
                                                subsample_with_replacement, mmap_dir, result, false);
}

ComputeStatus one_off_matrix_inmem_fp32_v2(const support_biom_t* table_data, const support_bptree_t* tree_data,
                                           const char* unifrac_method, bool variance_adjust, double alpha,
                                           bool bypass_tips, unsigned int n_substeps, unsigned int subsample_depth,
                                           bool subsample_with_replacement, const char* mmap_dir,
                                           mat_full_fp32_t** result) {
    return one_off_matrix_inmem_fp32_v3(table_data, tree_data, unifrac_method, variance_adjust, alpha, bypass_tips,
                                        true, n_substeps, subsample_depth, subsample_with_replacement, mmap_dir,
                                        result);
}

ComputeStatus one_off_inmem_fp32(const support_biom_t* table_data, const support_bptree_t* tree_data,
                                 const char* unifrac_method, bool variance_adjust, double alpha, bool bypass_tips,
                                 unsigned int n_substeps, mat_full_fp32_t** result) {
    return one_off_matrix_inmem_fp32_v3(table_data, tree_data, unifrac_method, variance_adjust, alpha, bypass_tips,
                                        true, n_substeps, 0, true, nullptr, result);
}

// ================================================================ file matrix

ComputeStatus one_off_matrix_v3(const char* biom_filename, const char* tree_filename, const char* unifrac_method,
                                bool variance_adjust, double alpha, bool bypass_tips, bool normalize_sample_counts,
                                unsigned int, unsigned int subsample_depth, bool subsample_with_replacement,
                                const char* mmap_dir, mat_full_fp64_t** result) {
    return matrix_files<double, mat_full_fp64_t>(biom_filename, nullptr, tree_filename, unifrac_method,
                                                 variance_adjust, alpha, bypass_tips, normalize_sample_counts,
                                                 subsample_depth, subsample_with_replacement, mmap_dir, result);
}

ComputeStatus one_off_matrix_v3t(const char* biom_filename, const opaque_bptree_t* tree_data,
                                 const char* unifrac_method, bool variance_adjust, double alpha, bool bypass_tips,
                                 bool normalize_sample_counts, unsigned int, unsigned int subsample_depth,
                                 bool subsample_with_replacement, const char* mmap_dir, mat_full_fp64_t** result) {
    if (!tree_data) return tree_missing;
    return matrix_files<double, mat_full_fp64_t>(biom_filename, (const PhyloTree*)tree_data, nullptr, unifrac_method,
                                                 variance_adjust, alpha, bypass_tips, normalize_sample_counts,
                                                 subsample_depth, subsample_with_replacement, mmap_dir, result);
}

ComputeStatus one_off_matrix_v2(const char* biom_filename, const char* tree_filename, const char* unifrac_method,
                                bool variance_adjust, double alpha, bool bypass_tips, unsigned int n_substeps,
                                unsigned int subsample_depth, bool subsample_with_replacement, const char* mmap_dir,
                                mat_full_fp64_t** result) {
    return one_off_matrix_v3(biom_filename, tree_filename, unifrac_method, variance_adjust, alpha, bypass_tips, true,
                             n_substeps, subsample_depth, subsample_with_replacement, mmap_dir, result);
}

ComputeStatus one_off_matrix_v2t(const char* biom_filename, const opaque_bptree_t* tree_data,
                                 const char* unifrac_method, bool variance_adjust, double alpha, bool bypass_tips,
                                 unsigned int n_substeps, unsigned int subsample_depth,
                                 bool subsample_with_replacement, const char* mmap_dir, mat_full_fp64_t** result) {
    return one_off_matrix_v3t(biom_filename, tree_data, unifrac_method, variance_adjust, alpha, bypass_tips, true,
                              n_substeps, subsample_depth, subsample_with_replacement, mmap_dir, result);
}

ComputeStatus one_off_matrix(const char* biom_filename, const char* tree_filename, const char* unifrac_method,
                             bool variance_adjust, double alpha, bool bypass_tips, unsigned int n_substeps,
                             const char* mmap_dir, mat_full_fp64_t** result) {
    return one_off_matrix_v3(biom_filename, tree_filename, unifrac_method, variance_adjust, alpha, bypass_tips, true,
                             n_substeps, 0, true, mmap_dir, result);
}

ComputeStatus one_off_matrix_fp32_v3(const char* biom_filename, const char* tree_filename, const char* unifrac_method,
                                     bool variance_adjust, double alpha, bool bypass_tips,
                                     bool normalize_sample_counts, unsigned int, unsigned int subsample_depth,
                                     bool subsample_with_replacement, const char* mmap_dir,
                                     mat_full_fp32_t** result) {
    return matrix_files<float, mat_full_fp32_t>(biom_filename, nullptr, tree_filename, unifrac_method,
                                                variance_adjust, alpha, bypass_tips, normalize_sample_counts,
                                                subsample_depth, subsample_with_replacement, mmap_dir, result);
}

ComputeStatus one_off_matrix_fp32_v3t(const char* biom_filename, const opaque_bptree_t* tree_data,
                                      const char* unifrac_method, bool variance_adjust, double alpha,
                                      bool bypass_tips, bool normalize_sample_counts, unsigned int,
                                      unsigned int subsample_depth, bool subsample_with_replacement,
                                      const char* mmap_dir, mat_full_fp32_t** result) {
    if (!tree_data) return tree_missing;
    return matrix_files<float, mat_full_fp32_t>(biom_filename, (const PhyloTree*)tree_data, nullptr, unifrac_method,
                                                variance_adjust, alpha, bypass_tips, normalize_sample_counts,
                                                subsample_depth, subsample_with_replacement, mmap_dir, result);
}

ComputeStatus one_off_matrix_fp32_v2(const char* biom_filename, const char* tree_filename, const char* unifrac_method,
                                     bool variance_adjust, double alpha, bool bypass_tips, unsigned int n_substeps,
                                     unsigned int subsample_depth, bool subsample_with_replacement,
                                     const char* mmap_dir, mat_full_fp32_t** result) {
    return one_off_matrix_fp32_v3(biom_filename, tree_filename, unifrac_method, variance_adjust, alpha, bypass_tips,
                                  true, n_substeps, subsample_depth, subsample_with_replacement, mmap_dir, result);
}

ComputeStatus one_off_matrix_fp32_v2t(const char* biom_filename, const opaque_bptree_t* tree_data,
                                      const char* unifrac_method, bool variance_adjust, double alpha,
                                      bool bypass_tips, unsigned int n_substeps, unsigned int subsample_depth,
                                      bool subsample_with_replacement, const char* mmap_dir,
                                      mat_full_fp32_t** result) {
    return one_off_matrix_fp32_v3t(biom_filename, tree_data, unifrac_method, variance_adjust, alpha, bypass_tips, true,
                                   n_substeps, subsample_depth, subsample_with_replacement, mmap_dir, result);
}

ComputeStatus one_off_matrix_fp32(const char* biom_filename, const char* tree_filename, const char* unifrac_method,
                                  bool variance_adjust, double alpha, bool bypass_tips, unsigned int n_substeps,
                                  const char* mmap_dir, mat_full_fp32_t** result) {
    return one_off_matrix_fp32_v3(biom_filename, tree_filename, unifrac_method, variance_adjust, alpha, bypass_tips,
                                  true, n_substeps, 0, true, mmap_dir, result);
}

// ================================================================ dense pair

ComputeStatus one_dense_pair_v3t(unsigned int n_obs, const char** obs_ids, const double* sample1,
                                 const double* sample2, const opaque_bptree_t* tree_data, const char* unifrac_method,
                                 bool variance_adjust, double alpha, bool bypass_tips, bool normalize_sample_counts,
                                 double* result) {
    if (!tree_data) return tree_missing;
    return dense_pair(n_obs, obs_ids, sample1, sample2, *(const PhyloTree*)tree_data, unifrac_method,
                      variance_adjust, alpha, bypass_tips, normalize_sample_counts, result);
}

ComputeStatus one_dense_pair_v3(unsigned int n_obs, const char** obs_ids, const double* sample1,
                                const double* sample2, const support_bptree_t* tree_data, const char* unifrac_method,
                                bool variance_adjust, double alpha, bool bypass_tips, bool normalize_sample_counts,
                                double* result) {
    if (!tree_data) return tree_missing;
    PhyloTree tree = PhyloTree::from_arrays(tree_data->structure, tree_data->lengths, tree_data->names,
                                            tree_data->n_parens);
    return dense_pair(n_obs, obs_ids, sample1, sample2, tree, unifrac_method, variance_adjust, alpha, bypass_tips,
                      normalize_sample_counts, result);
}

ComputeStatus one_dense_pair_v2t(unsigned int n_obs, const char** obs_ids, const double* sample1,
                                 const double* sample2, const opaque_bptree_t* tree_data, const char* unifrac_method,
                                 bool variance_adjust, double alpha, bool bypass_tips, double* result) {
    return one_dense_pair_v3t(n_obs, obs_ids, sample1, sample2, tree_data, unifrac_method, variance_adjust, alpha,
                              bypass_tips, true, result);
}

ComputeStatus one_dense_pair_v2(unsigned int n_obs, const char** obs_ids, const double* sample1,
                                const double* sample2, const support_bptree_t* tree_data, const char* unifrac_method,
                                bool variance_adjust, double alpha, bool bypass_tips, double* result) {
    return one_dense_pair_v3(n_obs, obs_ids, sample1, sample2, tree_data, unifrac_method, variance_adjust, alpha,
                             bypass_tips, true, result);
}

// ================================================================== Faith PD

ComputeStatus faith_pd_one_off(const char* biom_filename, const char* tree_filename, r_vec** result) {
    if (!biom_filename || !file_exists(biom_filename)) return table_missing;
    if (!tree_filename || !file_exists(tree_filename)) return tree_missing;
    PhyloTree tree = PhyloTree::from_newick(read_file(tree_filename).c_str());
    Table t;
    ComputeStatus st = load_table(biom_filename, t);
    if (st != okay) return st;
    st = validate(t, tree);
    if (st != okay) return st;
    return faith_common(t, tree, result);
}

ComputeStatus ssu_faith_pd_inmem(const support_biom_t* td, const support_bptree_t* tr, r_vec** result) {
    if (!td || !tr || td->n_samples <= 0 || td->n_obs <= 0) return table_empty;
    Table t = Table::borrow(td->obs_ids, td->sample_ids, td->indptr, td->indices, td->data, (uint32_t)td->n_obs,
                            (uint32_t)td->n_samples);
    PhyloTree tree = PhyloTree::from_arrays(tr->structure, tr->lengths, tr->names, tr->n_parens);
    ComputeStatus st = validate(t, tree);
    if (st != okay) return st;
    return faith_common(t, tree, result);
}

// =================================================================== partial

// src/api.cpp:527-567: stripes [stripe_start, stripe_stop) of the full matrix.
// Only the pairs of those stripes are computed (the Gram tiles, pair blocks and
// posting-list columns holding one).  n_substeps splits the reference's work
// into sequential tasks without changing any value (src/api.cpp:447-491); the
// range is computed here in one pass.
ComputeStatus partial_v3(const char* biom_filename, const char* tree_filename, const char* unifrac_method,
                         bool variance_adjust, double alpha, bool bypass_tips, bool normalize_sample_count,
                         unsigned int, unsigned int stripe_start, unsigned int stripe_stop, partial_mat_t** result) {
    if (!biom_filename || !file_exists(biom_filename)) return table_missing;
    if (!tree_filename || !file_exists(tree_filename)) return tree_missing;
    Method m;
    bool fp64;
    if (!parse_method(unifrac_method, m, fp64)) return unknown_method;
    PhyloTree tree = PhyloTree::from_newick(read_file(tree_filename).c_str());
    Table t;
    ComputeStatus st = load_table(biom_filename, t);
    if (st != okay) return st;
    st = validate(t, tree);
    if (st != okay) return st;
    const uint32_t n = t.n_samples;
    const uint32_t total = (n + 1) / 2;
    if (stripe_stop > total) {
        set_error("stopping stripe is out-of-bounds, max " + std::to_string(total));
        return output_error;
    }
    if (stripe_stop <= stripe_start) stripe_stop = total;  // src/api.cpp:456-457 (set_tasks)
    std::string err;
    ssu::PostorderNodes nodes = synced_nodes(t, tree);
    std::unique_ptr<Plan> plan =
        Plan::create(t, nodes,
                     ssu::spec_for_table(t, make_spec(m, fp64, variance_adjust, alpha, bypass_tips, normalize_sample_count)),
                     &err);
    if (!plan || !plan->set_stripes(stripe_start, stripe_stop, &err) || !plan->execute(&err)) {
        set_error(err);
        return output_error;
    }
    partial_mat_t* r = (partial_mat_t*)malloc(sizeof(partial_mat_t));
    r->n_samples = n;
    r->sample_ids = dup_ids(t.sample_ids);
    r->stripe_start = stripe_start;
    r->stripe_stop = stripe_stop;
    r->stripe_total = total;
    r->is_upper_triangle = true;
    const uint32_t rows = stripe_stop - stripe_start;
    r->stripes = (double**)malloc(sizeof(double*) * (rows ? rows : 1));
    for (uint32_t i = 0; i < rows; ++i) r->stripes[i] = (double*)malloc(sizeof(double) * (n ? n : 1));
    if (!plan->copy_stripes(stripe_start, stripe_stop, r->stripes, &err)) {
        set_error(err);
        destroy_partial_mat(&r);
        return output_error;
    }
    *result = r;
    return okay;
}

ComputeStatus partial(const char* biom_filename, const char* tree_filename, const char* unifrac_method,
                      bool variance_adjust, double alpha, bool bypass_tips, unsigned int n_substeps,
                      unsigned int stripe_start, unsigned int stripe_stop, partial_mat_t** result) {
    return partial_v3(biom_filename, tree_filename, unifrac_method, variance_adjust, alpha, bypass_tips, true,
                      n_substeps, stripe_start, stripe_stop, result);
}

// ============================================================= text writers

// src/api.cpp:1711-1757 (write_mat): tab-separated square matrix with a header row
IOStatus write_mat(const char* output_filename, mat_t* result) {
    FILE* f = fopen(output_filename, "w");
    if (!f) return open_error;
    const uint64_t n = result->n_samples;
    const uint64_t comb_N = n * (n - 1) / 2;
    for (uint64_t i = 0; i < n; ++i) fprintf(f, "\t%s", result->sample_ids[i]);
    fprintf(f, "\n");
    for (uint64_t i = 0; i < n; ++i) {
        fprintf(f, "%s", result->sample_ids[i]);
        for (uint64_t j = 0; j < n; ++j) {
            double v = 0.0;
            if (i < j) { const uint64_t ni = n - i; v = result->condensed_form[comb_N - ni * (ni - 1) / 2 + (j - i - 1)]; }
            else if (i > j) { const uint64_t nj = n - j; v = result->condensed_form[comb_N - nj * (nj - 1) / 2 + (i - j - 1)]; }
            fprintf(f, "\t%.16g", v);
        }
        fprintf(f, "\n");
    }
    fclose(f);
    return write_okay;
}

// src/api.cpp:1640-1775 (write_mat_from_matrix_txt_T): header of ids, then one row per sample,
// values in std::to_chars fixed (shortest round-trip) form
template <class TMat>
IOStatus write_full_txt(const char* output_filename, TMat* result) {
    FILE* f = fopen(output_filename, "w");
    if (!f) return write_error;
    const uint64_t n = result->n_samples;
    std::string line;
    for (uint64_t i = 0; i < n; ++i) { line += '\t'; line += result->sample_ids[i]; }
    line += '\n';
    IOStatus rc = write_okay;
    if (fwrite(line.data(), line.size(), 1, f) != 1) rc = write_error;
    char buf[400];
    for (uint64_t i = 0; i < n && rc == write_okay; ++i) {
        line.assign(result->sample_ids[i]);
        for (uint64_t j = 0; j < n; ++j) {
            line += '\t';
            auto r = std::to_chars(buf, buf + sizeof(buf), result->matrix[i * n + j], std::chars_format::fixed);
            line.append(buf, r.ptr);
        }
        line += '\n';
        if (fwrite(line.data(), line.size(), 1, f) != 1) rc = write_error;
    }
    fclose(f);
    return rc;
}

IOStatus write_mat_from_matrix(const char* output_filename, mat_full_fp64_t* result) {
    return write_full_txt(output_filename, result);
}

IOStatus write_mat_from_matrix_fp32(const char* output_filename, mat_full_fp32_t* result) {
    return write_full_txt(output_filename, result);
}


namespace {

// src/api.cpp:243-271 (is_fp64_method, is_fp64)
ComputeStatus method_precision(const char* method, bool& fp64) {
    Method m;
    return parse_method(method, m, fp64) ? okay : unknown_method;
}

ComputeStatus format_precision(const char* method, const char* format, bool& fp64, bool& save_dist) {
    const std::string f = format ? format : "";
    if (f == "hdf5_fp32") { fp64 = false; save_dist = true; return okay; }
    if (f == "hdf5_fp64") { fp64 = true; save_dist = true; return okay; }
    if (f == "hdf5") { save_dist = true; return method_precision(method, fp64); }
    if (f == "hdf5_nodist") { save_dist = false; return method_precision(method, fp64); }
    return unknown_method;
}

// PERMANOVA needs a grouping file and its columns (src/su.cpp:495-503)
bool missing_grouping(unsigned int permanova_perms, const char* grouping_filename, const char* grouping_columns) {
    if (permanova_perms > 0 && (!grouping_filename || !grouping_columns || !grouping_columns[0])) {
        set_error("PERMANOVA needs a grouping file and grouping columns");
        return true;
    }
    return false;
}

// PCoA of a full matrix (src/skbio_alt.cpp:170-214) into arrays of the
// result's element type; the seed comes from the ssu_set_random_seed stream.
template <class TReal>
bool pcoa_of(const TReal* matrix, uint32_t n, unsigned int dims, std::vector<TReal>& eig, std::vector<TReal>& samples,
             std::vector<TReal>& prop, std::string* err) {
    std::vector<double> e(dims), s((size_t)n * dims), p(dims);
    if (!ssu::pcoa(matrix, sizeof(TReal) == sizeof(double), n, dims, g_rng(), e.data(), s.data(), p.data(), err))
        return false;
    eig.assign(e.begin(), e.end());
    samples.assign(s.begin(), s.end());
    prop.assign(p.begin(), p.end());
    return true;
}

// the three PCoA datasets, in the order of src/api.cpp:1264-1290 (append_hdf5_pcoa)
template <class TReal>
bool write_pcoa(ssu::H5Writer& w, const TReal* matrix, uint32_t n, unsigned int dims, const std::string& suffix,
                std::string* err) {
    std::vector<TReal> eig, samples, prop;
    if (!pcoa_of(matrix, n, dims, eig, samples, prop, err)) return false;
    const bool fp64 = sizeof(TReal) == sizeof(double);
    return w.reals(("pcoa_eigvals" + suffix).c_str(), dims, eig.data(), fp64) &&
           w.reals(("pcoa_proportion_explained" + suffix).c_str(), dims, prop.data(), fp64) &&
           w.matrix(("pcoa_samples" + suffix).c_str(), n, dims, samples.data(), fp64);
}

// src/api.cpp:1787-1935, including PCoA (:1834-1860) and the statistics
// datasets (:1880-1916)
template <class TMat, class TReal>
IOStatus write_full_hdf5_stats(const char* filename, TMat* result, unsigned int pcoa_dims, int save_dist, bool fp64,
                               unsigned int n_vals, const char* const* methods, const char* const* names,
                               const TReal* vals, const TReal* pvals, const unsigned int* perms,
                               const char* const* group_names, const unsigned int* group_counts) {
    if (!filename || !result) return write_error;
    std::string err;
    auto w = ssu::H5Writer::create(filename, &err);
    if (!w) { set_error(err); return write_error; }
    const uint32_t n = result->n_samples;
    bool ok = w->string("format", "BDSM") && w->string("version", "2020.12") &&
              w->strings("order", n, result->sample_ids);
    if (ok && save_dist) ok = w->matrix("matrix", n, n, result->matrix, fp64);
    if (ok && pcoa_dims > 0) {
        ok = w->string("pcoa_method", "FSVD") && write_pcoa(*w, result->matrix, n, pcoa_dims, "", &err);
        if (!ok && !err.empty()) { set_error(err); return write_error; }
    }
    if (ok && n_vals > 0) {
        ok = methods && names && vals && pvals && perms && group_names && group_counts &&
             w->strings("stat_methods", n_vals, methods) && w->strings("stat_test_names", n_vals, names) &&
             w->strings("stat_grouping_names", n_vals, group_names) &&
             w->u32s("stat_n_groups", n_vals, group_counts) && w->reals("stat_values", n_vals, vals, fp64) &&
             w->reals("stat_pvalues", n_vals, pvals, fp64) && w->u32s("stat_n_permutations", n_vals, perms);
    }
    if (!ok) { set_error(std::string("HDF5 write failed for ") + filename); return write_error; }
    return write_okay;
}

// src/api.cpp:951-976 (compute_permanova_T): each grouping column separately
template <class TReal, class TMat>
ComputeStatus permanova_columns(const char* grouping_filename, unsigned int n_columns, const char* const* columns,
                                const TMat* result, unsigned int permanova_perms, TReal* fstats, TReal* pvalues,
                                unsigned int* n_groups) {
    if (!grouping_filename || !result || (n_columns && (!columns || !fstats || !pvalues || !n_groups)))
        return grouping_missing;
    std::vector<std::string> ids(result->sample_ids, result->sample_ids + result->n_samples);
    for (unsigned int c = 0; c < n_columns; ++c) {
        std::vector<uint32_t> grouping;
        uint32_t groups = 0;
        std::string err;
        if (!ssu::read_grouping(grouping_filename, ids, columns[c], grouping, groups, &err)) {
            set_error(err);
            return grouping_missing;
        }
        double f = 0.0, p = 0.0;
        if (!ssu::permanova(result->matrix, sizeof(TReal) == sizeof(double), result->n_samples, grouping.data(),
                            permanova_perms, g_rng(), &f, &p, &err)) {
            set_error(err);
            return output_error;
        }
        fstats[c] = (TReal)f;
        pvalues[c] = (TReal)p;
        n_groups[c] = groups;
    }
    return okay;
}

// Copies matrix r of method k of a batch into host memory of element type TReal.
template <class TReal>
bool batch_matrix(const ssu::RarefyBatch& b, uint32_t r, TReal* out, std::string* err) {
    const size_t nn = (size_t)b.n_samples() * b.n_samples();
    if (b.out_fp64(0) == (sizeof(TReal) == sizeof(double))) return b.copy_matrix(0, r, out, err);
    if (b.out_fp64(0)) {
        std::vector<double> tmp(nn);
        if (!b.copy_matrix(0, r, tmp.data(), err)) return false;
        for (size_t i = 0; i < nn; ++i) out[i] = (TReal)tmp[i];
    } else {
        std::vector<float> tmp(nn);
        if (!b.copy_matrix(0, r, tmp.data(), err)) return false;
        for (size_t i = 0; i < nn; ++i) out[i] = (TReal)tmp[i];
    }
    return true;
}

constexpr size_t MULTI_DEVICE_BUDGET = size_t(32) << 30;   // bytes of replicate matrices held on the device at once

// src/api.cpp:1444-1563 (unifrac_multi_to_file_T): n_subsamples rarefied
// tables, one matrix each, into an HDF5 file of format "SUJK" version
// "2023.03" with datasets "order" and "matrix:<i>".  Without bypass_tips all
// replicates of a chunk run from one device batch (one HIP graph); with it,
// each replicate's table is sheared to its own observations first, as the
// reference does, and computed on its own.
template <class TReal>
ComputeStatus multi_to_file(bool save_dist, const char* biom_filename, const char* tree_filename,
                            const char* out_filename, const char* method_name, bool vaw, double alpha, bool bypass,
                            bool normalize, unsigned int n_subsamples, unsigned int depth, bool with_replacement,
                            unsigned int pcoa_dims, unsigned int permanova_perms, const char* grouping_filename,
                            const char* grouping_columns) {
    if (depth < 1) {
        fprintf(stderr, "ERROR: subsampling depth cannot be 0.\n");
        return table_empty;
    }
    if (!biom_filename || !file_exists(biom_filename)) return table_missing;
    if (!tree_filename || !file_exists(tree_filename)) return tree_missing;
    if (missing_grouping(permanova_perms, grouping_filename, grouping_columns)) return grouping_missing;
    const std::vector<std::string> columns =
        permanova_perms > 0 ? ssu::split_commas(grouping_columns) : std::vector<std::string>();
    Method m;
    bool m64;
    if (!parse_method(method_name, m, m64)) return unknown_method;
    const PhyloTree tree = PhyloTree::from_newick(read_file(tree_filename).c_str());
    Table t;
    ComputeStatus st = load_table(biom_filename, t);
    if (st != okay) return st;
    st = validate(t, tree);
    if (st != okay) return st;
    const ComputeSpec spec = make_spec(m, m64, vaw, alpha, bypass, normalize);
    const bool fp64 = sizeof(TReal) == sizeof(double);
    std::string err;
    auto w = ssu::H5Writer::create(out_filename, &err);
    if (!w || !w->string("format", "SUJK") || !w->string("version", "2023.03") ||
        (pcoa_dims > 0 && !w->string("pcoa_method", "FSVD"))) {
        set_error(err.empty() ? std::string("HDF5 write failed for ") + out_filename : err);
        return output_error;
    }
    // statistics of every replicate x column (src/api.cpp:1509-1559)
    std::vector<TReal> st_f, st_p;
    std::vector<unsigned int> st_groups;
    std::vector<std::string> st_names;
    uint32_t n_kept = 0;
    for (uint32_t n = 0; n < t.n_samples; ++n) n_kept += t.totals[n] >= depth;
    const size_t per_rep = std::max<size_t>((size_t)n_kept * n_kept * (m64 ? 8 : 4), 1);
    const uint32_t chunk = (uint32_t)std::max<size_t>(1, std::min<size_t>(n_subsamples, MULTI_DEVICE_BUDGET / per_rep));
    std::vector<TReal> mat;
    uint32_t done = 0;
    bool order_written = false;
    ComputeStatus emit_status = okay;
    auto emit = [&](uint32_t n, const std::vector<std::string>& ids) -> bool {
        std::vector<const char*> p;
        for (const auto& x : ids) p.push_back(x.c_str());
        if (!order_written) {
            if (!w->strings("order", n, p.data())) return false;
            order_written = true;
        }
        const std::string tag = ":" + std::to_string(done);
        if (permanova_perms > 0) {   // before PCoA, which does not touch the matrix here either
            struct { uint32_t n_samples; uint32_t flags; TReal* matrix; char** sample_ids; } view{
                n, 0, mat.data(), const_cast<char**>(p.data())};
            std::vector<const char*> cols;
            for (const auto& c : columns) cols.push_back(c.c_str());
            std::vector<TReal> f(columns.size()), pv(columns.size());
            std::vector<unsigned int> g(columns.size());
            emit_status = permanova_columns<TReal>(grouping_filename, (unsigned int)cols.size(), cols.data(), &view,
                                                   permanova_perms, f.data(), pv.data(), g.data());
            if (emit_status != okay) return false;
            for (size_t c = 0; c < columns.size(); ++c) {
                st_f.push_back(f[c]);
                st_p.push_back(pv[c]);
                st_groups.push_back(g[c]);
                st_names.push_back(columns[c] + tag);
            }
        }
        if (save_dist && !w->matrix(("matrix" + tag).c_str(), n, n, mat.data(), fp64)) return false;
        if (pcoa_dims > 0 && !write_pcoa(*w, mat.data(), n, pcoa_dims, tag, &err)) return false;
        return true;
    };
    while (done < n_subsamples) {
        const uint32_t reps = std::min(chunk, n_subsamples - done);
        if (!bypass) {
            auto b = ssu::RarefyBatch::create(t, tree, {spec}, depth, with_replacement, g_rng(), reps, &err);
            if (!b) {
                if (err.find("reaches the rarefaction depth") != std::string::npos) return table_empty;
                set_error(err);
                return output_error;
            }
            if (!b->execute(&err)) { set_error(err); return output_error; }
            const uint32_t n = b->n_samples();
            mat.resize((size_t)n * n);
            for (uint32_t r = 0; r < reps; ++r, ++done) {
                if (!batch_matrix<TReal>(*b, r, mat.data(), &err)) { set_error(err); return output_error; }
                if (!emit(n, b->kept().sample_ids)) {
                    if (emit_status != okay) return emit_status;
                    set_error(err.empty() ? "HDF5 write failed" : err);
                    return output_error;
                }
            }
        } else {
            std::vector<Table> subs;
            st = rarefied_tables(t, tree, depth, with_replacement, reps, subs);
            if (st != okay) return st;
            for (const Table& sub : subs) {
                mat.resize((size_t)sub.n_samples * sub.n_samples);
                st = run_full<TReal>(sub, tree, spec, mat.data());
                if (st != okay) return st;
                if (!emit(sub.n_samples, sub.sample_ids)) {
                    if (emit_status != okay) return emit_status;
                    set_error(err.empty() ? "HDF5 write failed" : err);
                    return output_error;
                }
                ++done;
            }
        }
    }
    if (permanova_perms > 0 && !st_f.empty()) {   // src/api.cpp:1390-1424 (write_stats)
        const size_t k = st_f.size();
        std::vector<const char*> meth(k, "PERMANOVA"), test(k, "pseudo-F"), names;
        for (const auto& x : st_names) names.push_back(x.c_str());
        std::vector<unsigned int> perms(k, permanova_perms);
        const bool ok = w->strings("stat_methods", k, meth.data()) && w->strings("stat_test_names", k, test.data()) &&
                        w->strings("stat_grouping_names", k, names.data()) &&
                        w->u32s("stat_n_groups", k, st_groups.data()) && w->reals("stat_values", k, st_f.data(), fp64) &&
                        w->reals("stat_pvalues", k, st_p.data(), fp64) &&
                        w->u32s("stat_n_permutations", k, perms.data());
        if (!ok) { set_error("HDF5 write failed"); return output_error; }
    }
    return okay;
}

}  // namespace

IOStatus write_mat_from_matrix_hdf5_fp64(const char* filename, mat_full_fp64_t* result, unsigned int pcoa_dims,
                                         int save_dist) {
    return write_full_hdf5_stats<mat_full_fp64_t, double>(filename, result, pcoa_dims, save_dist, true, 0, nullptr,
                                                          nullptr, nullptr, nullptr, nullptr, nullptr, nullptr);
}

IOStatus write_mat_from_matrix_hdf5_fp32(const char* filename, mat_full_fp32_t* result, unsigned int pcoa_dims,
                                         int save_dist) {
    return write_full_hdf5_stats<mat_full_fp32_t, float>(filename, result, pcoa_dims, save_dist, false, 0, nullptr,
                                                         nullptr, nullptr, nullptr, nullptr, nullptr, nullptr);
}

ComputeStatus compute_permanova_fp64(const char* grouping_filename, unsigned int n_columns, const char** columns,
                                     mat_full_fp64_t* result, unsigned int permanova_perms, double* fstats,
                                     double* pvalues, unsigned int* n_groups) {
    return permanova_columns<double>(grouping_filename, n_columns, columns, result, permanova_perms, fstats, pvalues,
                                     n_groups);
}

ComputeStatus compute_permanova_fp32(const char* grouping_filename, unsigned int n_columns, const char** columns,
                                     mat_full_fp32_t* result, unsigned int permanova_perms, float* fstats,
                                     float* pvalues, unsigned int* n_groups) {
    return permanova_columns<float>(grouping_filename, n_columns, columns, result, permanova_perms, fstats, pvalues,
                                    n_groups);
}

ComputeStatus ssu_center_matrix_fp64(const double* mat, unsigned int n, double* centered) {
    if (!mat || !centered) return output_error;
    std::string err;
    if (!ssu::center_matrix(mat, true, n, centered, &err)) { set_error(err); return output_error; }
    return okay;
}

ComputeStatus ssu_center_matrix_fp64_noalloc(const double* mat, unsigned int n, double* centered, void* stream) {
    if (!mat || !centered) return output_error;
    std::string err;
    if (!ssu::center_matrix_noalloc(mat, n, centered, stream, &err)) { set_error(err); return output_error; }
    return okay;
}

ComputeStatus ssu_center_matrix_fp32(const float* mat, unsigned int n, float* centered) {
    if (!mat || !centered) return output_error;
    std::string err;
    std::vector<double> tmp((size_t)n * n);
    if (!ssu::center_matrix(mat, false, n, tmp.data(), &err)) { set_error(err); return output_error; }
    for (size_t q = 0; q < tmp.size(); ++q) centered[q] = (float)tmp[q];
    return okay;
}

template <class T>
ComputeStatus pcoa_api(const T* mat, unsigned int n, unsigned int dims, T* eigenvalues, T* samples, T* proportion) {
    if (!mat || !eigenvalues || !samples || !proportion) return output_error;
    std::vector<T> e, s, p;
    std::string err;
    if (!pcoa_of(mat, n, dims, e, s, p, &err)) { set_error(err); return output_error; }
    std::copy(e.begin(), e.end(), eigenvalues);
    std::copy(s.begin(), s.end(), samples);
    std::copy(p.begin(), p.end(), proportion);
    return okay;
}

ComputeStatus ssu_pcoa_fp64(const double* mat, unsigned int n, unsigned int dims, double* eigenvalues,
                            double* samples, double* proportion_explained) {
    return pcoa_api(mat, n, dims, eigenvalues, samples, proportion_explained);
}

ComputeStatus ssu_pcoa_fp32(const float* mat, unsigned int n, unsigned int dims, float* eigenvalues, float* samples,
                            float* proportion_explained) {
    return pcoa_api(mat, n, dims, eigenvalues, samples, proportion_explained);
}

template <class T>
ComputeStatus permanova_api(const T* mat, unsigned int n, const uint32_t* grouping, unsigned int n_perm, T* fstat,
                            T* pvalue) {
    if (!mat || !grouping || !fstat || !pvalue) return output_error;
    double f = 0.0, p = 0.0;
    std::string err;
    if (!ssu::permanova(mat, sizeof(T) == sizeof(double), n, grouping, n_perm, g_rng(), &f, &p, &err)) {
        set_error(err);
        return output_error;
    }
    *fstat = (T)f;
    *pvalue = (T)p;
    return okay;
}

ComputeStatus ssu_permanova_fp64(const double* mat, unsigned int n, const uint32_t* grouping, unsigned int n_perm,
                                 double* fstat, double* pvalue) {
    return permanova_api(mat, n, grouping, n_perm, fstat, pvalue);
}

ComputeStatus ssu_permanova_fp32(const float* mat, unsigned int n, const uint32_t* grouping, unsigned int n_perm,
                                 float* fstat, float* pvalue) {
    return permanova_api(mat, n, grouping, n_perm, fstat, pvalue);
}

IOStatus write_mat_from_matrix_hdf5_fp64_v2(const char* output_filename, mat_full_fp64_t* result,
                                            unsigned int pcoa_dims, int save_dist, unsigned int stat_n_vals,
                                            const char** stat_method_arr, const char** stat_name_arr,
                                            const double* stat_val_arr, const double* stat_pval_arr,
                                            const unsigned int* stat_perm_count_arr, const char** stat_group_name_arr,
                                            const unsigned int* stat_group_count_arr) {
    return write_full_hdf5_stats(output_filename, result, pcoa_dims, save_dist, true, stat_n_vals, stat_method_arr,
                                 stat_name_arr, stat_val_arr, stat_pval_arr, stat_perm_count_arr, stat_group_name_arr,
                                 stat_group_count_arr);
}

IOStatus write_mat_from_matrix_hdf5_fp32_v2(const char* output_filename, mat_full_fp32_t* result,
                                            unsigned int pcoa_dims, int save_dist, unsigned int stat_n_vals,
                                            const char** stat_method_arr, const char** stat_name_arr,
                                            const float* stat_val_arr, const float* stat_pval_arr,
                                            const unsigned int* stat_perm_count_arr, const char** stat_group_name_arr,
                                            const unsigned int* stat_group_count_arr) {
    return write_full_hdf5_stats(output_filename, result, pcoa_dims, save_dist, false, stat_n_vals, stat_method_arr,
                                 stat_name_arr, stat_val_arr, stat_pval_arr, stat_perm_count_arr, stat_group_name_arr,
                                 stat_group_count_arr);
}

// src/api.cpp:989-1117
template <class TReal, class TMat>
ComputeStatus to_file_T(TMat* result, const char* out_filename, bool save_dist, unsigned int pcoa_dims,
                        unsigned int permanova_perms, const char* grouping_filename, const char* grouping_columns) {
    const bool fp64 = sizeof(TReal) == sizeof(double);
    const std::vector<std::string> columns = ssu::split_commas(grouping_columns);
    const unsigned int k = permanova_perms > 0 ? (unsigned int)columns.size() : 0;
    std::vector<const char*> cols;
    for (const auto& c : columns) cols.push_back(c.c_str());
    std::vector<TReal> f(k), p(k);
    std::vector<unsigned int> g(k);
    if (k) {
        const ComputeStatus st = permanova_columns<TReal>(grouping_filename, k, cols.data(), result, permanova_perms,
                                                          f.data(), p.data(), g.data());
        if (st != okay) return st;
    }
    std::vector<const char*> meth(k, "PERMANOVA"), test(k, "pseudo-F");
    std::vector<unsigned int> perms(k, permanova_perms);
    const IOStatus io = write_full_hdf5_stats<TMat, TReal>(out_filename, result, pcoa_dims, save_dist, fp64, k,
                                                           meth.data(), test.data(), f.data(), p.data(), perms.data(),
                                                           cols.data(), g.data());
    return io == write_okay ? okay : output_error;
}

ComputeStatus unifrac_to_file_v3(const char* biom_filename, const char* tree_filename, const char* out_filename,
                                 const char* unifrac_method, bool variance_adjust, double alpha, bool bypass_tips,
                                 bool normalize_sample_counts, unsigned int n_substeps, const char* format,
                                 unsigned int subsample_depth, bool subsample_with_replacement, unsigned int pcoa_dims,
                                 unsigned int permanova_perms, const char* grouping_filename,
                                 const char* grouping_columns, const char* mmap_dir) {
    bool fp64 = false, save_dist = true;
    ComputeStatus rc = format_precision(unifrac_method, format, fp64, save_dist);
    if (rc != okay) return rc;
    if (missing_grouping(permanova_perms, grouping_filename, grouping_columns)) return grouping_missing;
    if (fp64) {
        mat_full_fp64_t* result = nullptr;
        rc = one_off_matrix_v3(biom_filename, tree_filename, unifrac_method, variance_adjust, alpha, bypass_tips,
                               normalize_sample_counts, n_substeps, subsample_depth, subsample_with_replacement,
                               mmap_dir, &result);
        if (rc == okay)
            rc = to_file_T<double>(result, out_filename, save_dist, pcoa_dims, permanova_perms, grouping_filename,
                                   grouping_columns);
        if (result) destroy_mat_full_fp64(&result);
    } else {
        mat_full_fp32_t* result = nullptr;
        rc = one_off_matrix_fp32_v3(biom_filename, tree_filename, unifrac_method, variance_adjust, alpha, bypass_tips,
                                    normalize_sample_counts, n_substeps, subsample_depth, subsample_with_replacement,
                                    mmap_dir, &result);
        if (rc == okay)
            rc = to_file_T<float>(result, out_filename, save_dist, pcoa_dims, permanova_perms, grouping_filename,
                                  grouping_columns);
        if (result) destroy_mat_full_fp32(&result);
    }
    return rc;
}

ComputeStatus unifrac_to_file_v2(const char* biom_filename, const char* tree_filename, const char* out_filename,
                                 const char* unifrac_method, bool variance_adjust, double alpha, bool bypass_tips,
                                 unsigned int n_substeps, const char* format, unsigned int subsample_depth,
                                 bool subsample_with_replacement, unsigned int pcoa_dims, unsigned int permanova_perms,
                                 const char* grouping_filename, const char* grouping_columns, const char* mmap_dir) {
    return unifrac_to_file_v3(biom_filename, tree_filename, out_filename, unifrac_method, variance_adjust, alpha,
                              bypass_tips, true, n_substeps, format, subsample_depth, subsample_with_replacement,
                              pcoa_dims, permanova_perms, grouping_filename, grouping_columns, mmap_dir);
}

ComputeStatus unifrac_to_file(const char* biom_filename, const char* tree_filename, const char* out_filename,
                              const char* unifrac_method, bool variance_adjust, double alpha, bool bypass_tips,
                              unsigned int n_substeps, const char* format, unsigned int pcoa_dims,
                              const char* mmap_dir) {
    return unifrac_to_file_v2(biom_filename, tree_filename, out_filename, unifrac_method, variance_adjust, alpha,
                              bypass_tips, n_substeps, format, 0, true, pcoa_dims, 0, nullptr, nullptr, mmap_dir);
}

// src/api.cpp:1119-1165
ComputeStatus unifrac_to_txt_file_v3(const char* biom_filename, const char* tree_filename, const char* out_filename,
                                     const char* unifrac_method, bool variance_adjust, double alpha, bool bypass_tips,
                                     bool normalize_sample_counts, unsigned int n_substeps, const char* mmap_dir) {
    bool fp64 = false;
    ComputeStatus rc = method_precision(unifrac_method, fp64);
    if (rc != okay) return rc;
    if (fp64) {
        mat_full_fp64_t* result = nullptr;
        rc = one_off_matrix_v3(biom_filename, tree_filename, unifrac_method, variance_adjust, alpha, bypass_tips,
                               normalize_sample_counts, n_substeps, 0, true, mmap_dir, &result);
        if (rc == okay && write_mat_from_matrix(out_filename, result) != write_okay) rc = output_error;
        if (result) destroy_mat_full_fp64(&result);
    } else {
        mat_full_fp32_t* result = nullptr;
        rc = one_off_matrix_fp32_v3(biom_filename, tree_filename, unifrac_method, variance_adjust, alpha, bypass_tips,
                                    normalize_sample_counts, n_substeps, 0, true, mmap_dir, &result);
        if (rc == okay && write_mat_from_matrix_fp32(out_filename, result) != write_okay) rc = output_error;
        if (result) destroy_mat_full_fp32(&result);
    }
    return rc;
}

// src/api.cpp:1565-1598
ComputeStatus unifrac_multi_to_file_v3(const char* biom_filename, const char* tree_filename, const char* out_filename,
                                       const char* unifrac_method, bool variance_adjust, double alpha,
                                       bool bypass_tips, bool normalize_sample_counts, unsigned int,
                                       const char* format, unsigned int n_subsamples, unsigned int subsample_depth,
                                       bool subsample_with_replacement, unsigned int pcoa_dims,
                                       unsigned int permanova_perms, const char* grouping_filename,
                                       const char* grouping_columns, const char*) {
    bool fp64 = false, save_dist = true;
    const ComputeStatus rc = format_precision(unifrac_method, format, fp64, save_dist);
    if (rc != okay) return rc;
    if (fp64)
        return multi_to_file<double>(save_dist, biom_filename, tree_filename, out_filename, unifrac_method,
                                     variance_adjust, alpha, bypass_tips, normalize_sample_counts, n_subsamples,
                                     subsample_depth, subsample_with_replacement, pcoa_dims, permanova_perms,
                                     grouping_filename, grouping_columns);
    return multi_to_file<float>(save_dist, biom_filename, tree_filename, out_filename, unifrac_method, variance_adjust,
                                alpha, bypass_tips, normalize_sample_counts, n_subsamples, subsample_depth,
                                subsample_with_replacement, pcoa_dims, permanova_perms, grouping_filename,
                                grouping_columns);
}

ComputeStatus unifrac_multi_to_file_v2(const char* biom_filename, const char* tree_filename, const char* out_filename,
                                       const char* unifrac_method, bool variance_adjust, double alpha,
                                       bool bypass_tips, unsigned int n_substeps, const char* format,
                                       unsigned int n_subsamples, unsigned int subsample_depth,
                                       bool subsample_with_replacement, unsigned int pcoa_dims,
                                       unsigned int permanova_perms, const char* grouping_filename,
                                       const char* grouping_columns, const char* mmap_dir) {
    return unifrac_multi_to_file_v3(biom_filename, tree_filename, out_filename, unifrac_method, variance_adjust,
                                    alpha, bypass_tips, true, n_substeps, format, n_subsamples, subsample_depth,
                                    subsample_with_replacement, pcoa_dims, permanova_perms, grouping_filename,
                                    grouping_columns, mmap_dir);
}

struct ssu_rarefy {
    Table table;
    std::unique_ptr<ssu::RarefyBatch> batch;
};

ComputeStatus ssu_rarefy_create_inmem(const support_biom_t* td, const support_bptree_t* tr,
                                      const char* const* methods, int n_methods, bool variance_adjust, double alpha,
                                      bool normalize_sample_counts, unsigned int depth, bool with_replacement,
                                      uint64_t seed, unsigned int n_reps, ssu_rarefy_t** out) {
    if (!out) return output_error;
    *out = nullptr;
    if (!methods || n_methods <= 0) return unknown_method;
    std::vector<ComputeSpec> specs;
    for (int k = 0; k < n_methods; ++k) {
        Method m;
        bool fp64;
        if (!parse_method(methods[k], m, fp64)) return unknown_method;
        specs.push_back(make_spec(m, fp64, variance_adjust, alpha, false, normalize_sample_counts));
    }
    if (!td || !tr || td->n_samples <= 0 || td->n_obs <= 0) return table_empty;
    auto r = std::make_unique<ssu_rarefy>();
    r->table = Table::borrow(td->obs_ids, td->sample_ids, td->indptr, td->indices, td->data, (uint32_t)td->n_obs,
                             (uint32_t)td->n_samples);
    PhyloTree tree = PhyloTree::from_arrays(tr->structure, tr->lengths, tr->names, tr->n_parens);
    const ComputeStatus v = validate(r->table, tree);
    if (v != okay) return v;
    std::string err;
    r->batch = ssu::RarefyBatch::create(r->table, tree, specs, depth, with_replacement, seed, n_reps, &err);
    if (!r->batch) {
        set_error(err);
        return err.find("reaches the rarefaction depth") != std::string::npos ? table_empty : output_error;
    }
    *out = r.release();
    return okay;
}

ComputeStatus ssu_rarefy_execute(ssu_rarefy_t* r) {
    if (!r || !r->batch) return output_error;
    std::string err;
    if (!r->batch->execute(&err)) { set_error(err); return output_error; }
    return okay;
}

unsigned int ssu_rarefy_n_samples(const ssu_rarefy_t* r) { return r && r->batch ? r->batch->n_samples() : 0; }
unsigned int ssu_rarefy_n_obs(const ssu_rarefy_t* r) { return r && r->batch ? r->batch->n_obs() : 0; }
uint64_t ssu_rarefy_nnz(const ssu_rarefy_t* r) { return r && r->batch ? r->batch->nnz() : 0; }

void ssu_rarefy_kept(const ssu_rarefy_t* r, uint32_t* sample_index, uint32_t* obs_index, uint32_t* indptr,
                     uint32_t* indices) {
    if (!r || !r->batch) return;
    const auto& b = *r->batch;
    const Table& t = b.kept();
    if (sample_index) std::copy(b.sample_index().begin(), b.sample_index().end(), sample_index);
    if (obs_index) std::copy(b.obs_index().begin(), b.obs_index().end(), obs_index);
    const uint32_t base = t.n_obs ? t.indptr[0] : 0;
    if (indptr)
        for (uint32_t i = 0; i <= t.n_obs; ++i) indptr[i] = t.indptr[i] - base;
    if (indices) std::copy(t.indices + base, t.indices + base + t.nnz(), indices);
}

ComputeStatus ssu_rarefy_copy_counts(const ssu_rarefy_t* r, unsigned int rep, double* data) {
    if (!r || !r->batch || !data) return output_error;
    std::string err;
    if (!r->batch->copy_counts(rep, data, &err)) { set_error(err); return output_error; }
    return okay;
}

ComputeStatus ssu_rarefy_copy_matrix(const ssu_rarefy_t* r, int method, unsigned int rep, void* host_out) {
    if (!r || !r->batch || !host_out || method < 0) return output_error;
    std::string err;
    if (!r->batch->copy_matrix((size_t)method, rep, host_out, &err)) { set_error(err); return output_error; }
    return okay;
}

double ssu_rarefy_last_ms(const ssu_rarefy_t* r) { return r && r->batch ? r->batch->last_ms() : 0.0; }
int ssu_rarefy_graph_used(const ssu_rarefy_t* r) { return r && r->batch && r->batch->graph_used() ? 1 : 0; }
uint64_t ssu_rarefy_pair_nodes(const ssu_rarefy_t* r) { return r && r->batch ? r->batch->pair_nodes() : 0; }

uint64_t ssu_rarefy_refined_pairs(const ssu_rarefy_t* r) { return r && r->batch ? r->batch->refined_pairs() : 0; }

void ssu_rarefy_destroy(ssu_rarefy_t** r) {
    if (!r || !*r) return;
    delete *r;
    *r = nullptr;
}

struct ssu_rarefy_group {
    std::unique_ptr<ssu::RarefyGroup> group;
};

ComputeStatus ssu_rarefy_group_create(ssu_rarefy_t* const* batches, int n_batches, ssu_rarefy_group_t** out) {
    if (!out || !batches || n_batches <= 0) return output_error;
    *out = nullptr;
    std::vector<ssu::RarefyBatch*> v;
    for (int i = 0; i < n_batches; ++i) {
        if (!batches[i] || !batches[i]->batch) { set_error("null rarefaction batch"); return output_error; }
        v.push_back(batches[i]->batch.get());
    }
    std::string err;
    auto g = ssu::RarefyGroup::create(v, &err);
    if (!g) { set_error(err); return output_error; }
    *out = new ssu_rarefy_group{std::move(g)};
    return okay;
}

ComputeStatus ssu_rarefy_group_execute(ssu_rarefy_group_t* g) {
    if (!g || !g->group) return output_error;
    std::string err;
    if (!g->group->execute(&err)) { set_error(err); return output_error; }
    return okay;
}

double ssu_rarefy_group_last_ms(const ssu_rarefy_group_t* g) { return g && g->group ? g->group->last_ms() : 0.0; }

void ssu_rarefy_group_destroy(ssu_rarefy_group_t** g) {
    if (!g || !*g) return;
    delete *g;
    *g = nullptr;
}

long ssu_biom_n_samples(const char* biom_filename) {
    uint64_t n = 0;
    std::string err;
    if (!biom_filename || !ssu::biom_n_samples(biom_filename, n, &err)) {
        set_error(err);
        return -1;
    }
    return (long)n;
}

// src/api.cpp:1759-1775 (write_vec): "#SampleID\tfaith_pd" then one line per sample
IOStatus write_vec(const char* output_filename, r_vec* result) {
    FILE* f = fopen(output_filename, "w");
    if (!f) return open_error;
    fprintf(f, "#SampleID\tfaith_pd\n");
    for (unsigned i = 0; i < result->n_samples; ++i) fprintf(f, "%s\t%.16g\n", result->sample_ids[i], result->values[i]);
    fclose(f);
    return write_okay;
}

// ================================================================ extensions

struct ssu_plan {
    Table table;
    std::unique_ptr<Plan> plan;
};

ComputeStatus ssu_plan_create_inmem(const support_biom_t* td, const support_bptree_t* tr, const char* method_name,
                                    bool variance_adjust, double alpha, bool bypass_tips, bool normalize_sample_counts,
                                    int plain_kernel, ssu_plan_t** out) {
    Method m;
    bool fp64;
    if (!parse_method(method_name, m, fp64)) return unknown_method;
    if (!td || !tr || td->n_samples <= 0 || td->n_obs <= 0) return table_empty;
    std::unique_ptr<ssu_plan> p(new ssu_plan);
    p->table = Table::borrow(td->obs_ids, td->sample_ids, td->indptr, td->indices, td->data, (uint32_t)td->n_obs,
                             (uint32_t)td->n_samples);
    PhyloTree tree = PhyloTree::from_arrays(tr->structure, tr->lengths, tr->names, tr->n_parens);
    ComputeStatus st = validate(p->table, tree);
    if (st != okay) return st;
    ComputeSpec spec = make_spec(m, fp64, variance_adjust, alpha, bypass_tips, normalize_sample_counts);
    spec.plain_kernel = plain_kernel;
    std::string err;
    ssu::PostorderNodes nodes = synced_nodes(p->table, tree);
    p->plan = Plan::create(p->table, nodes, ssu::spec_for_table(p->table, spec), &err);
    if (!p->plan) { set_error(err); return output_error; }
    *out = p.release();
    return okay;
}

ComputeStatus ssu_plan_execute(ssu_plan_t* plan) {
    if (!plan) return output_error;
    std::string err;
    if (!plan->plan->execute(&err)) { set_error(err); return output_error; }
    return okay;
}

int ssu_plan_result_bytes(const ssu_plan_t* plan) { return plan && plan->plan->out_fp64() ? 8 : 4; }

ComputeStatus ssu_plan_copy_matrix(const ssu_plan_t* plan, void* host_out) {
    std::string err;
    if (!plan || !plan->plan->copy_matrix(host_out, &err)) { set_error(err); return output_error; }
    return okay;
}

ComputeStatus ssu_plan_copy_rows(const ssu_plan_t* plan, unsigned int row0, unsigned int nrows, void* host_out) {
    std::string err;
    if (!plan || !plan->plan->copy_rows(row0, nrows, host_out, &err)) { set_error(err); return output_error; }
    return okay;
}

int ssu_plan_timings(const ssu_plan_t* plan, double* ms, int cap) {
    if (!plan) return 0;
    const float* t = plan->plan->timings();
    int k = 0;
    for (; k < ssu::PH_COUNT && k < cap; ++k) ms[k] = t[k];
    return k;
}

uint64_t ssu_plan_sparse_terms(const ssu_plan_t* plan) { return plan ? plan->plan->sparse_terms() : 0; }

uint64_t ssu_plan_refined_pairs(const ssu_plan_t* plan) { return plan ? plan->plan->refined_pairs() : 0; }

double ssu_plan_split_theta(const ssu_plan_t* plan) { return plan ? plan->plan->split_theta() : 0.0; }

ComputeStatus ssu_plan_execute_stripes(ssu_plan_t* plan, unsigned int start, unsigned int stop) {
    std::string err;
    if (!plan) return output_error;
    if (!plan->plan->set_stripes(start, stop, &err) || !plan->plan->execute(&err)) {
        set_error(err);
        return output_error;
    }
    return okay;
}

ComputeStatus ssu_plan_copy_stripes(const ssu_plan_t* plan, unsigned int start, unsigned int stop, double* host_out) {
    std::string err;
    if (!plan) return output_error;
    const uint32_t n = plan->plan->n_samples();
    if (stop > (n + 1) / 2 || stop < start) {
        set_error("ssu_plan_copy_stripes: stripes outside the matrix");
        return output_error;
    }
    std::vector<double*> rows(stop - start);
    for (uint32_t r = 0; r < stop - start; ++r) rows[r] = host_out + (size_t)r * n;
    if (!plan->plan->copy_stripes(start, stop, rows.data(), &err)) { set_error(err); return output_error; }
    return okay;
}

void ssu_release_device_memory(void) { ssu::pool_release_all(); }

unsigned int ssu_choose_split(int kind, unsigned int n_samples, const uint32_t* occupancy, unsigned int n_nodes,
                              double* model_ms, double* model_ms_none) {
    if (kind < 0 || kind > 6) return 0;
    std::vector<uint32_t> occ(occupancy, occupancy + n_nodes);
    // VAW (5, 6): one-sided terms at every sample's total (no two totals equal)
    const uint32_t cols = kind >= 5 ? n_samples : 0u;
    const ssu::SplitChoice c = ssu::choose_split((ssu::SplitKind)kind, n_samples, occ, 1ull << 30, cols);
    if (model_ms) *model_ms = c.model_ms;
    if (model_ms_none) *model_ms_none = c.model_ms_none;
    return c.thr;
}

int ssu_last_call_timings(double* ms, int cap) {
    const int k = std::min<int>(cap, CP_COUNT);
    for (int i = 0; i < k; ++i) ms[i] = g_call_ms[i];
    return k;
}

uint64_t ssu_plan_pair_nodes(const ssu_plan_t* plan) {
    if (!plan) return 0;
    const uint64_t n = plan->plan->n_samples();
    return n * (n - 1) / 2 * plan->plan->processed_nodes();
}

uint64_t ssu_plan_kernel_terms(const ssu_plan_t* plan) { return plan ? plan->plan->kernel_terms() : 0; }

const char* ssu_plan_main_kernel(const ssu_plan_t* plan) { return plan ? plan->plan->main_kernel() : ""; }

void ssu_plan_destroy(ssu_plan_t** plan) {
    if (plan) {
        delete *plan;
        *plan = nullptr;
    }
}

int ssu_postorder(const support_bptree_t* tr, int n_obs, const char* const* obs_ids, int32_t* parent, double* length,
                  int32_t* is_leaf, int32_t* obs_row) {
    ssu::NameIndex idx(obs_ids, (uint32_t)std::max(n_obs, 0));
    ssu::PostorderNodes p;
    uint32_t matched = 0;
    if (tr->n_parens < 0 ||
        !ssu::synced_postorder(tr->structure, tr->lengths, tr->names, (uint32_t)tr->n_parens, idx, p, matched))
        return -1;
    for (uint32_t k = 0; k < p.m; ++k) {
        parent[k] = p.parent[k];
        length[k] = p.length[k];
        is_leaf[k] = p.is_leaf[k];
        obs_row[k] = p.obs_row[k];
    }
    return (int)p.m;
}

int ssu_newick_to_bp(const char* newick, int cap, uint8_t* structure, double* lengths, char** names_buf,
                     int name_cap) {
    PhyloTree t = PhyloTree::from_newick(newick);
    const int n = (int)t.nparens();
    if (n > cap) return -n;
    for (int i = 0; i < n; ++i) {
        structure[i] = t.open[i];
        lengths[i] = t.length[i];
        strncpy(names_buf[i], t.name[i].c_str(), name_cap - 1);
        names_buf[i][name_cap - 1] = 0;
    }
    return n;
}

int ssu_device_available(void) {
    std::string err;
    return ssu::device_available(&err) ? 1 : 0;
}

const char* ssu_last_error(void) { return ssu::last_error().c_str(); }
