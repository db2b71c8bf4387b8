/*
 * ORACLE -- test infrastructure only.
 *
 * A harness around the reference implementation's own CPU code, compiled
 * straight from the reference sources by oracle/Makefile into oracle/_ref/:
 *   src/tree.cpp             (BPTree: newick parse, shear, collapse, postorder)
 *   src/biom_inmem.cpp       (in-memory CSR table, sample totals)
 *   src/unifrac_internal.cpp (PropStack, set_proportions_range)
 *   src/unifrac_task_impl.hpp, included below (the run_*_T stripe kernels,
 *                            CPU branch, and compute_stripes_totals_T)
 *
 * The reference's own driver (src/unifrac_cmp.cpp unifracTT / unifrac_vawTT)
 * and task classes (src/unifrac_task.hpp) depend on generated glue code and
 * are not built.  This file restates that driver loop -- node batches in
 * postorder, the per-batch embedding layouts of the CPU build, the stripe
 * buffers -- and hands every batch to the reference kernels unchanged.
 */
#include "tree.hpp"
#include "biom_inmem.hpp"
#include "unifrac_internal.hpp"
#include "unifrac_task_impl.hpp"

#include <cstring>
#include <string>
#include <unordered_set>
#include <vector>

namespace {

su::BPTree make_tree(int n_parens, const unsigned char* structure, const double* lengths, const char* const* names) {
    std::vector<bool> s(n_parens);
    for (int i = 0; i < n_parens; i++) s[i] = structure[i] != 0;
    // src/tree.cpp:72-99 (c-array constructor) needs bool*, so go through a bool copy
    bool* bs = new bool[n_parens > 0 ? n_parens : 1];
    for (int i = 0; i < n_parens; i++) bs[i] = s[i];
    su::BPTree t(bs, lengths, names, n_parens);
    delete[] bs;
    return t;
}

// src/api.cpp:90-92 (SYNC_TREE_TABLE)
su::BPTree sync_tree(const su::BPTree& tree, const su::biom_inmem& table) {
    std::unordered_set<std::string> keep(table.get_obs_ids().begin(), table.get_obs_ids().end());
    return tree.shear(keep).collapse();
}

// Straight / transposed / bit-packed batch layouts, after src/unifrac_task.hpp:208-319 (CPU build: transposed)
template <class TF>
struct Batch {
    uint64_t n, nr;  // samples, samples rounded up to 64
    unsigned max_embs;
};

template <class TF>
void run_task(int method, bool vaw, const su::BPTree& tree, const su::biom_inmem& table, double alpha,
              bool bypass_tips, bool normalize, unsigned start, unsigned stop, double* out) {
    const uint64_t n = table.n_samples;
    const uint64_t nr = ((n + 63) / 64) * 64;
    const uint64_t n_str = stop - start;
    const bool want_total = (method == 0 || method == 2 || method == 4);
    const bool bool_emb = (method == 0 || method == 1);

    // batch sizes: src/unifrac_task.hpp:392-398 (CPU), :661 (VAW)
    unsigned max_emb = vaw ? 128u : (bool_emb ? 64u * 32u : 512u - 16u);

    std::vector<TF> dm(n_str * nr, TF(0)), tot(n_str * nr, TF(0));
    std::vector<TF> lengths(max_emb);
    // embedding buffers
    const uint64_t emb_els_u64 = (max_emb + 63) / 64;   // unweighted, uint64 packed
    const uint64_t emb_els_u32 = (max_emb + 31) / 32;   // VAW unweighted, uint32 packed
    std::vector<TF> emb_f;           // transposed [nr][max_emb] or VAW straight [max_emb][nr]
    std::vector<uint64_t> emb_b64;   // transposed [nr][emb_els_u64]
    std::vector<uint32_t> emb_b32;   // VAW straight [emb_els_u32][nr]
    std::vector<TF> emb_c;           // VAW counts straight [max_emb][nr]
    std::vector<TF> stc;             // VAW sample totals
    if (vaw) {
        if (bool_emb) emb_b32.assign(emb_els_u32 * nr, 0u);
        else emb_f.assign((uint64_t)max_emb * nr, TF(0));
        emb_c.assign((uint64_t)max_emb * nr, TF(0));
        stc.assign(nr, TF(0));
        const double* sc = table.get_sample_counts();
        for (uint64_t i = 0; i < n; i++) stc[i] = sc[i];
    } else if (bool_emb) {
        emb_b64.assign(nr * emb_els_u64, 0ull);
    } else {
        emb_f.assign(nr * (uint64_t)max_emb, TF(0));
    }
    std::vector<bool> zcheck_v(n);
    bool* zcheck = new bool[n];
    std::vector<TF> sums(std::max<uint64_t>(n, (uint64_t)max_emb * (0x400 / 32)));
    std::vector<TF> stripe_sums(n);

    // src/unifrac_cmp.cpp:104-136: sample chunks of PropStackFixed::DEF_VEC_SIZE, one OpenMP task each
    su::PropStackMulti<TF> props(n), counts(n);
    const unsigned n_chunks = props.get_num_stacks();
    const unsigned max_k = (tree.nparens > 1) ? ((tree.nparens / 2) - 1) : 0;
    unsigned k = 0;
    while (k < max_k) {
        const unsigned k_start = k;
        unsigned filled = 0;
#pragma omp parallel for
        for (unsigned ck = 0; ck < n_chunks; ck++) {
            su::PropStack<TF>& ps = props.get_prop_stack(ck);
            su::PropStack<TF>& cs = counts.get_prop_stack(ck);
            const unsigned ts = props.get_start(ck), te = props.get_end(ck);
            unsigned my_filled = 0, my_k = k_start;
            while (my_filled < max_emb && my_k < max_k) {
                const uint32_t node = tree.postorderselect(my_k);
                my_k++;
                TF* np = ps.pop(node);
                su::set_proportions_range(np, tree, node, table, ts, te, ps, normalize);
                TF* nc = nullptr;
                if (vaw) {
                    nc = cs.pop(node);
                    su::set_proportions_range(nc, tree, node, table, ts, te, cs, false);
                }
                if (bypass_tips && tree.isleaf(node)) continue;
                if (ck == 0) { lengths[my_filled] = tree.lengths[node]; }
                const unsigned e = my_filled;
                if (vaw) {
                    if (bool_emb) {  // src/unifrac_task.hpp:252-283 (straight bool, uint32)
                        const unsigned blk = e / 32, bit = e % 32;
                        for (uint64_t i = ts; i < te; i++) {
                            uint32_t v = (np[i - ts] > 0);
                            if (bit == 0) emb_b32[blk * nr + i] = v; else emb_b32[blk * nr + i] |= (v << bit);
                        }
                        if (bit == 0 && te == n) for (uint64_t i = n; i < nr; i++) emb_b32[blk * nr + i] = 0;
                    } else {         // src/unifrac_task.hpp:208-224 (straight)
                        for (uint64_t i = ts; i < te; i++) emb_f[(uint64_t)e * nr + i] = np[i - ts];
                        if (te == n) for (uint64_t i = n; i < nr; i++) emb_f[(uint64_t)e * nr + i] = 0;
                    }
                    for (uint64_t i = ts; i < te; i++) emb_c[(uint64_t)e * nr + i] = nc[i - ts];
                    if (te == n) for (uint64_t i = n; i < nr; i++) emb_c[(uint64_t)e * nr + i] = 0;
                } else if (bool_emb) {  // src/unifrac_task.hpp:286-319 (transposed bool, uint64)
                    const unsigned blk = e / 64, bit = e % 64;
                    for (uint64_t i = ts; i < te; i++) {
                        uint64_t v = (np[i - ts] > 0);
                        if (bit == 0) emb_b64[i * emb_els_u64 + blk] = v; else emb_b64[i * emb_els_u64 + blk] |= (v << bit);
                    }
                    if (bit == 0 && te == n) for (uint64_t i = n; i < nr; i++) emb_b64[i * emb_els_u64 + blk] = 0;
                } else {                // src/unifrac_task.hpp:228-244 (transposed)
                    for (uint64_t i = ts; i < te; i++) emb_f[i * max_emb + e] = np[i - ts];
                    if (te == n) for (uint64_t i = n; i < nr; i++) emb_f[i * max_emb + e] = 0;
                }
                my_filled++;
            }
            if (ck == 0) { k = my_k; filled = my_filled; }
        }
        if (filled == 0) continue;
        if (vaw) {
            switch (method) {
            case 0: run_VawUnweightedTask_T<TF>(filled, start, stop, n, nr, lengths.data(), emb_b32.data(), emb_c.data(), stc.data(), dm.data(), tot.data()); break;
            case 1: run_VawUnnormalizedUnweightedTask_T<TF>(filled, start, stop, n, nr, lengths.data(), emb_b32.data(), emb_c.data(), stc.data(), dm.data()); break;
            case 2: run_VawNormalizedWeightedTask_T<TF>(filled, start, stop, n, nr, lengths.data(), emb_f.data(), emb_c.data(), stc.data(), dm.data(), tot.data()); break;
            case 3: run_VawUnnormalizedWeightedTask_T<TF>(filled, start, stop, n, nr, lengths.data(), emb_f.data(), emb_c.data(), stc.data(), dm.data()); break;
            case 4: run_VawGeneralizedTask_T<TF>(filled, start, stop, n, nr, lengths.data(), emb_f.data(), emb_c.data(), stc.data(), dm.data(), tot.data(), (TF)alpha); break;
            }
        } else {
            switch (method) {
            case 0: run_UnweightedTask_T<TF>(emb_els_u64, filled, start, stop, n, nr, lengths.data(), emb_b64.data(), dm.data(), tot.data(), sums.data(), zcheck, nullptr, stripe_sums.data()); break;
            case 1: run_UnnormalizedUnweightedTask_T<TF>(emb_els_u64, filled, start, stop, n, nr, lengths.data(), emb_b64.data(), dm.data(), sums.data(), zcheck, nullptr, stripe_sums.data()); break;
            case 2: run_NormalizedWeightedTask_T<TF>(max_emb, filled, start, stop, n, nr, lengths.data(), emb_f.data(), dm.data(), tot.data(), zcheck, sums.data()); break;
            case 3: run_UnnormalizedWeightedTask_T<TF>(max_emb, filled, start, stop, n, nr, lengths.data(), emb_f.data(), dm.data(), zcheck, sums.data()); break;
            case 4: run_GeneralizedTask_T<TF>(max_emb, filled, start, stop, n, nr, lengths.data(), emb_f.data(), dm.data(), tot.data(), (TF)alpha); break;
            }
        }
    }
    if (want_total) compute_stripes_totals_T<TF>(dm.data(), tot.data(), n_str * nr);
    for (uint64_t s = 0; s < n_str; s++)
        for (uint64_t i = 0; i < n; i++) out[s * n + i] = dm[s * nr + i];
    delete[] zcheck;
}

}  // namespace

extern "C" {

/* Newick -> parentheses via src/tree.cpp:30-70 (BPTree(newick)).  Arrays must hold cap entries. */
int ref_newick_to_bp(const char* newick, int cap, unsigned char* structure, double* lengths, char** names_out, int name_cap) {
    su::BPTree t(newick);
    const std::vector<bool>& s = t.get_structure();
    if ((int)t.nparens > cap) return -(int)t.nparens;
    for (uint32_t i = 0; i < t.nparens; i++) {
        structure[i] = s[i] ? 1 : 0;
        lengths[i] = t.lengths[i];
        std::strncpy(names_out[i], t.names[i].c_str(), name_cap - 1);
        names_out[i][name_cap - 1] = 0;
    }
    return (int)t.nparens;
}

/*
 * Sheared+collapsed postorder arrays via the reference BPTree
 * (postorderselect, isleaf, leftchild, rightsibling).  Arrays hold n_parens/2.
 */
int ref_postorder(int n_parens, const unsigned char* structure, const double* lengths, const char* const* names,
                  int n_obs, int n_samples, const char* const* obs_ids, const char* const* sample_ids,
                  uint32_t* indptr, uint32_t* indices, double* data,
                  int* parent, double* len, int* is_leaf, int* obs_row) {
    su::BPTree t0 = make_tree(n_parens, structure, lengths, names);
    su::biom_inmem table(obs_ids, sample_ids, indices, indptr, data, n_obs, n_samples);
    su::BPTree t = sync_tree(t0, table);
    const unsigned max_k = (t.nparens > 1) ? ((t.nparens / 2) - 1) : 0;
    std::vector<int> post_of(t.nparens, -1);
    for (unsigned k = 0; k <= max_k && k < t.nparens / 2; k++) post_of[t.postorderselect(k)] = (int)k;
    for (unsigned k = 0; k < max_k; k++) parent[k] = (int)max_k;
    std::unordered_map<std::string, int> row;
    for (int i = 0; i < n_obs; i++) row[obs_ids[i]] = i;
    for (unsigned k = 0; k < max_k; k++) {
        const uint32_t node = t.postorderselect(k);
        len[k] = t.lengths[node];
        is_leaf[k] = t.isleaf(node) ? 1 : 0;
        obs_row[k] = -1;
        if (is_leaf[k]) {
            auto it = row.find(t.names[node]);
            if (it != row.end()) obs_row[k] = it->second;
        } else {
            uint32_t c = t.leftchild(node), r = t.rightchild(node);
            while (c <= r && c != 0) { parent[post_of[c]] = (int)k; c = t.rightsibling(c); }
        }
    }
    return (int)max_k;
}

/*
 * Stripes [start, stop) computed by the reference kernels.  method: 0 unweighted,
 * 1 unweighted_unnormalized, 2 weighted_normalized, 3 weighted_unnormalized,
 * 4 generalized.  fp32 != 0 selects the float instantiations (the default,
 * non-"_fp64" method names, src/api.cpp:55-88).
 */
int ref_unifrac_stripes(int n_parens, const unsigned char* structure, const double* lengths, const char* const* names,
                        int n_obs, int n_samples, const char* const* obs_ids, const char* const* sample_ids,
                        uint32_t* indptr, uint32_t* indices, double* data,
                        int method, int fp32, int variance_adjust, double alpha, int bypass_tips, int normalize,
                        unsigned start, unsigned stop, double* out) {
    su::BPTree t0 = make_tree(n_parens, structure, lengths, names);
    su::biom_inmem table(obs_ids, sample_ids, indices, indptr, data, n_obs, n_samples);
    su::BPTree t = sync_tree(t0, table);
    su::register_report_status();
    if (fp32) run_task<float>(method, variance_adjust != 0, t, table, alpha, bypass_tips != 0, normalize != 0, start, stop, out);
    else run_task<double>(method, variance_adjust != 0, t, table, alpha, bypass_tips != 0, normalize != 0, start, stop, out);
    su::remove_report_status();
    return 0;
}

/* Faith's PD: the loop of src/unifrac.cpp:362-387 over the reference set_proportions */
int ref_faith_pd(int n_parens, const unsigned char* structure, const double* lengths, const char* const* names,
                 int n_obs, int n_samples, const char* const* obs_ids, const char* const* sample_ids,
                 uint32_t* indptr, uint32_t* indices, double* data, double* out) {
    su::BPTree t0 = make_tree(n_parens, structure, lengths, names);
    su::biom_inmem table(obs_ids, sample_ids, indices, indptr, data, n_obs, n_samples);
    su::BPTree t = sync_tree(t0, table);
    su::PropStack<double> ps(table.n_samples);
    for (int s = 0; s < n_samples; s++) out[s] = 0.0;
    const unsigned max_k = (t.nparens > 1) ? ((t.nparens / 2) - 1) : 0;
    for (unsigned k = 0; k < max_k; k++) {
        const uint32_t node = t.postorderselect(k);
        const double length = t.lengths[node];
        double* np = ps.pop(node);
        su::set_proportions(np, t, node, table, ps);
        for (int s = 0; s < n_samples; s++) out[s] += (np[s] > 0) * length;
    }
    return 0;
}

/* src/biom_inmem.cpp:513-526 */
int ref_sample_counts(int n_obs, int n_samples, const char* const* obs_ids, const char* const* sample_ids,
                      uint32_t* indptr, uint32_t* indices, double* data, double* out) {
    su::biom_inmem table(obs_ids, sample_ids, indices, indptr, data, n_obs, n_samples);
    const double* sc = table.get_sample_counts();
    for (int s = 0; s < n_samples; s++) out[s] = sc[s];
    return 0;
}

}  // extern "C"
