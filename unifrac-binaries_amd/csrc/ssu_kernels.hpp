// Launchers for the gfx950 kernels.  Plain pointers and sizes only; every
// launcher enqueues on the given stream and returns the launch status.
#pragma once

#include <hip/hip_runtime.h>

#include <vector>
#include <cstddef>
#include <cstdint>

namespace ssu {

// ---- K-chunk descriptor of the exact integer Gram (unweighted) -----------
// A chunk is a run of K' columns of the presence matrix whose branch lengths
// share one power-of-two scale.  Column c of the chunk carries 7-bit digits
// d_0..d_3 of the integer mantissa for digit group g:
//   length = scale * sum_q d_q * 2^(7q)    (this group's share)
struct GramChunk {
    uint32_t k0, k1;   // [k0, k1) in padded K' columns, multiples of GRAM_BK
    double scale;      // 2^e of digit 0 of this group
};

// Sample pairs (i, j), i < j, selected by their gap g = j - i: a0 <= g <= a1
// or b0 <= g <= b1.  The stripes [s0, s1) of an n-sample matrix are such a
// set (stripe s holds the pairs of gaps s + 1 and n - s - 1,
// src/unifrac_task_impl.hpp:367); the default selects every pair.
struct GapBand {
    uint32_t a0 = 1, a1 = 0xffffffffu, b0 = 1, b1 = 0;
    __host__ __device__ bool has(uint32_t g) const { return (g >= a0 && g <= a1) || (g >= b0 && g <= b1); }
    // some gap in [lo, hi] is selected
    __host__ __device__ bool meets(uint32_t lo, uint32_t hi) const {
        return lo <= hi && ((lo <= a1 && hi >= a0) || (lo <= b1 && hi >= b0));
    }
    // the tile rows [r0, r1) x columns [c0, c1) holds a selected pair i < j
    __host__ __device__ bool meets_tile(uint32_t r0, uint32_t r1, uint32_t c0, uint32_t c1) const {
        if (c1 == 0 || r1 <= r0 || c1 <= c0 || c1 - 1 <= r0) return false;
        const uint32_t lo = c0 + 1 > r1 ? c0 + 1 - r1 : 1u;
        return meets(lo > 0 ? lo : 1u, c1 - 1 - r0);
    }
    bool all(uint32_t n) const { return a0 <= 1 && (uint64_t)a1 + 1 >= n; }
};
// the pairs of stripes [s0, s1) of an n-sample matrix (s1 <= (n + 1) / 2)
GapBand stripe_band(uint32_t n, uint32_t s0, uint32_t s1);

constexpr int GRAM_BM = 128;   // output tile rows per workgroup
constexpr int GRAM_BK = 128;   // K' columns per LDS stage
constexpr int PAIR_TILE = 128; // samples per side of a VALU pair tile

// p[0..n) = v (a kernel, so that it can sit in a captured graph)
hipError_t launch_fill(double* p, size_t n, double v, hipStream_t st);
hipError_t launch_fill_u32(uint32_t* p, size_t n, uint32_t v, hipStream_t st);

// ---------------------------------------------------------------- near-duplicate refinement
// FP32 weighted / generalized distances carry the FP32 rounding of the
// proportions: |u - v| computed from rounded values is off by at most
// 2^-24 (u + v), so d is off by at most 2^-24 of its denominator, a relative
// error of up to 2^-24 / d.  Selected pairs (i < j in band) with d below tau --
// below tau * (S_i + S_j) when S is given (unnormalised weighted) -- are
// marked: flags[0] = their number, flags[1 + s] = 1 for every sample s of
// one (flags holds n + 1 words and is cleared here first).
hipError_t launch_flag_close(uint32_t* flags, const float* out, uint32_t n, const double* S, float tau,
                             const GapBand& band, hipStream_t st);
// Pairs among the sub-table samples idx[0..ns) that launch_flag_close marked
// (the same test on the same FP32 value and S) take the FP64 distance of the
// sub-table's ns x ns matrix, in both triangles; with cross, only pairs
// a < na <= b.
hipError_t launch_refine_scatter(float* out, uint32_t n, const double* sub, uint32_t ns, const uint32_t* idx,
                                 uint32_t na, bool cross, const double* S, float tau, hipStream_t st);
// sub (ns x ns) = the entries of out (n x n) at rows and columns idx[0..ns)
hipError_t launch_gather_square(float* sub, const float* out, uint32_t n, const uint32_t* idx, uint32_t ns,
                                hipStream_t st);

// ---------------------------------------------------------------- embedding
// E holds the sample columns [col0, col0 + ld) of the node-major embedding
hipError_t launch_embed_leaves(double* E, size_t ld, uint32_t col0, const int32_t* leaf_k, const int32_t* leaf_row,
                               uint32_t n_leaves, const uint32_t* indptr, const uint32_t* indices, const double* data,
                               const double* totals, bool normalize, hipStream_t st);
hipError_t launch_embed_level(double* E, size_t ldn, const int32_t* nodes, uint32_t n_nodes, const int32_t* child_ptr,
                              const int32_t* child_idx, hipStream_t st);

// Internal nodes along heavy paths (each node continues the path of its
// largest child; the paths of a round depend only on earlier rounds): path p
// holds nodes[off[p] .. off[p + 1]), bottom to top, light[q] the node's other
// child when it has two (-1 otherwise), heavy[k] every internal node's heavy
// child.  The launches run the paths [p0, p1) of one round; the FP64 sums are
// those of launch_embed_level (children left to right from 0).
struct HeavyPaths {
    const int32_t* nodes;
    const int32_t* light;
    const uint32_t* off;
    const int32_t* heavy;
    const int32_t* child_ptr;
    const int32_t* child_idx;
};
hipError_t launch_embed_paths(double* E, size_t ldn, const HeavyPaths& hp, uint32_t p0, uint32_t p1, hipStream_t st);
hipError_t launch_bits_paths(uint32_t* B, uint32_t ldb, const HeavyPaths& hp, uint32_t p0, uint32_t p1,
                             hipStream_t st);

// Presence bits (node-major, ldb words of 32 samples per row; every sample of
// the table in one block): bit s%32 of B[k][s/32] = sample s holds node k.
// Leaves from the CSR (value > 0, or every stored entry when `structural`),
// internal nodes level by level as the OR of their children.
constexpr uint32_t BITS_MAX_WORDS = 8192;   // 262,144 samples: a leaf row is gathered in 32 KB of LDS
hipError_t launch_bits_leaves(uint32_t* B, uint32_t ldb, const int32_t* leaf_k, const int32_t* leaf_row,
                              uint32_t n_leaves, const uint32_t* indptr, const uint32_t* indices, const double* data,
                              bool structural, hipStream_t st);
hipError_t launch_bits_level(uint32_t* B, uint32_t ldb, const int32_t* nodes, uint32_t n_nodes,
                             const int32_t* child_ptr, const int32_t* child_idx, hipStream_t st);
// cnt[r] = popcount of B[knodes[r]]
hipError_t launch_count_bits(uint32_t* cnt, const uint32_t* B, uint32_t ldb, const int32_t* knodes, uint32_t K,
                             hipStream_t st);
// smask[p][r] = B[nodes[r]][p], p < npan (the sample bits launch_gather_sums writes)
hipError_t launch_bits_smask(uint32_t* smask, const uint32_t* B, uint32_t ldb, const int32_t* nodes, uint32_t K,
                             uint32_t npan, hipStream_t st);

// One segmented pass over the processed rows r = 0..K-1 (row r is E[knodes[r]])
// of an embedding block holding sample columns [col0, col0 + lde):
//   U (if non-null): node-major FP32 copy, U[r][col0 + s] = (float)E[knodes[r]][s] (row stride ldu);
//   S (if non-null): S[s] = sum_r klen[r] * u_rs in FP64 (u as the pair kernels
//     see it: FP32-rounded when U is written), summed per segment then over
//     segments in order, into S[col0 ..]; Spart holds gather_segments(K, lde) * lde doubles;
//   masks (if non-null; lde, col0 multiples of 64): [ldn/32][ceil(K/32)] words, bit r%32 of
//     word r/32 of panel p set iff some sample of p*32..p*32+31 has u_rs > 0;
//   smask (if non-null, with masks): [ldn/32][K] words, bit s%32 of smask[s/32][r] = u_rs > 0;
//   skip_u (optional, [ceil(K/32)] words): rows r with bit r set are not written to U (S and
//     the masks still see their FP32-rounded values).
uint32_t gather_segments(uint32_t K, size_t ldn);
//   rps_fixed (optional, a multiple of 32): rows per segment, in place of the
//     default that depends on lde (S then depends on a sample's column alone).
hipError_t launch_gather_sums(float* U, size_t ldu, double* S, double* Spart, uint32_t* masks, uint32_t* smask,
                              const double* E,
                              size_t lde, uint32_t col0, const int32_t* knodes, const double* klen, uint32_t K,
                              uint32_t n, hipStream_t st, const uint32_t* skip_u = nullptr, uint32_t rps_fixed = 0);
// rows per segment that depends on K alone (for launch_gather_sums)
uint32_t fixed_rows_per_segment(uint32_t K);
// Faith PD: out[s] = sum over all non-root nodes k in postorder of (E[k][s] > 0) * len[k]
hipError_t launch_faith(double* out, const double* E, size_t ldn, const double* len, uint32_t m, uint32_t n,
                        hipStream_t st);

// ---------------------------------------------------------------- Gram path
// Pt[s][c] = 0xFF if E[col_node[c]][s] > 0 else 0 (col_node < 0 -> padding column)
// for the embedding block of sample columns [col0, col0 + lde): Pt rows col0 ..
hipError_t launch_presence_T(int8_t* Pt, size_t ldp, const double* E, size_t lde, uint32_t col0, const int32_t* col_node,
                             uint32_t Kp, uint32_t n, hipStream_t st);
// Pt[s][c] for s < n_pad from presence bits (rows past n zero)
hipError_t launch_presence_T_bits(int8_t* Pt, size_t ldp, const uint32_t* B, uint32_t ldb, const int32_t* col_node,
                                  uint32_t Kp, uint32_t n_pad, uint32_t n, hipStream_t st);
// G = sum_c len_c * P[i][c] * P[j][c] for the upper-triangle 128x128 tiles, FP64,
// stored tile by tile in MFMA fragment order (n_pad/128)^2 * 128^2 doubles;
// order lists the tiles computed (every tile, or a subset)
hipError_t launch_gram_i8(double* G, const int8_t* Pt, size_t ldp, const int8_t* digits, const GramChunk* chunks,
                          uint32_t n_chunks, uint32_t n_pad, const uint32_t* order, uint32_t n_tiles, hipStream_t st);
// out[i][j] from G and T = diag(G): (T_i+T_j-2G)/(T_i+T_j-G) or T_i+T_j-2G, rows i in [row0, row1)
template <class TOut>
hipError_t launch_unweighted_combine(TOut* out, const double* G, uint32_t n_pad, uint32_t n, bool normalized,
                                     hipStream_t st, uint32_t row0 = 0, uint32_t row1 = 0xffffffffu);
// the same values for the pairs (i, j), i in [row0, row1) (multiples of 64, or row1 = n),
// j >= row0, written at (i, j) and (j, i): afterwards rows [row0, row1) from column row0 on
// and columns [row0, row1) of the rows below are final (G's tiles of those rows and its
// diagonal must be complete)
template <class TOut>
hipError_t launch_unweighted_combine_band(TOut* out, const double* G, uint32_t n_pad, uint32_t n, bool normalized,
                                          hipStream_t st, uint32_t row0, uint32_t row1);

// ---------------------------------------------------------------- VALU pair tiles
enum class PairOp : int {
    weighted = 0,        // sum l |u - v|
    unweighted_plain,    // sum l [u>0 xor v>0], sum l [u>0 or v>0]  (FP32 stripe baseline)
    generalized,         // sum l (u+v)^a |u-v|/(u+v), sum l (u+v)^a
    vaw_weighted,        // sum l |u-v| / vaw, sum l (u+v) / vaw
    vaw_generalized,
    vaw_unweighted,      // sum l xor / vaw, sum l or / vaw
};

struct PairArgs {
    const void* U;          // node rows, float or double; row of pair-node k is rows[k] (or k)
    const void* Cn;         // raw-count rows, same indexing (VAW only)
    const int32_t* rows;    // nullptr: identity
    size_t ldn;
    const void* len;        // [K] lengths, float or double
    uint32_t K;
    uint32_t n;
    const double* S;        // per-sample denominators for weighted_normalized (nullptr otherwise)
    const double* totals;   // per-sample raw totals (VAW)
    double alpha;
    bool normalized;        // divide by the accumulated / S-based total
    void* out;              // n x n output
    bool out_fp64;
    GapBand band;           // tiles without a selected pair are skipped
    // two-accumulator ops (VAW): with out2, out[i * n + j] and out2[i * n + j] (i < j) hold
    // the posting-list shares of the numerator and denominator, added before the division
    const void* out2 = nullptr;
    // every length >= 0 (VAW weighted / presence terms stage l u instead of u)
    bool len_nonneg = false;
    // unnormalised VAW weighted: with flags, a pair (i < j) whose distance is below
    // flag_tau of its VAW denominator sets flags[1 + i], flags[1 + j] and adds 1 to flags[0]
    uint32_t* flags = nullptr;
    float flag_tau = 0.0f;
};
hipError_t launch_pairs(PairOp op, bool fp64, const PairArgs& a, hipStream_t st);

// Weighted (normalized or not).  FP64 (launch_pairs_sparse): d_ij = S_i + S_j -
// 2 sum_k l_k min(u_ik, u_jk), visiting per 64x64 block only the nodes whose
// bit is set in both panels' masks.  FP32 (launch_pairs_sparse32): sum_k l_k
// |u_ik - u_jk| itself over the nodes either panel holds; S normalises.
struct SparseArgs {
    const void* U;          // node rows, float (row r) or double (row rows[r])
    const int32_t* rows;    // nullptr: identity
    size_t ldn;             // multiple of 64
    const double* len;      // [K] lengths as the accumulation uses them
    const uint32_t* masks;  // from launch_gather_sums
    uint32_t Kw;            // ceil(K / 32)
    uint32_t n;
    const double* S;        // per-sample sums from launch_gather_sums
    bool normalized;
    void* out;              // n x n
    bool out_fp64;
    uint32_t* visits;       // optional [nb][nb], nb = ceil(n/64): nodes visited per block (bi <= bj)
    const uint32_t* order;  // device copy of upper_block_order(ceil(n/64), 16)
    uint32_t n_blocks;      // its length, nb (nb + 1) / 2
    const uint32_t* dmask = nullptr;   // visit only these nodes and add the posting-list shares in `out`
};

// Occupancy split.  Weighted (shared-node path, FP32 or FP64 values): nodes
// held by at least `thr` samples are dense (bit set in dmask, [ceil(K/32)]
// words, set here by launch_sparse_lists); for the others the posting lists
// are built from the value rows and the sample masks at offsets scanned from
// their counts, and each row's share over them is written, in the value type,
// to out[a * n + b] (b > a) for the tiled kernel to add (SparseArgs::dmask):
// with FP32 values the share of sum l |u_a - u_b| (ssp[a] + ssp[b] - 2 M,
// formed in FP64), with FP64 values the share of M.  Unweighted (gram != nullptr): the K nodes are the
// plan's sparse nodes with offsets fixed when the plan is built (dmask all
// zero, U and pval unused); each row's share of the shared length
// sum_r len[r] [a holds r][b holds r] is added to the FP64 Gram tile entries
// G(a, b), b >= a, after launch_gram_i8 has written them.  Posting and
// row-list arrays hold `slots` entries, which must bound the postings of any
// execution (fewer than 2^30: 32-bit slot offsets).
struct SparseNodesArgs {
    uint32_t* dmask;
    const uint32_t* smask;  // [npan][K] sample bits, from launch_gather_sums / launch_bits_smask
    uint32_t* pc;           // [npan + 1][K]
    void* pidx;             // [slots] sample indices: u16 when n <= 65,536, else u32
    void* pval;             // [slots] (weighted) values: float, or double with values_fp64
    uint32_t* poff;         // [K + 1] posting offsets (weighted: written here; unweighted: given)
    uint32_t* pcnt;         // [K] (weighted) sparse nodes' sample counts
    uint32_t* obsum;        // [ceil(K / 4096) + 1] (weighted) scan scratch
    uint32_t slots;
    uint32_t thr;           // (weighted) dense from this many samples
    const void* U;          // (weighted) value rows: row urows[r] (or r) of U, stride ldu
    const int32_t* urows;   // nullptr: identity
    bool values_fp64;       // (weighted) FP64 values, FP64 output
    bool src_fp64 = false;  // (weighted) U holds FP64 rows (FP32 values: rounded when copied)
    size_t ldu;
    const double* len;      // [K]
    const double* S;        // [n] (weighted)
    void* out;              // (weighted)
    bool out_fp64;
    uint32_t* walked;       // optional [sparse_rows_blocks(n)]: postings walked per wave
    unsigned long long* rows;   // [slots] row-list entries (node << 32 | position in its postings)
    uint32_t* rcount;       // [n * row_list_chunks(n, K)]
    uint32_t* roff;         // [n * row_list_chunks(n, K) + 1]
    uint32_t* rbsum;        // [row_list_scan_blocks(n, K) + 1]
    uint32_t K, n, npan;
    double* ssp = nullptr;  // (weighted, FP32 values) [n]: per-sample sum of l * u (GEN: l * u^alpha) over its sparse nodes
    bool gen = false;       // generalized: numerator shares to out, denominator shares to out2
    double alpha = 1.0;
    void* out2 = nullptr;
    double* pspmax = nullptr;   // (GEN) [1] scratch: max of ssp
    double* gram = nullptr; // unweighted: the Gram tiles of launch_gram_i8
    uint32_t n_pad = 0;     // their padded edge
    GapBand band;           // rows walk (and write) only the columns of selected pairs (G(a, a) always)
    // dense nodes fixed in advance ([ceil(K/32)] words, in place of the occupancy test)
    const uint32_t* fixed_dense = nullptr;
    // VAW (1: weighted values, 2: presence, 3: generalized, 0 <= alpha <= 1): numerator shares
    // to out, denominator shares to out2, each H(a, T_b) + H(b, T_a) plus the walk's
    // corrections of the shared nodes
    int vaw = 0;
    const void* Cv = nullptr;       // count rows, indexed as U
    void* pcv = nullptr;            // [slots] counts, in the value type
    const double* tot = nullptr;    // [n] sample totals
    const double* tau = nullptr;    // [D] the totals the one-sided sums are taken at
    const uint32_t* kidx = nullptr; // [n] sample -> its total's index in tau (nullptr: identity, D = n)
    uint32_t D = 0;
    double* H = nullptr;            // [n][D] one-sided sums
    double* nbound = nullptr;       // [K] per-node correction bounds (ssp: their row sums)
};
// posting lists, per-panel prefix counts and row lists
hipError_t launch_sparse_lists(const SparseNodesArgs& a, hipStream_t st);
// each row's share over the posting lists (after launch_sparse_lists; in the
// unweighted case after launch_gram_i8 as well), for rows [row0, row1): the
// columns past each row, and (unweighted) the diagonal G(a, a), as `parts`
// says.  Weighted and generalized sums need every row in one call.
constexpr int ROWS_OFF_DIAGONAL = 1, ROWS_DIAGONAL = 2;
hipError_t launch_sparse_rows(const SparseNodesArgs& a, hipStream_t st, uint32_t row0 = 0,
                              uint32_t row1 = 0xffffffffu, int parts = ROWS_OFF_DIAGONAL | ROWS_DIAGONAL);
// both, in order (weighted)
hipError_t launch_sparse_nodes(const SparseNodesArgs& a, hipStream_t st);
// cnt[r] += number of samples s in [col0, min(n, col0 + lde)) with E[knodes[r]][s - col0] > 0
hipError_t launch_count_present(uint32_t* cnt, const double* E, size_t lde, uint32_t col0, const int32_t* knodes,
                                uint32_t K, uint32_t n, hipStream_t st);
uint32_t sparse_rows_blocks(uint32_t n);
// workgroups launch_sparse_rows launches (the entries of walked it writes)
uint32_t sparse_rows_launched(uint32_t n, bool gen);
size_t offset_scan_blocks(uint32_t K);
// row lists are built per (sample, chunk of row_list_width(n, K) nodes)
uint32_t row_list_width(uint32_t n, uint32_t K);
uint32_t row_list_chunks(uint32_t n, uint32_t K);
size_t row_list_scan_blocks(uint32_t n, uint32_t K);
// The upper-triangle blocks (bi <= bj < nb) as bi << 16 | bj, super x super
// super-tile by super-tile; kernels walk it in XCD-contiguous runs.
std::vector<uint32_t> upper_block_order(uint32_t nb, uint32_t super);
hipError_t launch_pairs_sparse(bool in_fp64, const SparseArgs& a, hipStream_t st);
// FP32 operands: 32 x 64 blocks (32-sample row panel pi <= 2 pj + 1, 64-sample
// column panel pj), FP32 terms merged into FP64; `order` = rect_block_order(n)
// and visits is [ceil(n/32)][ceil(n/64)].
std::vector<uint32_t> rect_block_order(uint32_t n);
hipError_t launch_pairs_sparse32(const SparseArgs& a, hipStream_t st);
// generalized (FP32, 0 <= alpha <= 1) on the same blocks: numerator and denominator summed
// directly over the nodes either panel holds; with dmask, adds the posting-list shares in
// out (numerator) and out2 (denominator)
hipError_t launch_pairs_gen32(const SparseArgs& a, float alpha, void* out2, hipStream_t st);
// the same on FP64 operands (U: the FP64 embedding, row rows[r] for processed node r),
// FP64 terms, sums and output
hipError_t launch_pairs_gen64(const SparseArgs& a, double alpha, void* out2, hipStream_t st);

// --------------------------------------------------------- output layouts
// stripes[s - start][k] = M[k][(k+s+1) % n]   (double output)
template <class TIn>
hipError_t launch_matrix_to_stripes(double* stripes, const TIn* M, uint32_t n, uint32_t start, uint32_t stop,
                                    hipStream_t st);

}  // namespace ssu
