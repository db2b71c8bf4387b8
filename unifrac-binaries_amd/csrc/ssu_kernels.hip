// gfx950 kernels for the Strided State UniFrac hot path.
//
//   embedding     level-synchronous postorder propagation of per-sample values
//                 (src/unifrac_internal.cpp:230-258 semantics, FP64)
//   Gram          unweighted UniFrac as an exact integer Gram on int8 MFMA:
//                 G = A diag(l) A^T over 0/1 presence, l split into 7-bit digits
//   pair tiles    register-tiled VALU kernels for weighted / generalized / VAW
//                 (src/unifrac_task_impl.hpp formulas), FP32 with FP64 merges or FP64
//   Faith PD      src/unifrac.cpp:362-387, same FP64 summation order
#include "ssu_kernels.hpp"

#include <hip/hip_runtime.h>
#include <algorithm>
#include <vector>
#include <cmath>

namespace ssu {

namespace {

constexpr int EMB_THREADS = 256;
constexpr uint32_t N_XCDS = 8;   // MI355X: 8 XCDs, each with its own L2

inline unsigned grid1(size_t n, unsigned b) { return (unsigned)((n + b - 1) / b); }

// ------------------------------------------------------------------ embedding

// One workgroup per leaf: clear the row (16-byte stores; ld is even), then
// scatter the leaf's CSR row (src/biom_inmem.cpp:473-497: value / total, or
// the raw value).  E holds the sample columns [col0, col0 + ld) of the embedding.
__global__ void __launch_bounds__(EMB_THREADS) k_embed_leaves(double* __restrict__ E, size_t ld, uint32_t col0,
                                                              const int32_t* __restrict__ leaf_k,
                                                              const int32_t* __restrict__ leaf_row,
                                                              const uint32_t* __restrict__ indptr,
                                                              const uint32_t* __restrict__ indices,
                                                              const double* __restrict__ data,
                                                              const double* __restrict__ totals, int normalize) {
    const uint32_t li = blockIdx.x;
    double* row = E + (size_t)leaf_k[li] * ld;
    double2* row2 = reinterpret_cast<double2*>(row);
    for (size_t s = threadIdx.x; s < ld / 2; s += blockDim.x) {   // streaming stores: read back from HBM later
        __builtin_nontemporal_store(0.0, &row2[s].x);
        __builtin_nontemporal_store(0.0, &row2[s].y);
    }
    __syncthreads();
    const int32_t r = leaf_row[li];
    if (r < 0) return;
    for (uint32_t j = indptr[r] + threadIdx.x; j < indptr[r + 1]; j += blockDim.x) {
        const uint32_t s = indices[j];
        if (s < col0 || s - col0 >= ld) continue;
        row[s - col0] = normalize ? data[j] / totals[s] : data[j];
    }
}

// Internal nodes of one height level: E[k] = ((0 + E[c0]) + E[c1]) + ..., children left to right.
// blockIdx.x covers samples, two per thread (16-byte accesses; ldn is even);
// blockIdx.y strides over the level's nodes.
__global__ void __launch_bounds__(EMB_THREADS) k_embed_level(double* __restrict__ E, size_t ldn,
                                                             const int32_t* __restrict__ nodes, uint32_t n_nodes,
                                                             const int32_t* __restrict__ child_ptr,
                                                             const int32_t* __restrict__ child_idx) {
    const size_t s2 = (size_t)blockIdx.x * EMB_THREADS + threadIdx.x;   // samples 2 s2, 2 s2 + 1
    const size_t ld2 = ldn / 2;
    if (s2 >= ld2) return;
    double2* E2 = reinterpret_cast<double2*>(E);
    for (uint32_t ni = blockIdx.y; ni < n_nodes; ni += gridDim.y) {
        const int32_t k = nodes[ni];
        double2 acc = make_double2(0.0, 0.0);
        for (int32_t c = child_ptr[k]; c < child_ptr[k + 1]; ++c) {
            const double2 v = E2[(size_t)child_idx[c] * ld2 + s2];
            acc.x += v.x;
            acc.y += v.y;
        }
        E2[(size_t)k * ld2 + s2] = acc;
    }
}

// Heavy paths (each internal node continues the path of its largest child):
// path p holds internal nodes nodes[off[p] .. off[p + 1]) bottom to top, the
// first one's heavy child a leaf.  Every other child of a path node heads a
// path of an earlier round (or is a leaf), so the paths of one round are
// independent and a thread walks its path keeping the running value in
// registers: a node reads only its other children.  light[q] is the other
// child of a two-child node (E = heavy + other: the same FP64 sum as
// (0 + c0) + c1 in either order), or -1 for any other child count (children
// added left to right from 0, the heavy child's value from the registers).
// The other children of up to PATH_BATCH consecutive two-child nodes are
// loaded before they are added, so the walk keeps that many loads in flight.
constexpr int PATH_BATCH = 8;
__global__ void __launch_bounds__(EMB_THREADS) k_embed_paths(double* __restrict__ E, size_t ldn,
                                                             const int32_t* __restrict__ nodes,
                                                             const int32_t* __restrict__ light,
                                                             const uint32_t* __restrict__ off, uint32_t n_paths,
                                                             const int32_t* __restrict__ heavy,
                                                             const int32_t* __restrict__ child_ptr,
                                                             const int32_t* __restrict__ child_idx) {
    const size_t s2 = (size_t)blockIdx.x * EMB_THREADS + threadIdx.x;   // samples 2 s2, 2 s2 + 1
    const size_t ld2 = ldn / 2;
    if (s2 >= ld2) return;
    double2* E2 = reinterpret_cast<double2*>(E);
    for (uint32_t p = blockIdx.y; p < n_paths; p += gridDim.y) {
        const uint32_t a = off[p], b = off[p + 1];
        double2 run = E2[(size_t)heavy[nodes[a]] * ld2 + s2];
        for (uint32_t q = a; q < b;) {
            int32_t lt[PATH_BATCH];
            int g = 0;
            bool more = true;
#pragma unroll
            for (int u = 0; u < PATH_BATCH; ++u) {
                lt[u] = q + u < b ? light[q + u] : -1;
                more = more && lt[u] >= 0;
                g += more ? 1 : 0;
            }
            double2 v[PATH_BATCH];
#pragma unroll
            for (int u = 0; u < PATH_BATCH; ++u)
                if (u < g) v[u] = E2[(size_t)lt[u] * ld2 + s2];
#pragma unroll
            for (int u = 0; u < PATH_BATCH; ++u)
                if (u < g) {
                    run.x += v[u].x;
                    run.y += v[u].y;
                    E2[(size_t)nodes[q + u] * ld2 + s2] = run;
                }
            q += g;
            if (g < PATH_BATCH && q < b) {   // a node with one or three or more children
                const int32_t k = nodes[q], h = heavy[k];
                double2 acc = make_double2(0.0, 0.0);
                for (int32_t c = child_ptr[k]; c < child_ptr[k + 1]; ++c) {
                    const int32_t ch = child_idx[c];
                    const double2 w = ch == h ? run : E2[(size_t)ch * ld2 + s2];
                    acc.x += w.x;
                    acc.y += w.y;
                }
                run = acc;
                E2[(size_t)k * ld2 + s2] = run;
                ++q;
            }
        }
    }
}

// Presence bits along heavy paths (as k_embed_paths, with OR): one thread per (path, word).
__global__ void __launch_bounds__(256) k_bits_paths(uint32_t* __restrict__ B, uint32_t ldb,
                                                    const int32_t* __restrict__ nodes,
                                                    const int32_t* __restrict__ light,
                                                    const uint32_t* __restrict__ off, uint32_t n_paths,
                                                    const int32_t* __restrict__ heavy,
                                                    const int32_t* __restrict__ child_ptr,
                                                    const int32_t* __restrict__ child_idx) {
    const uint32_t w = blockIdx.x * 256 + threadIdx.x;
    if (w >= ldb) return;
    for (uint32_t p = blockIdx.y; p < n_paths; p += gridDim.y) {
        const uint32_t a = off[p], b = off[p + 1];
        uint32_t run = B[(size_t)heavy[nodes[a]] * ldb + w];
        for (uint32_t q = a; q < b;) {
            int32_t lt[PATH_BATCH];
            int g = 0;
            bool more = true;
#pragma unroll
            for (int u = 0; u < PATH_BATCH; ++u) {
                lt[u] = q + u < b ? light[q + u] : -1;
                more = more && lt[u] >= 0;
                g += more ? 1 : 0;
            }
            uint32_t v[PATH_BATCH];
#pragma unroll
            for (int u = 0; u < PATH_BATCH; ++u)
                if (u < g) v[u] = B[(size_t)lt[u] * ldb + w];
#pragma unroll
            for (int u = 0; u < PATH_BATCH; ++u)
                if (u < g) {
                    run |= v[u];
                    B[(size_t)nodes[q + u] * ldb + w] = run;
                }
            q += g;
            if (g < PATH_BATCH && q < b) {
                const int32_t k = nodes[q], h = heavy[k];
                for (int32_t c = child_ptr[k]; c < child_ptr[k + 1]; ++c)
                    if (child_idx[c] != h) run |= B[(size_t)child_idx[c] * ldb + w];
                B[(size_t)k * ldb + w] = run;
                ++q;
            }
        }
    }
}

// Segmented pass over the processed rows: blockIdx.y owns rows [k0, k1), k0 a
// multiple of 32.  Per thread (sample s) and row r, e = E[knodes[r]][s]:
//   GATHER: U[r][s] = (float)e
//   SUMS:   Spart[seg][s] = sum_{r in seg} klen[r] * e', increasing r, where e'
//           is e as the pair kernels see it ((float)e when GATHER, else e)
//   MASKS:  bit (r % 32) of masks[panel][r / 32] = any e > 0 over the 32
//           samples of the panel (one wavefront = two panels)
// E holds the sample columns [col0, col0 + lde); U (row stride ldu) and the
// masks are indexed by absolute sample, Spart by column within the block.
template <bool GATHER, bool SUMS, bool MASKS>
__global__ void __launch_bounds__(EMB_THREADS) k_gather_sums(float* __restrict__ U, size_t ldu, uint32_t col0,
                                                             double* __restrict__ Spart,
                                                             uint32_t* __restrict__ masks, uint32_t Kw,
                                                             uint32_t* __restrict__ smask,
                                                             const double* __restrict__ E, size_t lde,
                                                             const int32_t* __restrict__ knodes,
                                                             const double* __restrict__ klen, uint32_t K,
                                                             uint32_t rows_per_seg, uint32_t n_valid,
                                                             const uint32_t* __restrict__ skip_u) {
    const size_t s = (size_t)blockIdx.x * EMB_THREADS + threadIdx.x;
    if (s >= lde) return;
    const bool real = s < n_valid;   // columns past the last sample never enter a mask
    const uint32_t k0 = blockIdx.y * rows_per_seg;
    const uint32_t k1 = min(K, k0 + rows_per_seg);
    const size_t panel = (col0 + s) / 64 * 2;   // the wave's lower 32-sample panel
    double acc = 0.0;
    uint32_t lo = 0, hi = 0;
#pragma unroll 8
    for (uint32_t r = k0; r < k1; ++r) {
        double e = E[(size_t)knodes[r] * lde + s];
        if constexpr (GATHER) {
            const float f = (float)e;
            if (!skip_u || !((skip_u[r >> 5] >> (r & 31)) & 1u)) U[(size_t)r * ldu + col0 + s] = f;
            e = (double)f;
        }
        if constexpr (SUMS) acc += klen[r] * e;
        if constexpr (MASKS) {
            const uint64_t b = __ballot(real && e > 0.0);
            lo |= ((uint32_t)b != 0u ? 1u : 0u) << (r & 31);
            hi |= ((uint32_t)(b >> 32) != 0u ? 1u : 0u) << (r & 31);
            if (smask && (threadIdx.x & 63) == 0) {   // the samples themselves, [panel][K]
                smask[panel * K + r] = (uint32_t)b;
                smask[(panel + 1) * K + r] = (uint32_t)(b >> 32);
            }
            if ((r & 31) == 31 || r + 1 == k1) {
                if ((threadIdx.x & 63) == 0) {
                    masks[panel * Kw + (r >> 5)] = lo;
                    masks[(panel + 1) * Kw + (r >> 5)] = hi;
                }
                lo = hi = 0;
            }
        }
    }
    if constexpr (SUMS) Spart[(size_t)blockIdx.y * lde + s] = acc;
}

// bits 0..15 of x to the even bits 0, 2, ..., 30
__device__ __forceinline__ uint32_t spread16(uint32_t x) {
    x &= 0xffffu;
    x = (x | (x << 8)) & 0x00ff00ffu;
    x = (x | (x << 4)) & 0x0f0f0f0fu;
    x = (x | (x << 2)) & 0x33333333u;
    return (x | (x << 1)) & 0x55555555u;
}

// k_gather_sums with two samples per thread: 16-byte loads of E (8-byte U
// stores), so a wave covers 128 samples, four 32-sample panels; the outputs,
// and the order of every sum, are those of k_gather_sums.  The samples' words
// (smask) of 32 rows are collected in two registers, lane j (j < 32) holding
// the word of the group's row j for panel 0 (2), lane 32 + j that for panel 1
// (3), and stored as one coalesced store per panel pair, not a one-lane store
// per row and panel.
// Needs lde, col0 and ldu even; with masks, lde and col0 multiples of 64, so a
// wave that runs past lde does so by exactly its upper two panels: all its
// lanes stay (for the stores), those past lde neither load nor store samples.
template <bool GATHER, bool SUMS, bool MASKS>
__global__ void __launch_bounds__(EMB_THREADS) k_gather_sums2(float* __restrict__ U, size_t ldu, uint32_t col0,
                                                              double* __restrict__ Spart,
                                                              uint32_t* __restrict__ masks, uint32_t Kw,
                                                              uint32_t* __restrict__ smask,
                                                              const double* __restrict__ E, size_t lde,
                                                              const int32_t* __restrict__ knodes,
                                                              const double* __restrict__ klen, uint32_t K,
                                                              uint32_t rows_per_seg, uint32_t n_valid,
                                                              const uint32_t* __restrict__ skip_u) {
    const size_t s = 2 * ((size_t)blockIdx.x * EMB_THREADS + threadIdx.x);   // samples s, s + 1
    const size_t wave0 = s & ~(size_t)127;            // the wave's first sample
    if (wave0 >= lde) return;                         // (whole waves only)
    const bool in = s < lde;
    const bool real0 = s < n_valid, real1 = s + 1 < n_valid;   // (n_valid <= lde)
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t k0 = blockIdx.y * rows_per_seg;
    const uint32_t k1 = min(K, k0 + rows_per_seg);
    const size_t panel = (col0 + wave0) / 32;         // the wave's first 32-sample panel
    const uint32_t npan = lde - wave0 >= 128 ? 4u : 2u;   // panels of the wave inside [0, lde)
    double acc0 = 0.0, acc1 = 0.0;
    uint32_t pb[4] = {0u, 0u, 0u, 0u};
    uint32_t w01 = 0u, w23 = 0u;
#pragma unroll 8
    for (uint32_t r = k0; r < k1; ++r) {
        double e0 = 0.0, e1 = 0.0;
        if (in) {
            const double2 e2 = *reinterpret_cast<const double2*>(E + (size_t)knodes[r] * lde + s);
            e0 = e2.x;
            e1 = e2.y;
        }
        if constexpr (GATHER) {
            const float f0 = (float)e0, f1 = (float)e1;
            if (in && (!skip_u || !((skip_u[r >> 5] >> (r & 31)) & 1u)))
                *reinterpret_cast<float2*>(U + (size_t)r * ldu + col0 + s) = make_float2(f0, f1);
            e0 = (double)f0;
            e1 = (double)f1;
        }
        if constexpr (SUMS) {
            acc0 += klen[r] * e0;
            acc1 += klen[r] * e1;
        }
        if constexpr (MASKS) {
            // lane l holds samples wave0 + 2l (b0) and wave0 + 2l + 1 (b1): panel q is lanes 16q .. 16q + 15
            const uint64_t b0 = __ballot(real0 && e0 > 0.0);
            const uint64_t b1 = __ballot(real1 && e1 > 0.0);
            const uint64_t bo = b0 | b1;
#pragma unroll
            for (uint32_t q = 0; q < 4; ++q)
                pb[q] |= (((bo >> (16 * q)) & 0xffffu) != 0u ? 1u : 0u) << (r & 31);
            if (smask && (lane & 31) == (r & 31)) {   // row r's words go to lanes r % 32 and 32 + r % 32
                const uint32_t sh = 16 * (lane >> 5);
                w01 = spread16((uint32_t)(b0 >> sh)) | (spread16((uint32_t)(b1 >> sh)) << 1);
                w23 = spread16((uint32_t)(b0 >> (32 + sh))) | (spread16((uint32_t)(b1 >> (32 + sh))) << 1);
            }
            if ((r & 31) == 31 || r + 1 == k1) {
                const uint32_t rc = r & ~31u, rows = (r & 31) + 1;   // this group's first row, its rows
                if (smask && (lane & 31) < rows) {
                    const uint32_t q = lane >> 5;
                    smask[(panel + q) * K + rc + (lane & 31)] = w01;
                    if (npan == 4) smask[(panel + 2 + q) * K + rc + (lane & 31)] = w23;
                }
                if (lane == 0)
                    for (uint32_t q = 0; q < npan; ++q) masks[(panel + q) * Kw + (r >> 5)] = pb[q];
#pragma unroll
                for (uint32_t q = 0; q < 4; ++q) pb[q] = 0u;
            }
        }
    }
    if constexpr (SUMS)
        if (in) *reinterpret_cast<double2*>(Spart + (size_t)blockIdx.y * lde + s) = make_double2(acc0, acc1);
}

__global__ void __launch_bounds__(EMB_THREADS) k_sum_segments(double* __restrict__ S, const double* __restrict__ Spart,
                                                              size_t ldn, uint32_t segs, uint32_t n) {
    const uint32_t s = blockIdx.x * EMB_THREADS + threadIdx.x;
    if (s >= n) return;
    double acc = 0.0;
#pragma unroll 16
    for (uint32_t g = 0; g < segs; ++g) acc += Spart[(size_t)g * ldn + s];   // (loads ahead, adds in order)
    S[s] = acc;
}

// Faith PD: identical FP64 summation order to src/unifrac.cpp:373-386. The
// chain of adds is inherently sequential per sample, so each thread issues
// FAITH_BATCH independent loads ahead of the ordered adds.  E holds n sample
// columns (row stride ldn); out points at the first of them.
constexpr uint32_t FAITH_BATCH = 16;
__global__ void __launch_bounds__(64) k_faith(double* __restrict__ out, const double* __restrict__ E, size_t ldn,
                                              const double* __restrict__ len, uint32_t m, uint32_t n) {
    const uint32_t s = blockIdx.x * 64 + threadIdx.x;
    if (s >= n) return;
    double acc = 0.0;
    uint32_t k = 0;
    for (; k + FAITH_BATCH <= m; k += FAITH_BATCH) {
        double v[FAITH_BATCH];
#pragma unroll
        for (uint32_t i = 0; i < FAITH_BATCH; ++i) v[i] = E[(size_t)(k + i) * ldn + s];
#pragma unroll
        for (uint32_t i = 0; i < FAITH_BATCH; ++i) acc += (double)(v[i] > 0.0) * len[k + i];
    }
    for (; k < m; ++k) acc += (double)(E[(size_t)k * ldn + s] > 0.0) * len[k];
    out[s] = acc;
}

// ------------------------------------------------------------------ presence bits
// B is node-major with ldb 32-bit words per row: bit j of B[k][w] says that
// sample 32 w + j holds node k.  With non-negative values a node is present
// exactly when one of its children is, so the rows are ORs up the tree.

// One workgroup per leaf: the row's bits from its CSR entries (value > 0, or
// every stored entry when `structural`), gathered in LDS, then stored whole.
__global__ void __launch_bounds__(256) k_bits_leaves(uint32_t* __restrict__ B, uint32_t ldb,
                                                     const int32_t* __restrict__ leaf_k,
                                                     const int32_t* __restrict__ leaf_row,
                                                     const uint32_t* __restrict__ indptr,
                                                     const uint32_t* __restrict__ indices,
                                                     const double* __restrict__ data, int structural) {
    extern __shared__ uint32_t bits_row[];
    for (uint32_t w = threadIdx.x; w < ldb; w += 256) bits_row[w] = 0u;
    __syncthreads();
    const int32_t r = leaf_row[blockIdx.x];
    if (r >= 0)
        for (uint32_t j = indptr[r] + threadIdx.x; j < indptr[r + 1]; j += 256) {
            const uint32_t s = indices[j];
            if (structural || data[j] > 0.0) atomicOr(&bits_row[s >> 5], 1u << (s & 31));
        }
    __syncthreads();
    uint32_t* dst = B + (size_t)leaf_k[blockIdx.x] * ldb;
    for (uint32_t w = threadIdx.x; w < ldb; w += 256) dst[w] = bits_row[w];
}

// Internal nodes of one height level: one thread per (node, word).
__global__ void __launch_bounds__(256) k_bits_level(uint32_t* __restrict__ B, uint32_t ldb,
                                                    const int32_t* __restrict__ nodes, uint32_t n_nodes,
                                                    const int32_t* __restrict__ child_ptr,
                                                    const int32_t* __restrict__ child_idx) {
    const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= (uint64_t)n_nodes * ldb) return;
    const uint32_t ni = (uint32_t)(i / ldb), w = (uint32_t)(i % ldb);
    const int32_t k = nodes[ni];
    uint32_t acc = 0u;
    for (int32_t c = child_ptr[k]; c < child_ptr[k + 1]; ++c) acc |= B[(size_t)child_idx[c] * ldb + w];
    B[(size_t)k * ldb + w] = acc;
}

// cnt[r] = samples holding knodes[r]: one workgroup per node
__global__ void __launch_bounds__(256) k_count_bits(uint32_t* __restrict__ cnt, const uint32_t* __restrict__ B,
                                                    uint32_t ldb, const int32_t* __restrict__ knodes) {
    __shared__ uint32_t part[4];
    const uint32_t* row = B + (size_t)knodes[blockIdx.x] * ldb;
    uint32_t c = 0;
    for (uint32_t w = threadIdx.x; w < ldb; w += 256) c += __popc(row[w]);
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) c += __shfl_xor(c, d, 64);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) cnt[blockIdx.x] = part[0] + part[1] + part[2] + part[3];
}

// smask[p][r] = B[nodes[r]][p] for the K listed nodes and panels p < npan:
// 64 nodes x 64 words per workgroup, transposed through LDS.
__global__ void __launch_bounds__(256) k_bits_smask(uint32_t* __restrict__ smask, const uint32_t* __restrict__ B,
                                                    uint32_t ldb, const int32_t* __restrict__ nodes, uint32_t K,
                                                    uint32_t npan) {
    __shared__ uint32_t tile[64][65];
    const uint32_t r0 = blockIdx.x * 64, p0 = blockIdx.y * 64, l = threadIdx.x & 63;
    for (uint32_t q = threadIdx.x >> 6; q < 64; q += 4) {
        const uint32_t r = r0 + q, p = p0 + l;
        tile[q][l] = (r < K && p < npan) ? B[(size_t)nodes[r] * ldb + p] : 0u;
    }
    __syncthreads();
    for (uint32_t q = threadIdx.x >> 6; q < 64; q += 4) {
        const uint32_t p = p0 + q, r = r0 + l;
        if (p < npan && r < K) smask[(size_t)p * K + r] = tile[l][q];
    }
}

// ------------------------------------------------------------------ Gram path

// 64 K'-columns x 64 samples per workgroup: read node rows (coalesced over
// samples), write sample rows (coalesced over columns).
// E holds the sample columns [col0, col0 + gridDim.y * 64) (row stride lde);
// Pt rows are absolute samples.
__global__ void __launch_bounds__(256) k_presence_T(int8_t* __restrict__ Pt, size_t ldp, const double* __restrict__ E,
                                                    size_t lde, uint32_t col0, const int32_t* __restrict__ col_node,
                                                    uint32_t n) {
    __shared__ uint8_t tile[64][64 + 4];
    const uint32_t c0 = blockIdx.x * 64, s0 = blockIdx.y * 64;
    const uint32_t t = threadIdx.x;
    const uint32_t ls = t & 63;
    for (uint32_t q = t >> 6; q < 64; q += 4) {
        const int32_t node = col_node[c0 + q];
        uint8_t v = 0;
        const uint32_t s = s0 + ls;
        if (node >= 0 && col0 + s < n) v = E[(size_t)node * lde + s] > 0.0 ? 0xFF : 0x00;
        tile[q][ls] = v;
    }
    __syncthreads();
    // each thread writes 16 consecutive columns of one sample row
    const uint32_t row = t >> 2, cb = (t & 3) * 16;
    uint32_t w[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        uint32_t x = 0;
#pragma unroll
        for (int b = 0; b < 4; ++b) x |= (uint32_t)tile[cb + q * 4 + b][row] << (8 * b);
        w[q] = x;
    }
    *reinterpret_cast<uint4*>(Pt + (size_t)(col0 + s0 + row) * ldp + c0 + cb) = make_uint4(w[0], w[1], w[2], w[3]);
}

// Pt from presence bits (every sample in B, n_pad rows): as k_presence_T.
__global__ void __launch_bounds__(256) k_presence_T_bits(int8_t* __restrict__ Pt, size_t ldp,
                                                         const uint32_t* __restrict__ B, uint32_t ldb,
                                                         const int32_t* __restrict__ col_node, uint32_t n) {
    __shared__ uint8_t tile[64][64 + 4];
    const uint32_t c0 = blockIdx.x * 64, s0 = blockIdx.y * 64;
    const uint32_t t = threadIdx.x;
    const uint32_t ls = t & 63;
    for (uint32_t q = t >> 6; q < 64; q += 4) {
        const int32_t node = col_node[c0 + q];
        const uint32_t s = s0 + ls;
        uint8_t v = 0;
        if (node >= 0 && s < n) v = ((B[(size_t)node * ldb + (s >> 5)] >> (s & 31)) & 1u) ? 0xFF : 0x00;
        tile[q][ls] = v;
    }
    __syncthreads();
    const uint32_t row = t >> 2, cb = (t & 3) * 16;
    uint32_t w[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        uint32_t x = 0;
#pragma unroll
        for (int b = 0; b < 4; ++b) x |= (uint32_t)tile[cb + q * 4 + b][row] << (8 * b);
        w[q] = x;
    }
    *reinterpret_cast<uint4*>(Pt + (size_t)(s0 + row) * ldp + c0 + cb) = make_uint4(w[0], w[1], w[2], w[3]);
}

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));

// G is kept in MFMA fragment order: tile (bi, bj) occupies 128*128 consecutive
// doubles; inside it, quadrant w (rows (w>>1)*64, columns (w&1)*64), block
// (mi, nj), register r and lane l hold element ((((w*2 + mi)*2 + nj)*16 + r)*64 + l).
// With the 32x32 C layout (lane l, register r -> row 8(r/4) + 4(l/32) + r%4,
// column l%32) this maps back to (i, j) as below.
__device__ __forceinline__ size_t gram_slot(uint32_t i, uint32_t j, uint32_t tiles) {
    const uint32_t bi = i / GRAM_BM, bj = j / GRAM_BM;
    const uint32_t il = i % GRAM_BM, jl = j % GRAM_BM;
    const uint32_t w = (il / 64) * 2 + (jl / 64);
    const uint32_t mi = (il % 64) / 32, nj = (jl % 64) / 32;
    const uint32_t r32 = il % 32, c32 = jl % 32;
    const uint32_t r = (r32 / 8) * 4 + (r32 % 4);
    const uint32_t lane = c32 + 32 * ((r32 / 4) % 2);
    return ((size_t)bi * tiles + bj) * (GRAM_BM * GRAM_BM) + ((((w * 2 + mi) * 2 + nj) * 16 + r) * 64 + lane);
}

constexpr int GLDA = GRAM_BK + 16;  // padded LDS row in bytes: ds_read_b128 stays conflict-free
constexpr int GRAM_THREADS = 512;

// Exact Gram on v_mfma_i32_32x32x32_i8.
//   A operand: presence of the row samples, bytes 0x00 / 0xFF (0 / -1)
//   B operand: presence of the column samples AND digit byte d_q of the column's length
// so each MFMA accumulates -sum_c d_q(c) [i present][j present] exactly in int32.
// Workgroup: 8 waves on a 128x128 output tile.  Wave w owns rows (w>>2)*64 ..
// +64 and columns (w&3)*32 .. +32 of the tile (two 32x32 MFMA blocks) for all
// four digits: 128 accumulator registers, and per K step three operand
// fragments for eight MFMAs.  At the end of each K-chunk the wave merges its
// four digits exactly, acc_0 + 2^7 acc_1 + 2^14 acc_2 + 2^21 acc_3 (below
// 2^45), and adds it times the chunk scale into its FP64 partial of the tile
// held in registers (one FMA per chunk and element); G is written once, in the
// fragment order of gram_slot.
__global__ void __launch_bounds__(GRAM_THREADS) __attribute__((amdgpu_waves_per_eu(2, 2)))
k_gram_i8r(double* __restrict__ G, uint32_t tiles, const int8_t* __restrict__ Pt, size_t ldp,
           const int8_t* __restrict__ digits, const GramChunk* __restrict__ chunks, uint32_t n_chunks,
           uint32_t n_stages, const uint32_t* __restrict__ order, uint32_t n_tiles, uint32_t per_xcd) {
    const uint32_t seq = (blockIdx.x % N_XCDS) * per_xcd + blockIdx.x / N_XCDS;
    if (seq >= n_tiles) return;
    const uint32_t bi = order[seq] >> 16, bj = order[seq] & 0xffffu;
    __shared__ __attribute__((aligned(16))) int8_t sA[GRAM_BM * GLDA];
    __shared__ __attribute__((aligned(16))) int8_t sB[GRAM_BM * GLDA];
    __shared__ __attribute__((aligned(16))) int8_t sD[4 * GRAM_BK];

    const uint32_t t = threadIdx.x;
    const uint32_t lane = t & 63, wave = t >> 6;
    const uint32_t rh = wave >> 2, cq = wave & 3;   // 64 rows, 32 columns
    const size_t Kp = ldp;

    static_assert(GRAM_BK == 128 && GRAM_THREADS == 512, "staging below assumes 128-column stages, 512 threads");
    const uint32_t lrow = t >> 3, lc16 = (t & 7) * 16;
    const int8_t* gA0 = Pt + ((size_t)bi * GRAM_BM + lrow) * ldp + lc16;
    const int8_t* gB0 = Pt + ((size_t)bj * GRAM_BM + lrow) * ldp + lc16;
    const size_t rstep = (size_t)64 * ldp;
    int8_t* sA0 = sA + lrow * GLDA + lc16;
    int8_t* sB0 = sB + lrow * GLDA + lc16;
    const int8_t* gD0 = digits + (size_t)lrow * Kp + lc16;   // digit plane lrow, for t < 32
    int8_t* sD0 = sD + lrow * GRAM_BK + lc16;
    uint4 ra0, ra1, rb0, rb1, rd;
#define SSU_GLOAD(stage)                                                              \
    do {                                                                              \
        const size_t k0_ = (size_t)(stage) * GRAM_BK;                                 \
        ra0 = *reinterpret_cast<const uint4*>(gA0 + k0_);                             \
        ra1 = *reinterpret_cast<const uint4*>(gA0 + rstep + k0_);                     \
        rb0 = *reinterpret_cast<const uint4*>(gB0 + k0_);                             \
        rb1 = *reinterpret_cast<const uint4*>(gB0 + rstep + k0_);                     \
        if (t < 32) rd = *reinterpret_cast<const uint4*>(gD0 + k0_);                  \
    } while (0)
#define SSU_SSTORE()                                                                  \
    do {                                                                              \
        *reinterpret_cast<uint4*>(sA0) = ra0;                                         \
        *reinterpret_cast<uint4*>(sA0 + 64 * GLDA) = ra1;                             \
        *reinterpret_cast<uint4*>(sB0) = rb0;                                         \
        *reinterpret_cast<uint4*>(sB0 + 64 * GLDA) = rb1;                             \
        if (t < 32) *reinterpret_cast<uint4*>(sD0) = rd;                              \
    } while (0)

    v16i acc[4][2];   // [digit][mi]
#pragma unroll
    for (int d = 0; d < 4; ++d)
#pragma unroll
        for (int mi = 0; mi < 2; ++mi) acc[d][mi] = (v16i){0};
    double gacc[2][16];   // the wave's FP64 partial of G, C-fragment order
#pragma unroll
    for (int mi = 0; mi < 2; ++mi)
#pragma unroll
        for (int r = 0; r < 16; ++r) gacc[mi][r] = 0.0;

    const uint32_t arow = rh * 64 + (lane & 31);
    const uint32_t brow = cq * 32 + (lane & 31);
    const uint32_t khalf = (lane >> 5) * 16;

    uint32_t ch = 0;
    uint32_t ch_end_stage = n_chunks ? chunks[0].k1 / GRAM_BK : 0;
    if (n_stages) { SSU_GLOAD(0); SSU_SSTORE(); }
    __syncthreads();
    for (uint32_t s = 0; s < n_stages; ++s) {
        if (s + 1 < n_stages) SSU_GLOAD(s + 1);
#pragma unroll
        for (int kk = 0; kk < GRAM_BK / 32; ++kk) {
            const uint32_t ko = kk * 32 + khalf;
            const v4i a0 = *reinterpret_cast<const v4i*>(sA + arow * GLDA + ko);
            const v4i a1 = *reinterpret_cast<const v4i*>(sA + (arow + 32) * GLDA + ko);
            const v4i pb = *reinterpret_cast<const v4i*>(sB + brow * GLDA + ko);
#pragma unroll
            for (int d = 0; d < 4; ++d) {
                const v4i b = pb & *reinterpret_cast<const v4i*>(sD + d * GRAM_BK + ko);
                acc[d][0] = __builtin_amdgcn_mfma_i32_32x32x32_i8(a0, b, acc[d][0], 0, 0, 0);
                acc[d][1] = __builtin_amdgcn_mfma_i32_32x32x32_i8(a1, b, acc[d][1], 0, 0, 0);
            }
        }
        if (s + 1 == ch_end_stage) {
            // exact merge of the four digits (|acc| <= 127 * 2^17), then one FMA per element
            const double sc = -chunks[ch].scale;   // A carries -1 for "present"
#pragma unroll
            for (int mi = 0; mi < 2; ++mi)
#pragma unroll
                for (int r = 0; r < 16; ++r) {
                    double v = fma((double)acc[3][mi][r], 128.0, (double)acc[2][mi][r]);
                    v = fma(v, 128.0, (double)acc[1][mi][r]);
                    v = fma(v, 128.0, (double)acc[0][mi][r]);
                    gacc[mi][r] = fma(v, sc, gacc[mi][r]);
                }
#pragma unroll
            for (int d = 0; d < 4; ++d)
#pragma unroll
                for (int mi = 0; mi < 2; ++mi) acc[d][mi] = (v16i){0};
            ++ch;
            if (ch < n_chunks) ch_end_stage = chunks[ch].k1 / GRAM_BK;
        }
        __syncthreads();
        if (s + 1 < n_stages) {
            SSU_SSTORE();
            __syncthreads();
        }
    }
#undef SSU_GLOAD
#undef SSU_SSTORE
    // quadrant rh*2 + cq/2, block (mi, cq&1) of the fragment order of gram_slot
    double* gl = G + ((size_t)bi * tiles + bj) * (GRAM_BM * GRAM_BM) + lane;
#pragma unroll
    for (int mi = 0; mi < 2; ++mi)
#pragma unroll
        for (int r = 0; r < 16; ++r)
            gl[(size_t)(((((rh * 2 + (cq >> 1)) * 2 + mi) * 2 + (cq & 1)) * 16 + r) * 64)] = gacc[mi][r];
}

template <class TOut>
__global__ void __launch_bounds__(256) k_unweighted_combine(TOut* __restrict__ out, const double* __restrict__ G,
                                                            uint32_t tiles, uint32_t n, int normalized, uint32_t row0,
                                                            uint32_t row1) {
    const uint32_t j = blockIdx.x * 64 + (threadIdx.x & 63);
    const uint32_t i = row0 + blockIdx.y * 4 + (threadIdx.x >> 6);
    if (i >= row1 || j >= n) return;
    double d = 0.0;
    if (i != j) {
        const uint32_t a = i < j ? i : j, b = i < j ? j : i;
        const double g = G[gram_slot(a, b, tiles)];
        const double ta = G[gram_slot(a, a, tiles)], tb = G[gram_slot(b, b, tiles)];
        const double x = (ta + tb) - 2.0 * g;   // length-weighted XOR
        const double o = (ta + tb) - g;         // length-weighted OR
        d = normalized ? x / o : x;
    }
    out[(size_t)i * n + j] = (TOut)d;
}

// The same distances for the pairs of rows [row0, row1) (multiples of 64) with
// every column from row0 on, each written at (i, j) and (j, i): one workgroup
// per 64 x 64 upper block (bi <= bj), the block transposed through LDS so that
// both stores are row-contiguous.  After it, rows [row0, row1) from column
// row0 on and columns [row0, row1) of the rows below are final.
template <class TOut>
__global__ void __launch_bounds__(256) k_unweighted_combine_band(TOut* __restrict__ out, const double* __restrict__ G,
                                                                 uint32_t tiles, uint32_t n, int normalized,
                                                                 uint32_t row0, uint32_t row1) {
    const uint32_t bi = row0 / 64 + blockIdx.y, bj = blockIdx.x;
    if (bj < bi || bi * 64 >= row1) return;   // (uniform)
    const uint32_t tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    __shared__ TOut tile[64][65];
    for (uint32_t r = ty; r < 64; r += 4) {
        const uint32_t i = bi * 64 + r, j = bj * 64 + tx;
        TOut v = (TOut)0;
        if (i < n && j < n && i != j) {
            const uint32_t a = i < j ? i : j, b = i < j ? j : i;
            const double g = G[gram_slot(a, b, tiles)];
            const double ta = G[gram_slot(a, a, tiles)], tb = G[gram_slot(b, b, tiles)];
            const double x = (ta + tb) - 2.0 * g;
            const double o = (ta + tb) - g;
            v = (TOut)(normalized ? x / o : x);
        }
        if (i < n && j < n) out[(size_t)i * n + j] = v;
        tile[r][tx] = v;
    }
    if (bi == bj) return;   // (uniform) a diagonal block holds both triangles itself
    __syncthreads();
    for (uint32_t c = ty; c < 64; c += 4) {
        const uint32_t j = bj * 64 + c, i = bi * 64 + tx;
        if (i < n && j < n) out[(size_t)j * n + i] = tile[tx][c];
    }
}

// ------------------------------------------------------------------ VALU pair tiles

// |x| through the IEEE fabs builtins, so the compiler folds it into a source modifier
__device__ __forceinline__ float dabs(float x) { return __builtin_fabsf(x); }
__device__ __forceinline__ double dabs(double x) { return __builtin_fabs(x); }

template <class T> __device__ __forceinline__ T dsqrt(T x);
template <> __device__ __forceinline__ float dsqrt<float>(float x) { return sqrtf(x); }
template <> __device__ __forceinline__ double dsqrt<double>(double x) { return sqrt(x); }
template <class T> __device__ __forceinline__ T dpow(T x, T a);
template <> __device__ __forceinline__ float dpow<float>(float x, float a) { return powf(x, a); }

// FP64 log2 of a positive normal x: x = m 2^e with m in [sqrt(1/2), sqrt(2)),
// log2 m = (2 / ln 2) atanh t with t = (m - 1) / (m + 1), |t| <= 0.1716, summed
// as t (1 + t^2/3 + t^4/5 + ... + t^24/25) (the next term is below 2e-20 of the
// sum): a few ulps, against the many branches of the general pow.
__device__ __forceinline__ double log2_normal(double x) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(x);
    int e = (int)((b >> 52) & 0x7ff) - 1022;
    double m = __longlong_as_double((long long)((b & 0x000fffffffffffffull) | 0x3fe0000000000000ull));  // [0.5, 1)
    if (m < 0.70710678118654752440) { m *= 2.0; --e; }
    const double t = (m - 1.0) / (m + 1.0), t2 = t * t;
    double p = 1.0 / 25.0;
    p = fma(p, t2, 1.0 / 23.0);
    p = fma(p, t2, 1.0 / 21.0);
    p = fma(p, t2, 1.0 / 19.0);
    p = fma(p, t2, 1.0 / 17.0);
    p = fma(p, t2, 1.0 / 15.0);
    p = fma(p, t2, 1.0 / 13.0);
    p = fma(p, t2, 1.0 / 11.0);
    p = fma(p, t2, 1.0 / 9.0);
    p = fma(p, t2, 1.0 / 7.0);
    p = fma(p, t2, 1.0 / 5.0);
    p = fma(p, t2, 1.0 / 3.0);
    p = fma(p, t2, 1.0);
    return (double)e + (t * p) * 2.88539008177792681472;   // 2 / ln 2
}

// FP64 2^y for |y| < 1000: y = k + f with k = rint(y), |f| <= 1/2, and
// 2^f = e^z, z = f ln 2, from the Taylor series to z^16/16! (|z| <= 0.347:
// the next term is below 1e-21), scaled by 2^k in the exponent bits.
__device__ __forceinline__ double exp2_finite(double y) {
    const double k = rint(y);
    const double z = (y - k) * 0.69314718055994530942;   // ln 2
    double p = 1.0 / 20922789888000.0;                    // 1/16!
    p = fma(p, z, 1.0 / 1307674368000.0);
    p = fma(p, z, 1.0 / 87178291200.0);
    p = fma(p, z, 1.0 / 6227020800.0);
    p = fma(p, z, 1.0 / 479001600.0);
    p = fma(p, z, 1.0 / 39916800.0);
    p = fma(p, z, 1.0 / 3628800.0);
    p = fma(p, z, 1.0 / 362880.0);
    p = fma(p, z, 1.0 / 40320.0);
    p = fma(p, z, 1.0 / 5040.0);
    p = fma(p, z, 1.0 / 720.0);
    p = fma(p, z, 1.0 / 120.0);
    p = fma(p, z, 1.0 / 24.0);
    p = fma(p, z, 1.0 / 6.0);
    p = fma(p, z, 0.5);
    p = fma(p, z, 1.0);
    p = fma(p, z, 1.0);
    return ldexp(p, (int)k);
}

// x^a for x > 0 in FP64: 2^(a log2 x) for the normal range, the general pow otherwise
template <> __device__ __forceinline__ double dpow<double>(double x, double a) {
    if (x >= 2.2250738585072014e-308 && x < 1.0e300) {
        const double y = a * log2_normal(x);
        if (y > -1000.0 && y < 1000.0) return exp2_finite(y);
    }
    return pow(x, a);
}

// 1 / sqrt(x) for normal x > 0 in FP64: the hardware FP64 reciprocal square
// root approximation (5.2e-8 relative at most, measured over 1e-30..1e30)
// refined by STEPS Newton steps r += r (1 - x r^2) / 2, each of which squares
// the relative error (one step: 4.1e-15; two: FP64 rounding).  Straight-line
// code, no branch.
template <int STEPS> __device__ __forceinline__ double rsqrt_nr(double x) {
    double r = __builtin_amdgcn_rsq(x);
#pragma unroll
    for (int i = 0; i < STEPS; ++i) {
        const double e = fma(-x * r, r, 1.0);
        r = fma(0.5 * r, e, r);
    }
    return r;
}

// Newton steps after the hardware start: FP64 values two (FP64 rounding);
// FP32 values none (the start's 5.2e-8 relative, measured on this GPU, is
// below the FP32 rounding the reference's FP32 terms carry)
template <class TV> constexpr int vaw_steps() { return sizeof(TV) == 8 ? 2 : 0; }

// x^alpha for x > 0 from the hardware log2 / exp2, as the generalized terms of
// the shared-node path use them (0 <= alpha <= 1 there): exp2(a log2 x) carries
// the rounding of a log2 x, |a log2 x| ulp, which the weights (x + y)^alpha
// and the ratios |u - v| / (u + v) keep below 1e-6 of a distance for alpha in
// [0, 1] (every FP32 test, 1.1e-7 to 2.3e-7 at 8,192 samples).
__device__ __forceinline__ float pow_alpha(float x, float alpha) {
    return __builtin_amdgcn_exp2f(alpha * __builtin_amdgcn_logf(x));
}

// x^a and x^(a - 1) for x > 0 with the exponent of x kept exact, for the dense
// kernel's alpha outside [0, 1], where terms of small x can carry a distance
// (alpha < 0) and exp2(a log2 x) is off by up to 2.4e-6 (a = 1.5, x near 1e-8;
// simulated with correctly rounded log2 / exp2).  x = m 2^e (m in [0.5, 1)),
// a e = n + r with n the nearest integer and r exact (an FMA gives the
// product's rounding error), x^a = 2^n exp2(r + a log2 m): exp2 sees
// |argument| <= 1/2 + |a| and ldexp applies 2^n exactly; x^(a-1) = x^a / x.
// About 1.3e-7 relative for x^a, 2.5e-7 for x^(a-1).
__device__ __forceinline__ void pow_alpha2_split(float x, float a, float& pa, float& pa1) {
    const float lm = __builtin_amdgcn_logf(__builtin_amdgcn_frexp_mantf(x));
    const float fe = (float)__builtin_amdgcn_frexp_expf(x);
    const float ae = a * fe;
    const float lo = __builtin_fmaf(a, fe, -ae);   // a e - ae, exactly
    const float nr = __builtin_rintf(ae);
    const float f = __builtin_fmaf(a, lm, (ae - nr) + lo);
    pa = __builtin_amdgcn_ldexpf(__builtin_amdgcn_exp2f(f), (int)nr);
    pa1 = pa * __builtin_amdgcn_rcpf(x);
}

// One-sided VAW term (src/unifrac_task_impl.hpp:481-494 with one count zero):
// a sample holding a node with count c > 0 and value term k (l u, or l for
// presence), e = max(T - c, 0) the rest of its total, against a sample that
// lacks the node, of total tau: mi = c, m - mi = e + tau, so
// k / sqrt(c (e + tau)), 0 where c (e + tau) is 0.  The one-sided sums and the
// posting walk call this same function on the same operands, so the terms the
// walk takes back out are bit-equal to those the sums hold.
template <int STEPS> __device__ __forceinline__ double vaw_one(double k, double c, double e, double tau) {
    const double x = c * (e + tau);
    const double r = rsqrt_nr<STEPS>(x);   // (x = 0: not finite, and not taken)
    return x > 0.0 ? __dmul_rn(k, r) : 0.0;
}

template <PairOp OP> struct PairTraits {
    static constexpr bool two = OP != PairOp::weighted;       // numerator and denominator accumulators
    static constexpr bool counts = OP == PairOp::vaw_weighted || OP == PairOp::vaw_generalized ||
                                   OP == PairOp::vaw_unweighted;
};

// Per-node term for one pair.  u, v: embedded values; cu, cv: embedded counts
// (VAW); m: c_k + c_l sample totals (VAW); l: branch length.  LPRE (VAW
// presence terms, and FP64 weighted terms; lengths >= 0): u and v arrive
// multiplied by l, so l |u - v| = |lu - lv| and l (u + v) = lu + lv
// (presence: l max(u, v) = max(lu, lv), exact for 0/1 values).
// PW (FP32 generalized powers): 0 the hardware exp2(a log2 s), 1 the exponent of s kept
// exact (pow_alpha2_split, alpha outside [0, 1]), 2 alpha = 1/2 by the hardware square
// root and reciprocal square root (1 ulp each, and one instruction fewer than 0)
template <PairOp OP, class T, bool LPRE = false, int PW = 0>
__device__ __forceinline__ void pair_term(T u, T v, T cu, T cv, T m, T l, T alpha, T& num, T& den) {
    if constexpr (OP == PairOp::weighted) {
        // src/unifrac_task_impl.hpp:117-127
        num = fma(dabs(u - v), l, num);
    } else if constexpr (OP == PairOp::unweighted_plain) {
        // presence is staged as 0/1 values: |pu - pv| = xor, max = or
        num = fma(dabs(u - v), l, num);
        den = fma(u > v ? u : v, l, den);
    } else if constexpr (OP == PairOp::generalized) {
        // src/unifrac_task_impl.hpp:1016-1024: w = s^a l, num += w |u-v| / s, den += w (s = u+v > 0)
        const T s = u + v;
        const bool nz = s != T(0);
        const T ss = nz ? s : T(1);
        if constexpr (sizeof(T) == 4) {
            // FP32: s^a and s^(a-1) as PW selects
            float pa, pa1;
            if constexpr (PW == 1) {
                pow_alpha2_split(ss, alpha, pa, pa1);
            } else if constexpr (PW == 2) {
                pa = __builtin_amdgcn_sqrtf(ss);
                pa1 = __builtin_amdgcn_rsqf(ss);
            } else {
                const float lg = __builtin_amdgcn_logf(ss);
                pa = __builtin_amdgcn_exp2f(alpha * lg);
                pa1 = __builtin_amdgcn_exp2f((alpha - 1.0f) * lg);
            }
            const float w = nz ? pa * l : 0.0f;
            const float q = nz ? pa1 * l : 0.0f;
            num = fma(q, dabs(u - v), num);
            den += w;
        } else {
            const T w = nz ? dpow<T>(ss, alpha) * l : T(0);
            num = fma(w, dabs(u - v) / ss, num);
            den += w;
        }
    } else {
        // src/unifrac_task_impl.hpp:881-893, 1098-1109, 1817-1829: terms divided by
        // sqrt(mi (m - mi)) where that is > 0 (FP32: one hardware reciprocal square root)
        const T mi = cu + cv;
        const T vv = mi * (m - mi);
        const bool ok = vv > T(0);
        T rv;
        if constexpr (sizeof(T) == 4) rv = ok ? __builtin_amdgcn_rsqf(vv) : 0.0f;
        else rv = ok ? T(1) / dsqrt<T>(vv) : T(0);
        if constexpr (OP == PairOp::vaw_weighted && LPRE) {
            num = fma(dabs(u - v), rv, num);
            den = fma(u + v, rv, den);
        } else if constexpr (OP == PairOp::vaw_weighted) {
            num = fma(dabs(u - v) * l, rv, num);
            den = fma((u + v) * l, rv, den);
        } else if constexpr (OP == PairOp::vaw_generalized) {
            const T s1 = (u + v) * rv, d1 = dabs(u - v) * rv;
            const bool on = ok && s1 != T(0);
            const T ss = on ? s1 : T(1);
            if constexpr (sizeof(T) == 4) {
                float pa, pa1;
                if constexpr (PW == 1) {
                    pow_alpha2_split(ss, alpha, pa, pa1);
                } else if constexpr (PW == 2) {
                    pa = __builtin_amdgcn_sqrtf(ss);
                    pa1 = __builtin_amdgcn_rsqf(ss);
                } else {
                    const float lg = __builtin_amdgcn_logf(ss);
                    pa = __builtin_amdgcn_exp2f(alpha * lg);
                    pa1 = __builtin_amdgcn_exp2f((alpha - 1.0f) * lg);
                }
                const float w = on ? pa * l : 0.0f;
                const float q = on ? pa1 * l : 0.0f;
                num = fma(q, d1, num);
                den += w;
            } else {
                const T w = on ? dpow<T>(ss, alpha) * l : T(0);
                num = fma(w, d1 / ss, num);
                den += w;
            }
        } else if constexpr (LPRE) {   // vaw_unweighted: u, v are l or 0
            num = fma(dabs(u - v), rv, num);
            den = fma(u > v ? u : v, rv, den);
        } else {  // vaw_unweighted: u, v are 0/1 presence values
            const T lv = l * rv;
            num = fma(dabs(u - v), lv, num);
            den = fma(u > v ? u : v, lv, den);
        }
    }
}

// How many nodes an FP32 accumulator may collect before it is merged into FP64.
constexpr int PAIR_MERGE = 64;
template <class T> struct PairStage { static constexpr int EB = sizeof(T) == 4 ? 16 : 8; };

// One workgroup per 128x128 tile of the pair matrix (upper-triangle tiles);
// thread (ti, tj) owns rows RI*ti.. and columns RJ*tj.. of the tile.  EB nodes
// of both 128-sample panels (and of the count panels for VAW) are staged in
// LDS per step; the next step's rows are prefetched into registers while the
// current one is consumed, and LDS is double-buffered (one barrier per step).
// add_partial (two-accumulator ops): out[i * n + j] and out2[i * n + j], i < j,
// hold the posting-list shares of the numerator and the denominator, added
// before the division; each pair is then written by its i < j thread alone.
template <PairOp OP, class T, int RI, int RJ, int TT, bool LPRE = false, int PW = 0>
__global__ void __launch_bounds__((TT / RI) * (TT / RJ))
k_pairs(const T* __restrict__ U, const T* __restrict__ Cn, const int32_t* __restrict__ rows, size_t ldn,
        const T* __restrict__ len, uint32_t K,
        uint32_t n, const double* __restrict__ S, const double* __restrict__ totals, double alpha_d, int normalized,
        void* __restrict__ out_v, int out_fp64, GapBand band, int add_partial, const void* __restrict__ out2_v,
        uint32_t* __restrict__ flags, float flag_tau) {
    constexpr int NTI = TT / RI, NTJ = TT / RJ, NT = NTI * NTJ;
    constexpr bool TWO = PairTraits<OP>::two;
    constexpr bool CNT = PairTraits<OP>::counts;
    constexpr bool MIXED = sizeof(T) == 4;  // FP32 terms, FP64 merges
    constexpr bool BIN = OP == PairOp::unweighted_plain || OP == PairOp::vaw_unweighted;
    constexpr int EB = PairStage<T>::EB;
    constexpr int VEC = 16 / sizeof(T);                      // elements per 16-byte load
    constexpr int NPANEL = CNT ? 4 : 2;                      // u_i, u_j (, c_i, c_j)
    constexpr int LOADS = NPANEL * EB * TT / VEC;     // 16-byte loads per step
    static_assert(LOADS % NT == 0, "staging must divide evenly");
    constexpr int LPT = LOADS / NT;                          // per thread
    typedef __attribute__((ext_vector_type(VEC))) T vecT;
    const uint32_t bi = blockIdx.y, bj = blockIdx.x;
    if (bi > bj || !band.meets_tile(bi * TT, bi * TT + TT, bj * TT, bj * TT + TT)) return;
    const uint32_t t = threadIdx.x;
    const uint32_t ti = t % NTI, tj = t / NTI;

    __shared__ __attribute__((aligned(16))) T sp[2][NPANEL][EB][TT];
    __shared__ T sl[2][EB];

    const T alpha = (T)alpha_d;
    const size_t i0 = (size_t)bi * TT, j0 = (size_t)bj * TT;
    T mrow[CNT ? RI : 1][CNT ? RJ : 1];  // VAW: c_i + c_j
    if constexpr (CNT) {
#pragma unroll
        for (int r = 0; r < RI; ++r)
#pragma unroll
            for (int c = 0; c < RJ; ++c) {
                const size_t i = i0 + ti * RI + r, j = j0 + tj * RJ + c;
                mrow[r][c] = (T)((i < n ? totals[i] : 0.0) + (j < n ? totals[j] : 0.0));
            }
    }

    T num[RI][RJ], den[TWO ? RI : 1][TWO ? RJ : 1];
    double num64[MIXED ? RI : 1][MIXED ? RJ : 1], den64[MIXED && TWO ? RI : 1][MIXED && TWO ? RJ : 1];
#pragma unroll
    for (int r = 0; r < RI; ++r)
#pragma unroll
        for (int c = 0; c < RJ; ++c) {
            num[r][c] = T(0);
            if constexpr (TWO) den[r][c] = T(0);
            if constexpr (MIXED) { num64[r][c] = 0.0; if constexpr (TWO) den64[r][c] = 0.0; }
        }

    vecT pre[LPT];
    T pre_l = T(0);
    T pre_lq[LPRE ? LPT : 1];   // LPRE: each load's node length
    // load step k0 into registers: load q -> panel, node e, 16-byte column group x
    auto fetch = [&](uint32_t k0) __attribute__((always_inline)) {
#pragma unroll
        for (int q = 0; q < LPT; ++q) {
            const uint32_t idx = t + (uint32_t)NT * q;
            const uint32_t panel = idx / (EB * TT / VEC);
            const uint32_t rem = idx % (EB * TT / VEC);
            const uint32_t e = rem / (TT / VEC), x = (rem % (TT / VEC)) * VEC;
            const uint32_t k = k0 + e;
            vecT v = (vecT)(T(0));
            if (k < K) {
                const size_t rk = (size_t)(rows ? rows[k] : (int32_t)k) * ldn;
                const T* src = ((panel & 2) ? Cn : U) + rk + ((panel & 1) ? j0 : i0) + x;
                v = *reinterpret_cast<const vecT*>(src);
            }
            pre[q] = v;
            if constexpr (LPRE) pre_lq[q] = k < K ? len[k] : T(0);
        }
        if (t < EB) pre_l = (k0 + t < K) ? len[k0 + t] : T(0);
    };
    auto stash = [&](int buf) __attribute__((always_inline)) {
#pragma unroll
        for (int q = 0; q < LPT; ++q) {
            const uint32_t idx = t + (uint32_t)NT * q;
            const uint32_t panel = idx / (EB * TT / VEC);
            const uint32_t rem = idx % (EB * TT / VEC);
            const uint32_t e = rem / (TT / VEC), x = (rem % (TT / VEC)) * VEC;
            vecT v = pre[q];
            if constexpr (BIN) {
                if (panel < 2) {
#pragma unroll
                    for (int z = 0; z < VEC; ++z) v[z] = v[z] > T(0) ? T(1) : T(0);
                }
            }
            if constexpr (LPRE) {
                if (panel < 2) {
#pragma unroll
                    for (int z = 0; z < VEC; ++z) v[z] *= pre_lq[q];
                }
            }
            *reinterpret_cast<vecT*>(&sp[buf][panel][e][x]) = v;
        }
        if (t < EB) sl[buf][t] = pre_l;
    };

    if (K) { fetch(0); stash(0); }
    __syncthreads();
    int buf = 0;
    for (uint32_t k0 = 0; k0 < K; k0 += EB) {
        const bool more = k0 + EB < K;
        if (more) fetch(k0 + EB);
#pragma unroll 2
        for (int e = 0; e < EB; ++e) {
            T u[RI], v[RJ], cu[RI], cv[RJ];
#pragma unroll
            for (int r = 0; r < RI; ++r) {
                u[r] = sp[buf][0][e][ti * RI + r];
                cu[r] = CNT ? sp[buf][CNT ? 2 : 0][e][ti * RI + r] : T(0);
            }
#pragma unroll
            for (int c = 0; c < RJ; ++c) {
                v[c] = sp[buf][1][e][tj * RJ + c];
                cv[c] = CNT ? sp[buf][CNT ? 3 : 1][e][tj * RJ + c] : T(0);
            }
            const T l = sl[buf][e];
            T dummy = T(0);
#pragma unroll
            for (int r = 0; r < RI; ++r)
#pragma unroll
                for (int c = 0; c < RJ; ++c) {
                    if constexpr (TWO)
                        pair_term<OP, T, LPRE, PW>(u[r], v[c], cu[r], cv[c], CNT ? mrow[r][c] : T(0), l, alpha,
                                                    num[r][c], den[r][c]);
                    else
                        pair_term<OP, T>(u[r], v[c], cu[r], cv[c], T(0), l, alpha, num[r][c], dummy);
                }
        }
        if constexpr (MIXED) {
            if (((k0 / EB + 1) % (PAIR_MERGE / EB)) == 0 || !more) {
#pragma unroll
                for (int r = 0; r < RI; ++r)
#pragma unroll
                    for (int c = 0; c < RJ; ++c) {
                        num64[r][c] += (double)num[r][c]; num[r][c] = T(0);
                        if constexpr (TWO) { den64[r][c] += (double)den[r][c]; den[r][c] = T(0); }
                    }
            }
        }
        if (more) stash(buf ^ 1);
        __syncthreads();
        buf ^= 1;
    }

#pragma unroll
    for (int r = 0; r < RI; ++r)
#pragma unroll
        for (int c = 0; c < RJ; ++c) {
            const size_t i = i0 + ti * RI + r, j = j0 + tj * RJ + c;
            if (i >= n || j >= n) continue;
            double nm, dn = 0.0;
            if constexpr (MIXED) { nm = num64[r][c]; if constexpr (TWO) dn = den64[r][c]; }
            else { nm = (double)num[r][c]; if constexpr (TWO) dn = (double)den[r][c]; }
            if constexpr (TWO) {
                if (add_partial) {
                    if (i > j) continue;   // (the pair's i < j thread writes both entries)
                    if (i < j) {
                        if (out_fp64) {
                            nm += ((const double*)out_v)[i * n + j];
                            dn += ((const double*)out2_v)[i * n + j];
                        } else {
                            nm += (double)((const float*)out_v)[i * n + j];
                            dn += (double)((const float*)out2_v)[i * n + j];
                        }
                        nm = nm < 0.0 ? 0.0 : nm;   // the shares' rounding can undershoot 0
                    }
                }
            }
            double d;
            if (i == j) d = 0.0;
            else if (!normalized) d = nm;
            else if (TWO) d = nm / dn;
            else d = nm / (S[i] + S[j]);
            if constexpr (OP == PairOp::vaw_weighted) {
                // unnormalised: near-duplicates (distance below flag_tau of the VAW
                // denominator) marked for the FP64 recomputation
                if (flags && !normalized && i < j && nm < (double)flag_tau * dn) {
                    flags[1 + i] = 1u;
                    flags[1 + j] = 1u;
                    atomicAdd(flags, 1u);
                }
            }
            if (out_fp64) {
                double* o = (double*)out_v;
                o[i * n + j] = d; o[j * n + i] = d;
            } else {
                float* o = (float*)out_v;
                o[i * n + j] = (float)d; o[j * n + i] = (float)d;
            }
        }
}

// ------------------------------------------------------ sparse weighted path
//
// Weighted UniFrac through the minimum:  |u - v| = u + v - 2 min(u, v), so
//   sum_k l_k |u_ik - u_jk| = S_i + S_j - 2 M_ij,   M_ij = sum_k l_k min(u_ik, u_jk),
// with S the per-sample sums.  min(u, v) is zero unless both samples hold the
// node, so a 64x64 block of pairs only visits the nodes present in both of its
// sample panels: the AND of the two panels' node masks.  One wavefront owns one
// block; each lane keeps an 8x8 register tile of FP64 accumulators (v_min_f64
// and v_fma_f64 issue at the same rate as their FP32 forms on this GPU).

// v_min_f64 without the NaN-quieting canonicalisation the compiler adds for
// operands it cannot see produced (here: LDS loads of staged values)
__device__ __forceinline__ double vmin_f64(double a, double b) {
    double r;
    asm("v_min_f64 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
    return r;
}

// Compacts the set bits of the wave's 64 mask words (lane l holds word l) into
// list[], in increasing order of l * 32 + bit; returns their number (uniform).
__device__ __forceinline__ uint32_t build_list(uint16_t* list, uint32_t word, uint32_t lane) {
    const uint32_t cnt = __popc(word);
    uint32_t incl = cnt;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t y = __shfl_up(incl, d, 64);
        if ((int)lane >= d) incl += y;
    }
    const uint32_t total = __builtin_amdgcn_readfirstlane(__shfl(incl, 63, 64));
    uint32_t o = incl - cnt;
    while (word) {
        const uint32_t b = __ffs(word) - 1;
        list[o++] = (uint16_t)(lane * 32 + b);
        word &= word - 1;
    }
    __syncthreads();
    return total;
}

constexpr uint32_t SP_B = 64;        // block edge in samples
constexpr uint32_t SP_WORDS = 64;    // mask words (32 nodes each) per list chunk
// Staging: the 8 values a lane reads for its rows (or columns) are contiguous;
// groups of 8 sit SP_GS doubles (80 bytes) apart so that the 8 distinct
// 16-byte reads of one ds_read_b128 pass fall in different LDS banks.
constexpr uint32_t SP_GS = 10;
constexpr uint32_t SP_HALF = 8 * SP_GS;   // offset of the column panel in a staged node
// nodes per step; two steps are in flight in registers while one is consumed
template <class TIn> struct SpStep { static constexpr uint32_t EB = sizeof(TIn) == 4 ? 8 : 4; };

// Blocks are numbered so that each XCD (workgroups go to the XCDs round robin)
// walks a contiguous run of `order` (upper_block_order): the blocks resident
// on one XCD then share their sample panels in that XCD's L2.

template <class TIn, class TOut>
__global__ void __launch_bounds__(64) k_pairs_sparse(const TIn* __restrict__ U, const int32_t* __restrict__ rows,
                                                     size_t ldn, const double* __restrict__ len,
                                                     const uint32_t* __restrict__ masks, uint32_t Kw, uint32_t n,
                                                     const double* __restrict__ S, int normalized,
                                                     TOut* __restrict__ out, uint32_t* __restrict__ visits,
                                                     const uint32_t* __restrict__ order, uint32_t n_blocks,
                                                     uint32_t per_xcd, const uint32_t* __restrict__ dmask,
                                                     int add_partial) {
    constexpr uint32_t EB = SpStep<TIn>::EB;
    const uint32_t seq = (blockIdx.x % N_XCDS) * per_xcd + blockIdx.x / N_XCDS;
    if (seq >= n_blocks) return;
    const uint32_t bij = order[seq];
    const uint32_t bi = bij >> 16, bj = bij & 0xffffu;
    uint32_t visited = 0;
    const uint32_t lane = threadIdx.x;
    const uint32_t li = lane >> 3, lj = lane & 7;

    __shared__ uint16_t list[SP_WORDS * 32];
    __shared__ __attribute__((aligned(16))) double sv[EB][2 * SP_HALF];
    __shared__ double sl[EB];

    double acc[8][8];
#pragma unroll
    for (int r = 0; r < 8; ++r)
#pragma unroll
        for (int c = 0; c < 8; ++c) acc[r][c] = 0.0;

    const uint32_t* mI = masks + (size_t)(2 * bi) * Kw;   // two 32-sample panels per side
    const uint32_t* mJ = masks + (size_t)(2 * bj) * Kw;
    const size_t offI = (size_t)bi * SP_B + lane, offJ = (size_t)bj * SP_B + lane;
    const uint32_t slot = (lane >> 3) * SP_GS + (lane & 7);   // where this lane stages its sample

    for (uint32_t w0 = 0; w0 < Kw; w0 += SP_WORDS) {
        // active nodes of this chunk, in increasing order, as offsets from w0*32
        const uint32_t w = w0 + lane;
        uint32_t word = w < Kw ? ((mI[w] | mI[Kw + w]) & (mJ[w] | mJ[Kw + w])) : 0u;
        if (dmask && w < Kw) word &= dmask[w];   // the other nodes come from the posting-list kernel
        const uint32_t total = build_list(list, word, lane);
        visited += total;
        const uint32_t base = w0 * 32;
        // lane e < EB resolves node t0 + e; its panel rows land in registers.
        // The loads are unconditional (row 0 stands in past the list end and
        // is masked out at staging), so the compiler can wait for one step's
        // loads while the other step's stay in flight.
        TIn pA[EB], qA[EB], pB[EB], qB[EB];
        double lA = 0.0, lB = 0.0;
        uint64_t vA = 0, vB = 0;   // which of the step's EB nodes are real
        auto fetch = [&](TIn (&p)[EB], TIn (&q)[EB], double& l, uint64_t& valid, uint32_t t0)
                         __attribute__((always_inline)) {
            const uint32_t idx = t0 + lane;
            const bool real = lane < EB && idx < total;
            int32_t row = 0;
            l = 0.0;
            if (real) {
                const uint32_t r = base + list[idx];
                // FP32 operand rows are the processed rows themselves; only FP64
                // operands read the embedding through rows[]
                if constexpr (sizeof(TIn) == 8) row = rows[r];
                else row = (int32_t)r;
                l = len[r];
            }
            valid = __ballot(real);
#pragma unroll
            for (uint32_t e = 0; e < EB; ++e) {
                const size_t off = (size_t)__builtin_amdgcn_readlane(row, e) * ldn;
                p[e] = U[off + offI];
                q[e] = U[off + offJ];
            }
        };
        auto step = [&](TIn (&p)[EB], TIn (&q)[EB], double& l, uint64_t& valid, uint32_t t_next)
                        __attribute__((always_inline)) {
#pragma unroll
            for (uint32_t e = 0; e < EB; ++e) {
                const bool on = (valid >> e) & 1;
                sv[e][slot] = on ? (double)p[e] : 0.0;
                sv[e][SP_HALF + slot] = on ? (double)q[e] : 0.0;
            }
            if (lane < EB) sl[lane] = l;
            __syncthreads();
            if (t_next < total) fetch(p, q, l, valid, t_next);   // two steps ahead of its use
#pragma unroll 2
            for (uint32_t e = 0; e < EB; ++e) {
                double a[8], b[8];
#pragma unroll
                for (int r = 0; r < 8; ++r) a[r] = sv[e][li * SP_GS + r];
#pragma unroll
                for (int c = 0; c < 8; ++c) b[c] = sv[e][SP_HALF + lj * SP_GS + c];
                const double ln = sl[e];
#pragma unroll
                for (int r = 0; r < 8; ++r)
#pragma unroll
                    for (int c = 0; c < 8; ++c) acc[r][c] = __builtin_fma(vmin_f64(a[r], b[c]), ln, acc[r][c]);
            }
            __syncthreads();
        };
        if (total) fetch(pA, qA, lA, vA, 0);
        if (total > EB) fetch(pB, qB, lB, vB, EB);
        for (uint32_t t0 = 0; t0 < total; t0 += 2 * EB) {
            step(pA, qA, lA, vA, t0 + 2 * EB);
            if (t0 + EB >= total) break;
            step(pB, qB, lB, vB, t0 + 3 * EB);
        }
    }

    if (visits && lane == 0) visits[(size_t)bi * ((n + SP_B - 1) / SP_B) + bj] = visited;
    const size_t i0 = (size_t)bi * SP_B + li * 8, j0 = (size_t)bj * SP_B + lj * 8;
#pragma unroll
    for (int r = 0; r < 8; ++r)
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            const size_t i = i0 + r, j = j0 + c;
            // the i < j lane writes both entries (so the partial it reads is its own)
            if (i >= n || j >= n || i > j) continue;
            if (i == j) { out[i * n + i] = (TOut)0; continue; }
            double m = acc[r][c];
            if (add_partial) m += (double)out[i * n + j];   // the sparse nodes' share, from k_pairs_rows
            const double tot = S[i] + S[j];
            double x = tot - 2.0 * m;
            x = x < 0.0 ? 0.0 : x;   // rounding can undershoot 0; NaN (empty samples) passes
            const double d = normalized ? x / tot : x;
            out[i * n + j] = (TOut)d;
            out[j * n + i] = (TOut)d;
        }
}

// FP32 operands: 32 x 64 blocks of pairs (rows: a 32-sample panel, columns:
// a 64-sample panel), one wavefront each, a 4 x 8 tile per lane.  The kernel
// sums l * |u - v| directly over the nodes either panel holds (the dense
// nodes, with the occupancy split): v_sub_f32 + v_fma_f32 with an |x| source
// modifier, summed in FP32 over S3_MERGE listed nodes and then merged into
// FP64, so the distance never comes from a difference of large sums and stays
// accurate for near-identical samples.  S only normalises.  Only pairs i < j
// are written (both (i, j) and (j, i)), so every pair has one writer even
// where blocks straddle the diagonal.
constexpr uint32_t S3_BI = 32, S3_BJ = 64;
constexpr uint32_t S3_EB = 8;          // nodes per step; two steps in flight
constexpr uint32_t S3_MERGE = 64;      // FP32 accumulation span, in listed nodes
constexpr uint32_t S3_GSJ = 12;        // column groups of 8 floats sit 48 bytes apart (bank-conflict free)
constexpr uint32_t S3_JOFF = 32;       // the 32 row values come first, contiguous
constexpr uint32_t S3_NODE = S3_JOFF + 8 * S3_GSJ;


// generalized (GEN) terms, FP32 as in the dense kernel: w = s^a l, q = s^(a-1) l
// from the hardware log2 / exp2 (s = u + v > 0), or, HALF (a = 1/2), from the
// hardware square root and reciprocal square root; num += q |u - v|, den += w
template <bool HALF = false>
__device__ __forceinline__ void gen_term(float u, float v, float l, float alpha, float& num, float& den) {
    const float s = u + v;
    const bool nz = s != 0.0f;
    float w, q;
    if constexpr (HALF) {
        const float ss = nz ? s : 1.0f;
        w = nz ? __builtin_amdgcn_sqrtf(ss) * l : 0.0f;
        q = nz ? __builtin_amdgcn_rsqf(ss) * l : 0.0f;
    } else {
        const float lg = __builtin_amdgcn_logf(nz ? s : 1.0f);
        w = nz ? __builtin_amdgcn_exp2f(alpha * lg) * l : 0.0f;
        q = nz ? __builtin_amdgcn_exp2f((alpha - 1.0f) * lg) * l : 0.0f;
    }
    num = __builtin_fmaf(q, __builtin_fabsf(u - v), num);
    den += w;
}
// FP64 as in the dense FP64 kernel: w = s^a l, num += w |u - v| / s, den += w
template <bool HALF = false>   // (FP64: dpow for every alpha)
__device__ __forceinline__ void gen_term(double u, double v, double l, double alpha, double& num, double& den) {
    const double s = u + v;
    const bool nz = s != 0.0;
    const double ss = nz ? s : 1.0;
    const double w = nz ? dpow<double>(ss, alpha) * l : 0.0;
    num = fma(w, __builtin_fabs(u - v) / ss, num);
    den += w;
}

// V: the operand type.  FP32 operands are the gathered rows U[r] of the
// processed nodes, summed in FP32 over S3_MERGE listed nodes and then merged
// into FP64; FP64 operands (generalized only) are the embedding rows
// U[rows[r]], summed in FP64.
template <class TOut, bool GEN, class V = float, bool HALF = false>
__device__ __forceinline__ void pairs_sparse32_body(const V* __restrict__ U, const int32_t* __restrict__ rows,
                                                       size_t ldn, const double* __restrict__ len,
                                                       const uint32_t* __restrict__ masks, uint32_t Kw, uint32_t n,
                                                       const double* __restrict__ S, int normalized,
                                                       TOut* __restrict__ out, uint32_t* __restrict__ visits,
                                                       const uint32_t* __restrict__ order, uint32_t n_blocks,
                                                       uint32_t per_xcd, const uint32_t* __restrict__ dmask,
                                                       int add_partial, V alpha, TOut* __restrict__ out2) {
    constexpr bool F64 = sizeof(V) == 8;
    static_assert(!F64 || GEN, "FP64 operands take the generalized terms only");
    constexpr uint32_t EB = S3_EB;
    const uint32_t seq = (blockIdx.x % N_XCDS) * per_xcd + blockIdx.x / N_XCDS;
    if (seq >= n_blocks) return;
    const uint32_t pi = order[seq] >> 16, pj = order[seq] & 0xffffu;   // 32-panel, 64-panel
    uint32_t visited = 0;
    const uint32_t lane = threadIdx.x;
    const uint32_t li = lane >> 3, lj = lane & 7;

    __shared__ uint16_t list[SP_WORDS * 32];
    __shared__ __attribute__((aligned(16))) V sv[EB][S3_NODE];
    __shared__ V sl[EB];

    V acc[4][8];
    double acc64[4][8];
    V accd[GEN ? 4 : 1][GEN ? 8 : 1];        // GEN: the denominator
    double accd64[GEN ? 4 : 1][GEN ? 8 : 1];
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            acc[r][c] = V(0);
            acc64[r][c] = 0.0;
            if constexpr (GEN) { accd[r][c] = V(0); accd64[r][c] = 0.0; }
        }
    uint32_t pending = 0;

    const uint32_t* mI = masks + (size_t)pi * Kw;
    const uint32_t* mJ = masks + (size_t)(2 * pj) * Kw;
    const size_t offI = (size_t)pi * S3_BI + (lane & 31), offJ = (size_t)pj * S3_BJ + lane;
    const uint32_t slotJ = S3_JOFF + (lane >> 3) * S3_GSJ + (lane & 7);

    for (uint32_t w0 = 0; w0 < Kw; w0 += SP_WORDS) {
        const uint32_t w = w0 + lane;
        uint32_t word = w < Kw ? (mI[w] | mJ[w] | mJ[Kw + w]) : 0u;
        if (dmask && w < Kw) word &= dmask[w];   // the other nodes come from the posting-list kernel
        const uint32_t total = build_list(list, word, lane);
        visited += total;
        const uint32_t base = w0 * 32;
        V pA[EB], qA[EB], pB[EB], qB[EB];
        V lA = V(0), lB = V(0);
        uint64_t vA = 0, vB = 0;
        auto fetch = [&](V (&p)[EB], V (&q)[EB], V& l, uint64_t& valid, uint32_t t0)
                         __attribute__((always_inline)) {
            const uint32_t idx = t0 + lane;
            const bool real = lane < EB && idx < total;
            int32_t row = 0;
            l = V(0);
            if (real) {
                const uint32_t r = base + list[idx];
                row = F64 ? rows[r] : (int32_t)r;
                l = (V)len[r];
            }
            valid = __ballot(real);
#pragma unroll
            for (uint32_t e = 0; e < EB; ++e) {
                const size_t off = (size_t)__builtin_amdgcn_readlane(row, e) * ldn;
                p[e] = U[off + offI];      // lanes 32..63 repeat lanes 0..31
                q[e] = U[off + offJ];
            }
        };
        auto step = [&](V (&p)[EB], V (&q)[EB], V& l, uint64_t& valid, uint32_t t_next)
                        __attribute__((always_inline)) {
#pragma unroll
            for (uint32_t e = 0; e < EB; ++e) {
                const bool on = (valid >> e) & 1;
                if (lane < S3_BI) sv[e][lane] = on ? p[e] : V(0);
                sv[e][slotJ] = on ? q[e] : V(0);
            }
            if (lane < EB) sl[lane] = l;
            __syncthreads();
            if (t_next < total) fetch(p, q, l, valid, t_next);
#pragma unroll 2
            for (uint32_t e = 0; e < EB; ++e) {
                V a[4], b[8];
#pragma unroll
                for (int r = 0; r < 4; ++r) a[r] = sv[e][li * 4 + r];
#pragma unroll
                for (int c = 0; c < 8; ++c) b[c] = sv[e][S3_JOFF + lj * S3_GSJ + c];
                const V ln = sl[e];
#pragma unroll
                for (int r = 0; r < 4; ++r)
#pragma unroll
                    for (int c = 0; c < 8; ++c) {
                        if constexpr (GEN)
                            gen_term<HALF>(a[r], b[c], ln, alpha, acc[r][c], accd[r][c]);
                        else
                            acc[r][c] = __builtin_fmaf(__builtin_fabsf(a[r] - b[c]), ln, acc[r][c]);
                    }
            }
            if constexpr (!F64) {
                pending += EB;
                if (pending >= S3_MERGE) {
#pragma unroll
                    for (int r = 0; r < 4; ++r)
#pragma unroll
                        for (int c = 0; c < 8; ++c) {
                            acc64[r][c] += (double)acc[r][c];
                            acc[r][c] = 0.0f;
                            if constexpr (GEN) { accd64[r][c] += (double)accd[r][c]; accd[r][c] = 0.0f; }
                        }
                    pending = 0;
                }
            }
            __syncthreads();
        };
        if (total) fetch(pA, qA, lA, vA, 0);
        if (total > EB) fetch(pB, qB, lB, vB, EB);
        for (uint32_t t0 = 0; t0 < total; t0 += 2 * EB) {
            step(pA, qA, lA, vA, t0 + 2 * EB);
            if (t0 + EB >= total) break;
            step(pB, qB, lB, vB, t0 + 3 * EB);
        }
    }

    if (visits && lane == 0) visits[(size_t)pi * ((n + S3_BJ - 1) / S3_BJ) + pj] = visited;
    const size_t i0 = (size_t)pi * S3_BI + li * 4, j0 = (size_t)pj * S3_BJ + lj * 8;
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            const size_t i = i0 + r, j = j0 + c;
            if (j >= n || i > j) continue;
            if (i == j) { out[i * n + i] = (TOut)0; continue; }
            double x = acc64[r][c] + (double)acc[r][c];
            if (add_partial) x += (double)out[i * n + j];   // the sparse nodes' share, from k_pairs_rows
            x = x < 0.0 ? 0.0 : x;   // the share's rounding can undershoot 0; NaN (empty samples) passes
            double d;
            if constexpr (GEN) {
                double y = accd64[r][c] + (double)accd[r][c];
                if (add_partial) y += (double)out2[i * n + j];
                d = x / y;   // 0 / 0 (no node held by either sample) gives NaN, as the dense kernel
            } else {
                d = normalized ? x / (S[i] + S[j]) : x;
            }
            out[i * n + j] = (TOut)d;
            out[j * n + i] = (TOut)d;
        }
}

template <class TOut>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(3))) k_pairs_sparse32(
        const float* __restrict__ U, size_t ldn, const double* __restrict__ len, const uint32_t* __restrict__ masks,
        uint32_t Kw, uint32_t n, const double* __restrict__ S, int normalized, TOut* __restrict__ out,
        uint32_t* __restrict__ visits, const uint32_t* __restrict__ order, uint32_t n_blocks, uint32_t per_xcd,
        const uint32_t* __restrict__ dmask, int add_partial) {
    pairs_sparse32_body<TOut, false>(U, nullptr, ldn, len, masks, Kw, n, S, normalized, out, visits, order, n_blocks,
                                     per_xcd, dmask, add_partial, 0.0f, nullptr);
}

// generalized: numerator and denominator summed directly (two accumulator sets, 2 waves per SIMD)
template <class TOut, bool HALF = false>
__global__ void __launch_bounds__(64) k_pairs_gen32(
        const float* __restrict__ U, size_t ldn, const double* __restrict__ len, const uint32_t* __restrict__ masks,
        uint32_t Kw, uint32_t n, TOut* __restrict__ out, TOut* __restrict__ out2, uint32_t* __restrict__ visits,
        const uint32_t* __restrict__ order, uint32_t n_blocks, uint32_t per_xcd, const uint32_t* __restrict__ dmask,
        int add_partial, float alpha) {
    pairs_sparse32_body<TOut, true, float, HALF>(U, nullptr, ldn, len, masks, Kw, n, nullptr, 1, out, visits, order, n_blocks,
                                    per_xcd, dmask, add_partial, alpha, out2);
}

// generalized on FP64 operands: the FP64 embedding rows, FP64 terms and sums (one wave per SIMD)
__global__ void __launch_bounds__(64) k_pairs_gen64(
        const double* __restrict__ U, const int32_t* __restrict__ rows, size_t ldn, const double* __restrict__ len,
        const uint32_t* __restrict__ masks, uint32_t Kw, uint32_t n, double* __restrict__ out,
        double* __restrict__ out2, uint32_t* __restrict__ visits, const uint32_t* __restrict__ order,
        uint32_t n_blocks, uint32_t per_xcd, const uint32_t* __restrict__ dmask, int add_partial, double alpha) {
    pairs_sparse32_body<double, true, double>(U, rows, ldn, len, masks, Kw, n, nullptr, 1, out, visits, order,
                                              n_blocks, per_xcd, dmask, add_partial, alpha, out2);
}

// --------------------------------------------- sparse nodes: posting lists

// Node occupancy split: a node held by at least `thr` samples is dense (its
// bit set in dmask, it stays with the tiled shared-node kernel); the rest go
// to the posting-list kernels, pcnt[r] = their sample counts (0 for dense
// nodes).  With `fixed` ([ceil(K/32)] words), the dense nodes are its set bits
// instead.  One thread per node; smask is [npan][K].
__global__ void __launch_bounds__(64) k_node_split(uint32_t* __restrict__ dmask, uint32_t* __restrict__ pcnt,
                                                   const uint32_t* __restrict__ smask, uint32_t K, uint32_t npan,
                                                   uint32_t thr, const uint32_t* __restrict__ fixed) {
    const uint32_t r = blockIdx.x * 64 + threadIdx.x;
    uint32_t cnt = 0;
    if (r < K)
        for (uint32_t p = 0; p < npan; ++p) cnt += __popc(smask[(size_t)p * K + r]);
    const bool dense = r < K && (fixed ? ((fixed[r >> 5] >> (r & 31)) & 1u) != 0u : cnt >= thr);
    if (r < K) pcnt[r] = dense ? 0u : cnt;
    const uint64_t b = __ballot(dense);
    const uint32_t w = blockIdx.x * 2 + (threadIdx.x >> 5);
    if ((threadIdx.x & 31) == 0 && w * 32 < K) dmask[w] = (threadIdx.x < 32) ? (uint32_t)b : (uint32_t)(b >> 32);
}

// Posting lists of the sparse nodes: node r's samples in increasing order
// (pidx, when non-null) with their values (pval, when non-null: row urows[r] of U, or row r), in slots
// [poff[r], poff[r + 1]); pc[p * K + r] = number of node r's samples below
// 32 p, for p = 0..npan.  A workgroup holds 64 nodes (lane = node, so the
// [panel][K] reads coalesce) in PST_WAVES waves, each wave a contiguous range
// of panels: the waves count their ranges' samples, an exclusive sum over the
// waves gives each range its first slot, then every wave writes its postings.
// Slots and counts stop at the node's capacity (never reached otherwise: the
// capacities are these same counts).  Values of type TU are stored as TV
// (FP64 embedding rows rounded to FP32 the way the gathered operand is).
constexpr uint32_t PST_WAVES = 8;
template <class IX, class TV, class TU>
__global__ void __launch_bounds__(64 * PST_WAVES) k_postings(IX* __restrict__ pidx, TV* __restrict__ pval,
                                                 uint32_t* __restrict__ pc, const uint32_t* __restrict__ smask,
                                                 const uint32_t* __restrict__ dmask, const TU* __restrict__ U,
                                                 const int32_t* __restrict__ urows, size_t ldu, uint32_t K,
                                                 uint32_t npan, const uint32_t* __restrict__ poff, uint32_t slots) {
    const uint32_t lane = threadIdx.x & 63, g = threadIdx.x >> 6;
    const uint32_t r = blockIdx.x * 64 + lane;
    __shared__ uint32_t cnt_s[PST_WAVES][64];
    // (the last condition: never, by the bound)
    const bool live = r < K && !((dmask[r >> 5] >> (r & 31)) & 1u) && poff[K] <= slots;
    const uint32_t per = (npan + PST_WAVES - 1) / PST_WAVES;
    const uint32_t p0 = min(npan, g * per), p1 = min(npan, p0 + per);
    uint32_t c = 0;
    if (live)
        for (uint32_t p = p0; p < p1; ++p) c += __popc(smask[(size_t)p * K + r]);
    cnt_s[g][lane] = c;
    __syncthreads();
    if (!live) return;
    uint32_t cnt = 0;   // samples of node r below panel p0
    for (uint32_t h = 0; h < g; ++h) cnt += cnt_s[h][lane];
    const TU* ur = U ? U + (size_t)(urows ? urows[r] : (int32_t)r) * ldu : nullptr;
    const uint32_t o = poff[r], cap = poff[r + 1] - o;
    IX* ix = pidx ? pidx + o : nullptr;
    TV* vx = pval ? pval + o : nullptr;
    for (uint32_t p = p0; p < p1; ++p) {
        pc[(size_t)p * K + r] = min(cnt, cap);
        uint32_t w = smask[(size_t)p * K + r];
        while (w && cnt < cap) {   // the slots hold every sample of the node; the bound guards memory
            const uint32_t s = p * 32 + (uint32_t)(__ffs(w) - 1);
            if (ix) ix[cnt] = (IX)s;
            if (vx) vx[cnt] = (TV)ur[s];
            ++cnt;
            w &= w - 1;
        }
        if (w) cnt += __popc(w);   // past the capacity: counted, not written
    }
    if (g == PST_WAVES - 1) pc[(size_t)npan * K + r] = min(cnt, cap);
}

// Fixed point for the LDS accumulators: 0 <= x < 2^62, truncated.  Integer
// adds are associative, so the row sums do not depend on the order in which
// lanes add their terms.
__device__ __forceinline__ unsigned long long to_fixed(double x) {
    const uint32_t hi = (uint32_t)(x * 0x1.0p-32);
    const uint32_t lo = (uint32_t)(x - (double)hi * 0x1.0p32);
    return ((unsigned long long)hi << 32) | lo;
}

// round(m * ls) as an integer, for 0 <= m * ls < 2^51: the FMA adds 2^52 with
// one rounding (to nearest, ties to even), which leaves the integer in the
// low 52 bits of the mantissa.
__device__ __forceinline__ unsigned long long fixed_term(double m, double ls) {
    const double y = fma(m, ls, 0x1.0p52);
    return (unsigned long long)__double_as_longlong(y) & 0x000FFFFFFFFFFFFFull;
}

// min of two values >= 0: for FP32 the smaller float has the smaller bit pattern
__device__ __forceinline__ double vmin_nonneg(float a, float b) {
    return (double)__uint_as_float(min(__float_as_uint(a), __float_as_uint(b)));
}
__device__ __forceinline__ double vmin_nonneg(double a, double b) { return a < b ? a : b; }

// Row lists of the sparse nodes: for sample a, the sparse nodes r that a holds,
// in increasing r, each as r << 32 | (a's index in r's posting list).  One
// wave per (32-sample panel P, chunk of kcw nodes) transposes the panel's
// bits with ballots, so the order is deterministic.  With rows == nullptr the
// wave writes its per-sample counts to counts[a * nkc + chunk]; otherwise it
// writes its entries from offsets[a * nkc + chunk] (the exclusive scan of the
// counts in that order).
// Chunks are RL_KC nodes wide unless that would make more than RL_MAX_CHUNKS
// chunks or n * chunks above RL_MAX_ENTRIES (the count and offset arrays):
// then the width doubles until both hold.
constexpr uint32_t RL_KC = 256, RL_MAX_CHUNKS = 65535;
constexpr size_t RL_MAX_ENTRIES = size_t(1) << 27;
__global__ void __launch_bounds__(64) k_row_lists(unsigned long long* __restrict__ rows, uint32_t* __restrict__ counts,
                                                  const uint32_t* __restrict__ offsets,
                                                  const uint32_t* __restrict__ smask,
                                                  const uint32_t* __restrict__ dmask, const uint32_t* __restrict__ pc,
                                                  uint32_t K, uint32_t n, uint32_t nkc, uint32_t kcw,
                                                  const uint32_t* __restrict__ poff, uint32_t slots) {
    const uint32_t P = blockIdx.x, kc = blockIdx.y, lane = threadIdx.x;
    if (rows && poff[K] > slots) return;   // more postings than slots: never, by the plan's bound
    const uint32_t a_lane = P * 32 + (lane & 31);
    uint32_t len = 0;   // lane s < 32: entries so far of sample 32 P + s in this chunk
    unsigned long long* dst =
        (rows && lane < 32 && a_lane < n) ? rows + offsets[(size_t)a_lane * nkc + kc] : nullptr;
    const uint32_t k_end = (uint32_t)min((uint64_t)K, (uint64_t)(kc + 1) * kcw);
    for (uint32_t k0 = kc * kcw; k0 < k_end; k0 += 64) {
        const uint32_t r = k0 + lane;
        uint32_t w = 0, base = 0;
        if (r < k_end && !((dmask[r >> 5] >> (r & 31)) & 1u)) {
            w = smask[(size_t)P * K + r];
            if (rows && w) base = pc[(size_t)P * K + r];
        }
        if (__ballot(w != 0u) == 0ull) continue;   // uniform
#pragma unroll 4
        for (uint32_t s = 0; s < 32; ++s) {
            const bool has = (w >> s) & 1u;
            const uint64_t m = __ballot(has);
            if (rows) {
                // sample 32 P + s's list and its length so far live in lane s (read with every lane active)
                const uint32_t len_s = __shfl(len, (int)s, 64);
                unsigned long long* d = (unsigned long long*)__shfl((long long)dst, (int)s, 64);
                if (has) {
                    const uint32_t at = len_s + __popcll(m & ((1ull << lane) - 1ull));
                    d[at] = ((unsigned long long)r << 32) | (base + __popc(w & ((1u << s) - 1u)));
                }
            }
            if (lane == s) len += (uint32_t)__popcll(m);
        }
    }
    if (!rows && lane < 32 && a_lane < n) counts[(size_t)a_lane * nkc + kc] = len;
}

// Exclusive prefix sum over m elements in three launches: per-block sums,
// a one-block scan of those, then each block's elements.  offsets[m] = total.
constexpr uint32_t SCAN_T = 256, SCAN_PER = 16, SCAN_BLOCK = SCAN_T * SCAN_PER;
__device__ __forceinline__ uint32_t block_exclusive(uint32_t v, uint32_t* part, uint32_t& total) {
    const uint32_t t = threadIdx.x;
    part[t] = v;
    __syncthreads();
    for (uint32_t d = 1; d < SCAN_T; d <<= 1) {
        const uint32_t y = t >= d ? part[t - d] : 0u;
        __syncthreads();
        part[t] += y;
        __syncthreads();
    }
    total = part[SCAN_T - 1];
    const uint32_t r = part[t] - v;
    __syncthreads();
    return r;
}

__global__ void __launch_bounds__(SCAN_T) k_scan_sums(uint32_t* __restrict__ bsum, const uint32_t* __restrict__ in,
                                                      size_t m) {
    __shared__ uint32_t part[SCAN_T];
    const size_t base = (size_t)blockIdx.x * SCAN_BLOCK;
    uint32_t v = 0;
    for (uint32_t i = 0; i < SCAN_PER; ++i) {
        const size_t q = base + (size_t)i * SCAN_T + threadIdx.x;
        if (q < m) v += in[q];
    }
    uint32_t total;
    (void)block_exclusive(v, part, total);
    if (threadIdx.x == 0) bsum[blockIdx.x] = total;
}

__global__ void __launch_bounds__(SCAN_T) k_scan_top(uint32_t* __restrict__ bsum, uint32_t nb) {
    __shared__ uint32_t part[SCAN_T];
    uint32_t carry = 0;
    for (uint32_t b0 = 0; b0 < nb; b0 += SCAN_T) {
        const uint32_t q = b0 + threadIdx.x;
        const uint32_t v = q < nb ? bsum[q] : 0u;
        uint32_t total;
        const uint32_t ex = block_exclusive(v, part, total);
        if (q < nb) bsum[q] = carry + ex;
        carry += total;
    }
    if (threadIdx.x == 0) bsum[nb] = carry;
}

__global__ void __launch_bounds__(SCAN_T) k_scan_apply(uint32_t* __restrict__ out, const uint32_t* __restrict__ in,
                                                       const uint32_t* __restrict__ bsum, size_t m, uint32_t nb) {
    __shared__ uint32_t part[SCAN_T];
    const size_t base = (size_t)blockIdx.x * SCAN_BLOCK;
    // each thread owns SCAN_PER consecutive elements
    const size_t q0 = base + (size_t)threadIdx.x * SCAN_PER;
    uint32_t v[SCAN_PER], sum = 0;
#pragma unroll
    for (uint32_t i = 0; i < SCAN_PER; ++i) {
        v[i] = q0 + i < m ? in[q0 + i] : 0u;
        sum += v[i];
    }
    uint32_t total;
    uint32_t run = bsum[blockIdx.x] + block_exclusive(sum, part, total);
#pragma unroll
    for (uint32_t i = 0; i < SCAN_PER; ++i)
        if (q0 + i < m) { out[q0 + i] = run; run += v[i]; }
    if (blockIdx.x == 0 && threadIdx.x == 0) out[m] = bsum[nb];
}

constexpr uint32_t SR_CW = 4096;   // output columns per workgroup (64-bit LDS accumulators, 32 KB)
constexpr uint32_t SR_WAVES = 8;   // waves sharing one accumulator, each taking every 8th round of entries
constexpr uint32_t SR_UNROLL = 4;  // postings per lane with their loads in flight together
constexpr uint32_t SR_WIN = 1024;  // walk positions resolved to entries at a time (64 lanes x 16 bytes)
constexpr uint32_t SR_WALK = 1, SR_DIAG = 2;   // k_pairs_rows modes: the columns past a, G(a, a)

// Sparse-node share for row a.  Weighted: M = sum over sparse nodes r held by
// both a and b > a of l_r * min(u_ra, u_rb); out[a * n + b] gets, for the
// tiled kernel's epilogue, M itself (FP64 values) or, for FP32 values, the
// sparse nodes' share of sum l |u_a - u_b|, ssp[a] + ssp[b] - 2 M, formed in
// FP64 (ssp: each sample's sum of l u over its sparse nodes).  GEN
// (generalized, FP32 values, 2048-column chunks): per shared sparse node the
// signed corrections l (s^(alpha-1) |u - v| - u^alpha - v^alpha) and
// l (s^alpha - u^alpha - v^alpha), s = u + v, summed into two accumulators;
// out and out2 get ssp[a] + ssp[b] plus each (ssp: sums of l u^alpha).  GRAM
// (unweighted): sum of l_r over the sparse nodes a and b >= a share, added to
// the FP64 Gram entry G(a, b) (the diagonal by the chunk that holds column a).
// One workgroup of 8 waves per (row a, column chunk c), blockIdx = c * n + a.  Row a's list
// is taken 64 entries at a time (round q goes to wave q % 8); a round's
// posting ranges inside the chunk (past a) are walked with the wave's 64 lanes
// spread evenly over the concatenated postings, each position's entry read
// from a per-wave byte map (entry starts marked, then a prefix max).  Terms
// are accumulated in fixed point, integers in units of 2^(e - 51) (FP32
// values: the FP64 product rounded to nearest by one FMA) or 2^(e - 61) (FP64
// weighted values: l_r * min, and GRAM: l_r, truncated) where 2^e bounds the
// row's sum of term magnitudes (S[a] weighted; 3 (ssp[a] + max ssp) GEN; the
// FP64 sum of the row list's lengths unweighted).  Weighted partial sums go out in the value type.  Integer
// adds make the result independent of the order in which waves and lanes add,
// and two samples holding the same nodes get bit-equal G(a, b) and G(a, a).
// Posting slots are addressed with 32-bit offsets (fewer than 2^30 slots).
// A round entry as the walk reads it: one 16-byte LDS read for FP32 values
template <class TV, bool GEN> struct alignas(16) RowEntry {
    uint32_t base;            // slot of walk position t: base + t (mod 2^32)
    TV lu;                    // the row's value in this node (weighted)
    unsigned long long q8;    // GRAM: the fixed-point term; weighted: l * scale (FP64 bits)
};
template <class TV> struct alignas(16) RowEntry<TV, true> {   // generalized
    uint32_t base;
    TV lu;                    // u_a
    TV uap;                   // u_a^alpha
    unsigned long long q8;    // l * scale (FP64 bits)
};

// the generalized terms' x^alpha in the value type: FP32 as above, FP64 as the FP64 kernels
// (FP32: alpha = 1/2 by the hardware square root, as the dense kernel's PW 2; the
// test is uniform, alpha being the launch's)
__device__ __forceinline__ float gen_pow(float x, double alpha) {
    return alpha == 0.5 ? __builtin_amdgcn_sqrtf(x) : pow_alpha(x, (float)alpha);
}
__device__ __forceinline__ double gen_pow(double x, double alpha) { return dpow<double>(x, alpha); }

// round(x * ls) as a signed integer (two's complement in the 64-bit
// accumulators) for |x * ls| < 2^51: the FMA adds 1.5 * 2^52, so the sum lies
// in [2^52, 2^53) where the FP64 spacing is 1 and its bits exceed those of
// 1.5 * 2^52 by exactly the rounded product
__device__ __forceinline__ unsigned long long fixed_term_signed(double x, double ls) {
    const double y = fma(x, ls, 0x1.8p52);
    return (unsigned long long)__double_as_longlong(y) - 0x4338000000000000ull;
}

// p[i] through a 32-bit byte offset from the (wave-uniform) base, so the load
// takes the SGPR-base form: i * sizeof(T) < 2^32 (slots < 2^30, T of 4 bytes at most)
template <class T> __device__ __forceinline__ T ld_slot(const T* __restrict__ p, uint32_t i) {
    if constexpr (sizeof(T) <= 4)
        return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(p) + (uint32_t)(i * (uint32_t)sizeof(T)));
    else
        return p[i];
}

// One posting's term added to the row's accumulators; accb is the accumulator
// array shifted so that accb[b] is column b (GEN: the denominator CW further on).
template <class TV, bool GRAM, bool GEN, bool FINE, uint32_t CW>
__device__ __forceinline__ void row_term(unsigned long long* accb, const RowEntry<TV, GEN>& en, uint32_t b, TV v,
                                         double alpha) {
    if constexpr (GRAM) {
        atomicAdd(&accb[b], en.q8);
    } else if constexpr (GEN && sizeof(TV) == 8) {
        // FP64 values: the same corrections in FP64, truncated to integers in units
        // of 2^-61 of the row's bound (|x * ls| < 2^61)
        const double uu = en.lu, vv = v, sm = uu + vv;
        const double p1 = dpow<double>(sm, alpha - 1.0);
        const double one = en.uap + dpow<double>(vv, alpha);
        const double ls = __longlong_as_double((long long)en.q8);
        const double cn = p1 * __builtin_fabs(uu - vv) - one;
        const double cd = p1 * sm - one;
        atomicAdd(&accb[b], (unsigned long long)(long long)(cn * ls));
        atomicAdd(&accb[CW + b], (unsigned long long)(long long)(cd * ls));
    } else if constexpr (GEN) {
        // a pair's correction to u^a + v^a (held by both): numerator and denominator
        const float af = (float)alpha;
        const float uu = (float)en.lu, vv = (float)v, sm = uu + vv;
        const float p1 = alpha == 0.5 ? __builtin_amdgcn_rsqf(sm)
                                      : __builtin_amdgcn_exp2f((af - 1.0f) * __builtin_amdgcn_logf(sm));
        const double one = (double)en.uap + (double)gen_pow(vv, alpha);   // (as the sums' terms)
        const double ls = __longlong_as_double((long long)en.q8);
        const double cn = (double)p1 * (double)__builtin_fabsf(uu - vv) - one;
        const double cd = (double)p1 * (double)sm - one;
        atomicAdd(&accb[b], fixed_term_signed(cn, ls));
        atomicAdd(&accb[CW + b], fixed_term_signed(cd, ls));
    } else {
        const double m = vmin_nonneg(en.lu, v);
        const double ls = __longlong_as_double((long long)en.q8);
        if constexpr (FINE)
            atomicAdd(&accb[b], to_fixed(m * ls));
        else
            atomicAdd(&accb[b], fixed_term(m, ls));
    }
}

// Generalized VAW one-sided term (src/unifrac_task_impl.hpp:1097-1109 with one
// count zero): sum1 = sub1 = u / vaw, so the numerator and the denominator
// both gain l (u / vaw)^alpha, the power in the value type (FP32: the
// hardware log2 / exp2; FP64: the series power); shared by the one-sided sums
// and the walk as vaw_one is.
template <int STEPS, class TV>
__device__ __forceinline__ double vaw_one_gen(double l, double u, double c, double e, double tau, double alpha) {
    const double x = c * (e + tau);
    const double r = rsqrt_nr<STEPS>(x);   // (x = 0: not finite, and not taken)
    return x > 0.0 ? __dmul_rn(l, (double)gen_pow((TV)(u * r), alpha)) : 0.0;
}

// VAW (1: weighted values, 2: presence, 3: generalized): one posting's
// corrections for the pair (a, b) holding node r, the both-held terms less the
// one-sided terms the sums H hold, numerator and denominator; en.lu = u_a,
// en.uap = c_a, en.q8 = l_r (FP64 bits); fixed point in units of 1 / scale,
// rounded (|x * scale| < 2^61 for FP64 values, 2^51 for FP32 values).
template <class TV, int VAW, uint32_t CW>
__device__ __forceinline__ void vaw_row_term(unsigned long long* accb, const RowEntry<TV, true>& en, uint32_t b,
                                             TV v, TV cv, double Ta, double Tb, double scale, double alpha) {
    constexpr int ST = vaw_steps<TV>();
    const double l = __longlong_as_double((long long)en.q8);
    const double ua = (double)en.lu, ca = (double)en.uap, ub = (double)v, cb = (double)cv;
    const double ea = Ta - ca > 0.0 ? Ta - ca : 0.0, eb = Tb - cb > 0.0 ? Tb - cb : 0.0;
    const double x = (ca + cb) * (ea + eb);
    const double r = rsqrt_nr<ST>(x);   // (x = 0: not finite, and not taken)
    const double rb = x > 0.0 ? r : 0.0;
    double one;
    if constexpr (VAW == 3)
        one = vaw_one_gen<ST, TV>(l, ua, ca, ea, Tb, alpha) + vaw_one_gen<ST, TV>(l, ub, cb, eb, Ta, alpha);
    else
        one = vaw_one<ST>(VAW == 2 ? l : l * ua, ca, ea, Tb) + vaw_one<ST>(VAW == 2 ? l : l * ub, cb, eb, Ta);
    double cn, cd;
    if constexpr (VAW == 2) {   // xor = 0, or = 1 for a node both hold
        cn = -one;
        cd = l * rb - one;
    } else if constexpr (VAW == 3) {   // w = l ((u + v) / vaw)^alpha, num += w |u - v| / (u + v)
        const double s1 = (ua + ub) * rb;
        const double w = x > 0.0 ? l * (double)gen_pow((TV)s1, alpha) : 0.0;
        const double q = x > 0.0 ? __builtin_fabs(ua - ub) / (ua + ub) : 0.0;
        cn = w * q - one;
        cd = w - one;
    } else {
        cn = l * __builtin_fabs(ua - ub) * rb - one;
        cd = l * (ua + ub) * rb - one;
    }
    if constexpr (sizeof(TV) == 8) {   // FP64 values: 61-bit units, rounded
        atomicAdd(&accb[b], (unsigned long long)(long long)rint(cn * scale));
        atomicAdd(&accb[CW + b], (unsigned long long)(long long)rint(cd * scale));
    } else {                           // FP32 values: 51-bit units, rounded by one FMA
        atomicAdd(&accb[b], fixed_term_signed(cn, scale));
        atomicAdd(&accb[CW + b], fixed_term_signed(cd, scale));
    }
}

// A pair's share formed as base2 (the two samples' one-sided or per-sample sums,
// >= 0) plus the walk's corrections, which take back the terms of the nodes both
// samples hold.  When the pair's true share is 0 (for VAW: every node both hold
// has mi = m, so vaw = 0 and the reference adds no term) what remains is the
// rounding of that cancellation, FP64 sums and fixed-point units, below 1e-11 of
// base2: shares within 2^-30 of base2 are 0, so that such a pair is 0 / 0 as in
// the reference.  (A true share that small would be that rounding's anyway.)
__device__ __forceinline__ double snap_share(double x, double base2) {
    return __builtin_fabs(x) <= 0x1p-30 * base2 ? 0.0 : x;
}

template <class TOut, class IX, bool GRAM, class TV, bool GEN, int VAW = 0>
__global__ void __launch_bounds__(64 * SR_WAVES) k_pairs_rows(TOut* __restrict__ out, uint32_t* __restrict__ walked,
                                                   const unsigned long long* __restrict__ rows,
                                                   const uint32_t* __restrict__ offsets,
                                                   const uint32_t* __restrict__ pc,
                                                   const IX* __restrict__ pidx, const TV* __restrict__ pval,
                                                   const double* __restrict__ len, const double* __restrict__ S,
                                                   uint32_t K, uint32_t n, uint32_t npan,
                                                   const uint32_t* __restrict__ poff, uint32_t slots,
                                                   uint32_t nchunks, uint32_t nkc, double* __restrict__ G,
                                                   uint32_t tiles, const double* __restrict__ ssp,
                                                   TOut* __restrict__ out2, const double* __restrict__ pspmax,
                                                   double alpha, GapBand band, uint32_t a0, uint32_t na,
                                                   uint32_t mode, const TV* __restrict__ pcv,
                                                   const double* __restrict__ tot, const double* __restrict__ Hm,
                                                   uint32_t D, const uint32_t* __restrict__ kidx) {
    // GEN and VAW keep two accumulators (numerator, denominator) on half-width chunks
    constexpr bool TWO = GEN || VAW != 0;
    constexpr uint32_t CW = TWO ? SR_CW / 2 : SR_CW;
    // chunk-major: the rows in flight share one column chunk of every posting
    // list, a quarter of the posting data at 16k samples (cache-resident).
    // Rows a0 .. a0 + na - 1; mode SR_DIAG alone: one workgroup per row, on the
    // chunk that holds its diagonal.
    const uint32_t a = a0 + blockIdx.x % na;
    const uint32_t c = mode == SR_DIAG ? a / CW : blockIdx.x / na;
    const bool walk = (mode & SR_WALK) != 0;
    const uint32_t b0 = c * CW, b1 = min(n, b0 + CW);
    // the selected columns b > a of this chunk: the hull [lo, hi) of the band's two gap
    // windows, widened to whole 32-sample panels (inside the chunk: b0 is a panel start)
    uint32_t lo = 0xffffffffu, hi = 0;
    {
        const uint64_t c_lo = max(b0, a + 1), c_hi = b1;
        const uint64_t w[2][2] = {{(uint64_t)a + band.a0, (uint64_t)a + band.a1 + 1},
                                  {(uint64_t)a + band.b0, (uint64_t)a + band.b1 + 1}};
        for (int k = 0; k < 2; ++k) {
            const uint64_t x = max(c_lo, w[k][0]), y = min(c_hi, w[k][1]);
            if (x < y) { lo = min(lo, (uint32_t)x); hi = max(hi, (uint32_t)y); }
        }
        if (lo < hi) {
            lo &= ~31u;
            hi = min(b1, (hi + 31u) & ~31u);
        }
        if (!walk) lo = 0xffffffffu, hi = 0;
    }
    const bool has_diag = GRAM && (mode & SR_DIAG) && b0 <= a && a < b1;
    // uniform: no selected column here (nor, unweighted, the diagonal); more postings
    // than slots never happens (the plan's bound) and is skipped, not read
    if (a >= n || (lo >= hi && !has_diag) || poff[K] > slots) {
        if (walked && walk && threadIdx.x == 0 && a < n) walked[(size_t)c * n + a] = 0;
        return;
    }
    const uint32_t lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    __shared__ unsigned long long acc[SR_CW];
    __shared__ uint32_t wk[SR_WAVES];
    using Ent = RowEntry<TV, TWO>;
    __shared__ Ent ent_s[SR_WAVES][65];   // per wave: the round's 64 entries at 1..64
    __shared__ double tot_s[VAW ? CW : 1];   // VAW: the chunk's sample totals
    // per entry: GRAM the fixed-point term, weighted l_r * scale (FP64 bits)
    __shared__ unsigned long long diag_s;
    __shared__ double bsum_s[SR_WAVES];
    __shared__ uint4 own_s[SR_WAVES][SR_WIN / 16];   // per wave: entry + 1 of every window position
    uint4* own4 = own_s[wv];
    uint8_t* own = reinterpret_cast<uint8_t*>(own4);
    Ent* ents = ent_s[wv];
    for (uint32_t q = threadIdx.x; q < SR_CW; q += 64 * SR_WAVES) acc[q] = 0ull;
    if constexpr (VAW != 0)
        for (uint32_t q = threadIdx.x; q < CW; q += 64 * SR_WAVES) tot_s[q] = b0 + q < n ? tot[b0 + q] : 0.0;
    const uint32_t e0 = offsets[(size_t)a * nkc], e1 = offsets[(size_t)(a + 1) * nkc];
    double bound;
    if constexpr (GRAM) {
        // FP64 sum of the row list's lengths, in a fixed order: per thread, then a
        // butterfly per wave, then the waves in order
        double x = 0.0;
        for (uint32_t e = e0 + threadIdx.x; e < e1; e += 64 * SR_WAVES) x += len[(uint32_t)(rows[e] >> 32)];
#pragma unroll
        for (int d = 32; d >= 1; d >>= 1) x += __shfl_xor(x, d, 64);
        if (lane == 0) bsum_s[wv] = x;
        if (threadIdx.x == 0) diag_s = 0ull;
        __syncthreads();
        bound = 0.0;
        for (uint32_t w = 0; w < SR_WAVES; ++w) bound += bsum_s[w];
    } else if constexpr (VAW != 0) {
        bound = ssp[a];   // the sum of the row's node bounds
        bound = bound < 1e300 ? bound : 1e300;   // (extreme inputs: coarser units, no overflow)
    } else if constexpr (GEN) {
        // |l (s^(a-1) |u - v| - u^a - v^a)| and |l (s^a - u^a - v^a)| are at most 3 l (u^a + v^a)
        bound = 3.0 * (ssp[a] + pspmax[0]);
    } else {
        bound = S[a];
    }
    __syncthreads();
    int e2 = 0;
    (void)frexp(bound, &e2);   // bound < 2^e2 bounds every partial sum of the row
    // FP32 values: 51-bit rounding (one FMA); FP64 values and GRAM: 61 bits (truncated;
    // VAW: rounded)
    constexpr bool FINE = GRAM || sizeof(TV) == 8;
    const int fb = FINE ? 61 : 51;
    const double scale = ldexp(1.0, fb - e2), inv = ldexp(1.0, e2 - fb);
    const double Ta = VAW != 0 ? tot[a] : 0.0;
    unsigned long long* const accb = acc - b0;   // accb[b]: column b (b0 <= b < b1)
    if (GRAM && has_diag) {   // G(a, a): the same fixed-point terms, every one of them
        unsigned long long x = 0ull;
        for (uint32_t e = e0 + threadIdx.x; e < e1; e += 64 * SR_WAVES)
            x += to_fixed(len[(uint32_t)(rows[e] >> 32)] * scale);
        if (x) atomicAdd(&diag_s, x);
    }
    const uint32_t pb0 = lo >> 5, pb1 = lo >= hi ? pb0 : (hi == n ? npan : (hi >> 5));
    uint32_t total_walked = 0;
    for (uint32_t r0 = e0 + 64 * wv; r0 < e1 && pb0 < pb1; r0 += 64 * SR_WAVES) {   // uniform per wave
        uint32_t start = 0, end = 0, r = 0, slot0 = 0;
        TV ua = (TV)0;
        TV uap = (TV)0;
        unsigned long long q = 0ull;
        if (r0 + lane < e1) {
            const unsigned long long ent = rows[r0 + lane];
            r = (uint32_t)(ent >> 32);
            const uint32_t pa = (uint32_t)ent;
            const uint32_t cs = pc[(size_t)pb0 * K + r], ce = pc[(size_t)pb1 * K + r];
            start = max(pa + 1, cs);
            end = ce;
            slot0 = poff[r];
            if constexpr (GRAM) {
                q = to_fixed(len[r] * scale);
            } else if constexpr (VAW != 0) {
                ua = pval[slot0 + pa];
                uap = pcv[slot0 + pa];   // c_a
                q = (unsigned long long)__double_as_longlong(len[r]);
            } else {
                ua = pval[slot0 + pa];
                q = (unsigned long long)__double_as_longlong(len[r] * scale);
                if constexpr (GEN) uap = gen_pow(ua, alpha);
            }
        }
        const uint32_t cnt = end > start ? end - start : 0u;
        uint32_t incl = cnt;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const uint32_t y = __shfl_up(incl, d, 64);
            if ((int)lane >= d) incl += y;
        }
        const uint32_t T = __builtin_amdgcn_readfirstlane(__shfl(incl, 63, 64));
        if (T == 0) continue;   // uniform per wave
        Ent me;
        me.base = slot0 + start - (incl - cnt);   // slot of walk position t: base + t (mod 2^32)
        me.lu = ua;
        me.q8 = q;
        if constexpr (TWO) me.uap = uap;
        ents[lane + 1] = me;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        // The walk positions [0, T) go SR_WIN at a time.  Each non-empty entry marks
        // its first position (entry + 1) in the wave's window bytes; a prefix max
        // (starts increase with the entry index) then names every position's
        // entry, so a lane finds its postings without searching.
        const uint32_t my_pre = incl - cnt;
        uint32_t carry = 0;   // entry + 1 owning the position before the window
        for (uint32_t w0 = 0; w0 < T; w0 += SR_WIN) {   // uniform per wave
            const uint32_t wn = min(SR_WIN, T - w0);
            own4[lane] = make_uint4(0u, 0u, 0u, 0u);
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            if (cnt > 0 && my_pre >= w0 && my_pre < w0 + SR_WIN) own[my_pre - w0] = (uint8_t)(lane + 1);
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            // lane l resolves bytes [16 l, 16 l + 16): running max, then a wave prefix max
            uint4 v4 = own4[lane];
            uint32_t wd[4] = {v4.x, v4.y, v4.z, v4.w};
            uint32_t run = 0;
#pragma unroll
            for (int i = 0; i < 16; ++i) run = max(run, (wd[i >> 2] >> (8 * (i & 3))) & 0xffu);
            uint32_t incl_max = run;
#pragma unroll
            for (int d = 1; d < 64; d <<= 1) {
                const uint32_t y = __shfl_up(incl_max, d, 64);
                if ((int)lane >= d) incl_max = max(incl_max, y);
            }
            uint32_t cur = max(carry, __shfl_up(incl_max, 1, 64));
            if (lane == 0) cur = carry;
#pragma unroll
            for (int i = 0; i < 16; ++i) {
                cur = max(cur, (wd[i >> 2] >> (8 * (i & 3))) & 0xffu);
                wd[i >> 2] = (wd[i >> 2] & ~(0xffu << (8 * (i & 3)))) | (cur << (8 * (i & 3)));
            }
            own4[lane] = make_uint4(wd[0], wd[1], wd[2], wd[3]);
            carry = max(carry, (uint32_t)__builtin_amdgcn_readlane((int)incl_max, 63));
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            // SR_UNROLL postings per lane in flight: entries and slots first, then
            // every load, then the adds.  A full window (every round but the last
            // window of each) has each lane take positions lane + 64 i, no guards.
            const uint32_t wb = w0 + lane;
            if (wn == SR_WIN) {   // uniform
#pragma unroll
                for (uint32_t g = 0; g < SR_WIN / 64; g += SR_UNROLL) {
                    Ent en[SR_UNROLL];
                    uint32_t bq[SR_UNROLL];
                    TV vq[SR_UNROLL], cq[VAW ? SR_UNROLL : 1];
#pragma unroll
                    for (uint32_t u = 0; u < SR_UNROLL; ++u) en[u] = ents[own[lane + 64 * (g + u)]];
#pragma unroll
                    for (uint32_t u = 0; u < SR_UNROLL; ++u) {
                        const uint32_t qs = en[u].base + wb + 64 * (g + u);
                        bq[u] = (uint32_t)ld_slot(pidx, qs);
                        if constexpr (!GRAM) vq[u] = ld_slot(pval, qs);
                        if constexpr (VAW != 0) cq[u] = ld_slot(pcv, qs);
                    }
#pragma unroll
                    for (uint32_t u = 0; u < SR_UNROLL; ++u) {
                        if constexpr (VAW != 0)
                            vaw_row_term<TV, VAW, CW>(accb, en[u], bq[u], vq[u], cq[u], Ta, tot_s[bq[u] - b0], scale,
                                                      alpha);
                        else
                            row_term<TV, GRAM, GEN, FINE, CW>(accb, en[u], bq[u], vq[u], alpha);
                    }
                }
            } else {
                for (uint32_t j0 = lane; j0 < wn; j0 += 64 * SR_UNROLL) {
                    Ent en[SR_UNROLL];
                    uint32_t bq[SR_UNROLL];
                    TV vq[SR_UNROLL], cq[VAW ? SR_UNROLL : 1];
#pragma unroll
                    for (uint32_t u = 0; u < SR_UNROLL; ++u) {
                        const uint32_t j = min(j0 + 64 * u, wn - 1);
                        en[u] = ents[own[j]];
                        const uint32_t qs = en[u].base + w0 + j;
                        bq[u] = (uint32_t)ld_slot(pidx, qs);
                        if constexpr (!GRAM) vq[u] = ld_slot(pval, qs);
                        if constexpr (VAW != 0) cq[u] = ld_slot(pcv, qs);
                    }
#pragma unroll
                    for (uint32_t u = 0; u < SR_UNROLL; ++u) {
                        if (j0 + 64 * u >= wn) continue;
                        if constexpr (VAW != 0)
                            vaw_row_term<TV, VAW, CW>(accb, en[u], bq[u], vq[u], cq[u], Ta, tot_s[bq[u] - b0], scale,
                                                      alpha);
                        else
                            row_term<TV, GRAM, GEN, FINE, CW>(accb, en[u], bq[u], vq[u], alpha);
                    }
                }
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }
        total_walked += T;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
    if (lane == 0) wk[wv] = total_walked;
    __syncthreads();
    if (walked && walk && threadIdx.x == 0) {
        uint32_t sum = 0;
        for (uint32_t w = 0; w < SR_WAVES; ++w) sum += wk[w];
        walked[(size_t)c * n + a] = sum;
    }
    // the walked panels' columns past the diagonal (every posting of theirs was added)
    const uint32_t wlo = max(a + 1, pb0 * 32), whi = min(hi, pb1 * 32);
    if constexpr (GRAM) {
        if (has_diag && threadIdx.x == 0 && diag_s) G[gram_slot(a, a, tiles)] += (double)diag_s * inv;
        for (uint32_t b = wlo + threadIdx.x; b < whi; b += 64 * SR_WAVES) {
            const unsigned long long v = acc[b - b0];
            if (v) G[gram_slot(a, b, tiles)] += (double)v * inv;
        }
    } else if constexpr (VAW != 0) {
        // numerator and denominator shares: the one-sided sums of both samples at
        // the other's total plus the corrections of the shared nodes
        const uint32_t ka = kidx ? kidx[a] : a;
        for (uint32_t b = wlo + threadIdx.x; b < whi; b += 64 * SR_WAVES) {
            const uint32_t kb = kidx ? kidx[b] : b;
            const double base2 = Hm[(size_t)a * D + kb] + Hm[(size_t)b * D + ka];
            out[(size_t)a * n + b] = (TOut)snap_share(base2 + (double)(long long)acc[b - b0] * inv, base2);
            out2[(size_t)a * n + b] = (TOut)snap_share(base2 + (double)(long long)acc[CW + b - b0] * inv, base2);
        }
    } else if constexpr (GEN) {
        // numerator and denominator shares: P_a + P_b plus the corrections of the
        // shared nodes (ssp = per-sample sum of l u^a over the sparse nodes)
        const double sa = ssp[a];
        for (uint32_t b = wlo + threadIdx.x; b < whi; b += 64 * SR_WAVES) {
            const double base2 = sa + ssp[b];
            out[(size_t)a * n + b] = (TOut)snap_share(base2 + (double)(long long)acc[b - b0] * inv, base2);
            out2[(size_t)a * n + b] = (TOut)snap_share(base2 + (double)(long long)acc[CW + b - b0] * inv, base2);
        }
    } else if constexpr (sizeof(TV) == 4) {
        // FP32 values: the sparse nodes' share of sum l |u_a - u_b| itself, formed in
        // FP64 before the rounding to FP32
        const double sa = ssp[a];
        for (uint32_t b = wlo + threadIdx.x; b < whi; b += 64 * SR_WAVES)
            out[(size_t)a * n + b] = (TOut)((sa + ssp[b]) - 2.0 * ((double)acc[b - b0] * inv));
    } else {
        for (uint32_t b = wlo + threadIdx.x; b < whi; b += 64 * SR_WAVES)
            out[(size_t)a * n + b] = (TOut)((double)acc[b - b0] * inv);
    }
}

// ssp[a] = sum over sample a's row list (its sparse nodes r, increasing) of
// len[r] * value of a in r's postings, in FP64 in a fixed order: one wavefront
// per sample, lane-strided partial sums then a butterfly.
template <class TV, bool GEN>
__global__ void __launch_bounds__(64) k_sparse_sums(double* __restrict__ ssp,
                                                    const unsigned long long* __restrict__ rows,
                                                    const uint32_t* __restrict__ offsets, uint32_t nkc,
                                                    const uint32_t* __restrict__ poff, const TV* __restrict__ pval,
                                                    const double* __restrict__ len, uint32_t K, uint32_t n,
                                                    uint32_t slots, double alpha) {
    const uint32_t a = blockIdx.x, lane = threadIdx.x;
    if (a >= n) return;
    double x = 0.0;
    if (poff[K] <= slots) {   // (never otherwise, by the plan's bound)
        const uint32_t e0 = offsets[(size_t)a * nkc], e1 = offsets[(size_t)(a + 1) * nkc];
        for (uint32_t e = e0 + lane; e < e1; e += 64) {
            const unsigned long long ent = rows[e];
            const uint32_t r = (uint32_t)(ent >> 32);
            const TV v = pval[poff[r] + (uint32_t)ent];
            if constexpr (GEN)
                x += len[r] * (double)gen_pow(v, alpha);   // (GEN: l u^a)
            else
                x += len[r] * (double)v;
        }
    }
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) x += __shfl_xor(x, d, 64);
    if (lane == 0) ssp[a] = x;
}

// ------------------------------------------------ VAW on the shared-node path
//
// VAW terms (src/unifrac_task_impl.hpp:481-494, 897-911, 1097-1109,
// 1817-1829) divide each node's term by sqrt(mi (m - mi)), mi = c_a + c_b,
// m = T_a + T_b, so a node held by a alone still adds
// l u_a / sqrt(c_a (T_a - c_a + T_b)) (generalized: l (u_a / sqrt(...))^alpha),
// which depends on b's total.  Over the sparse nodes the pair's share is therefore
//   H(a, T_b) + H(b, T_a) + sum over the nodes both hold of (both-term - h_a - h_b),
// H(a, t) = sum over a's sparse nodes of the one-sided term at total t (taken
// at each distinct total t once), the sum over shared nodes from the posting
// walk.  m - mi is taken as e_a + e_b, e = max(T - c, 0) (equal for integer
// counts; it keeps FP32-rounded counts from producing a tiny positive m - mi).

// Per sparse node r, a bound on any one correction the walk adds for a pair
// holding r (numerator or denominator): |both-term - h_a - h_b| <= den_both +
// h_a + h_b, den_both = l (u_a + u_b) / sqrt((c_a + c_b)(e_a + e_b)) and
// h = l u / sqrt(c (e + T')), so with, over r's postings, umax the largest
// value (1 for presence), cmin the smallest positive count, emin the smallest
// positive e, tmin the smallest positive total, g = min(emin, tmin):
//   B_r = l umax (2 / sqrt(cmin emin) + 2 / sqrt(cmin g)).
// Generalized (GEN): the terms are l s^alpha with s at most the quotients above,
// so B_r = l ((2 umax / sqrt(cmin emin))^alpha + 2 (umax / sqrt(cmin g))^alpha).
// One wavefront per node: lane-strided maxima and minima over the postings,
// then a butterfly.
template <class IX, class TV, bool UNW, bool GEN>
__global__ void __launch_bounds__(64) k_vaw_node_bound(double* __restrict__ nb, const IX* __restrict__ pidx,
                                                       const TV* __restrict__ pval, const TV* __restrict__ pcv,
                                                       const uint32_t* __restrict__ poff,
                                                       const double* __restrict__ tot, const double* __restrict__ len,
                                                       uint32_t K, uint32_t slots, double alpha) {
    const uint32_t r = blockIdx.x, lane = threadIdx.x;
    if (r >= K) return;
    constexpr double INF = __builtin_huge_val();
    double umax = 0.0, cmin = INF, emin = INF, tmin = INF;
    if (poff[K] <= slots) {   // (never otherwise, by the plan's bound)
        for (uint32_t q = poff[r] + lane; q < poff[r + 1]; q += 64) {
            const double t = tot[pidx[q]], c = (double)pcv[q], u = (double)pval[q];
            const double e = t - c > 0.0 ? t - c : 0.0;
            umax = u > umax ? u : umax;
            if (c > 0.0) cmin = c < cmin ? c : cmin;
            if (e > 0.0) emin = e < emin ? e : emin;
            if (t > 0.0) tmin = t < tmin ? t : tmin;
        }
    }
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) {
        umax = fmax(umax, __shfl_xor(umax, d, 64));
        cmin = fmin(cmin, __shfl_xor(cmin, d, 64));
        emin = fmin(emin, __shfl_xor(emin, d, 64));
        tmin = fmin(tmin, __shfl_xor(tmin, d, 64));
    }
    if (lane) return;
    double b = 0.0;
    if (cmin < INF) {
        const double g = emin < tmin ? emin : tmin;
        const double uu = UNW ? 1.0 : umax;
        if constexpr (GEN)
            b = len[r] * ((emin < INF ? pow(2.0 * uu / sqrt(cmin * emin), alpha) : 0.0) +
                          (g < INF ? 2.0 * pow(uu / sqrt(cmin * g), alpha) : 0.0));
        else
            b = len[r] * uu * ((emin < INF ? 2.0 / sqrt(cmin * emin) : 0.0) + (g < INF ? 2.0 / sqrt(cmin * g) : 0.0));
    }
    nb[r] = b;
}

// out[a] = sum of w[r] over sample a's row list (the sparse nodes it holds), in
// FP64 in a fixed order: one wavefront per sample, lane-strided partial sums
// then a butterfly.
__global__ void __launch_bounds__(64) k_row_weights(double* __restrict__ out, const unsigned long long* __restrict__ rows,
                                                    const uint32_t* __restrict__ offsets, uint32_t nkc,
                                                    const double* __restrict__ w, const uint32_t* __restrict__ poff,
                                                    uint32_t K, uint32_t n, uint32_t slots) {
    const uint32_t a = blockIdx.x, lane = threadIdx.x;
    if (a >= n) return;
    double x = 0.0;
    if (poff[K] <= slots) {
        const uint32_t e0 = offsets[(size_t)a * nkc], e1 = offsets[(size_t)(a + 1) * nkc];
        for (uint32_t e = e0 + lane; e < e1; e += 64) x += w[(uint32_t)(rows[e] >> 32)];
    }
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) x += __shfl_xor(x, d, 64);
    if (lane == 0) out[a] = x;
}

// One-sided sums H[a][k] = sum over sample a's sparse nodes r (its row list, in
// order) of vaw_one(l_r u_ar, c_ar, max(T_a - c_ar, 0), tau[k]) (VAW 2,
// presence: l_r in place of l_r u_ar; VAW 3, generalized: vaw_one_gen), in
// FP64.  One wavefront per (sample a, 256 totals): the row's entries are
// staged 64 at a time in LDS and read by every lane, each lane summing
// VH_COLS totals.
constexpr uint32_t VH_COLS = 4;
template <class TV, int VAW>
__global__ void __launch_bounds__(64) k_vaw_onesided(double* __restrict__ H, uint32_t D, const double* __restrict__ tau,
                                                     const unsigned long long* __restrict__ rows,
                                                     const uint32_t* __restrict__ offsets, uint32_t nkc,
                                                     const uint32_t* __restrict__ poff, const TV* __restrict__ pval,
                                                     const TV* __restrict__ pcv, const double* __restrict__ len,
                                                     const double* __restrict__ tot, uint32_t K, uint32_t n,
                                                     uint32_t slots, double alpha) {
    constexpr int ST = vaw_steps<TV>();
    constexpr bool GEN = VAW == 3;
    const uint32_t a = blockIdx.x, lane = threadIdx.x;
    const uint32_t k0 = blockIdx.y * 64 * VH_COLS + lane;
    __shared__ double sk[64], sc[64], se[64], su[GEN ? 64 : 1];
    double tk[VH_COLS], acc[VH_COLS];
#pragma unroll
    for (uint32_t j = 0; j < VH_COLS; ++j) {
        const uint32_t k = k0 + 64 * j;
        tk[j] = k < D ? tau[k] : 0.0;
        acc[j] = 0.0;
    }
    if (a >= n) return;
    if (poff[K] <= slots) {   // (never otherwise, by the plan's bound)
        const double Ta = tot[a];
        const uint32_t e0 = offsets[(size_t)a * nkc], e1 = offsets[(size_t)(a + 1) * nkc];
        for (uint32_t q0 = e0; q0 < e1; q0 += 64) {   // uniform
            double kk = 0.0, cc = 0.0, ee = 0.0, uu = 0.0;
            if (q0 + lane < e1) {
                const unsigned long long ent = rows[q0 + lane];
                const uint32_t r = (uint32_t)(ent >> 32);
                const uint32_t slot = poff[r] + (uint32_t)ent;
                const double l = len[r];
                cc = (double)pcv[slot];
                uu = (double)pval[slot];
                kk = VAW == 1 ? l * uu : l;
                ee = Ta - cc > 0.0 ? Ta - cc : 0.0;
            }
            __syncthreads();   // the previous round's entries are read
            sk[lane] = kk;
            sc[lane] = cc;
            se[lane] = ee;
            if constexpr (GEN) su[lane] = uu;
            __syncthreads();
            const uint32_t cnt = min(64u, e1 - q0);
            for (uint32_t i = 0; i < cnt; ++i) {
                const double kv = sk[i], cv = sc[i], ev = se[i];
                if constexpr (GEN) {
                    const double uv = su[i];
#pragma unroll
                    for (uint32_t j = 0; j < VH_COLS; ++j) acc[j] += vaw_one_gen<ST, TV>(kv, uv, cv, ev, tk[j], alpha);
                } else {
#pragma unroll
                    for (uint32_t j = 0; j < VH_COLS; ++j) acc[j] += vaw_one<ST>(kv, cv, ev, tk[j]);
                }
            }
        }
    }
#pragma unroll
    for (uint32_t j = 0; j < VH_COLS; ++j) {
        const uint32_t k = k0 + 64 * j;
        if (k < D) H[(size_t)a * D + k] = acc[j];
    }
}

// mx[0] = max of v[0..n): one workgroup
__global__ void __launch_bounds__(256) k_max(double* __restrict__ mx, const double* __restrict__ v, uint32_t n) {
    __shared__ double part[256];
    double m = 0.0;
    for (uint32_t i = threadIdx.x; i < n; i += 256) m = fmax(m, v[i]);
    part[threadIdx.x] = m;
    __syncthreads();
    for (uint32_t d = 128; d > 0; d >>= 1) {
        if (threadIdx.x < d) part[threadIdx.x] = fmax(part[threadIdx.x], part[threadIdx.x + d]);
        __syncthreads();
    }
    if (threadIdx.x == 0) mx[0] = part[0];
}

// cnt[r] += samples of the block with E[knodes[r]][s] > 0: one workgroup per node
__global__ void __launch_bounds__(256) k_count_present(uint32_t* __restrict__ cnt, const double* __restrict__ E,
                                                       size_t lde, const int32_t* __restrict__ knodes,
                                                       uint32_t n_valid) {
    __shared__ uint32_t part[4];
    const uint32_t r = blockIdx.x;
    const double* row = E + (size_t)knodes[r] * lde;
    uint32_t c = 0;
    for (uint32_t s = threadIdx.x; s < n_valid; s += 256) c += row[s] > 0.0 ? 1u : 0u;
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) c += __shfl_xor(c, d, 64);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) cnt[r] += part[0] + part[1] + part[2] + part[3];
}

template <class TIn>
__global__ void __launch_bounds__(256) k_matrix_to_stripes(double* __restrict__ st, const TIn* __restrict__ M,
                                                           uint32_t n, uint32_t start) {
    const uint32_t k = blockIdx.x * 256 + threadIdx.x;
    const uint32_t s = start + blockIdx.y;
    if (k >= n) return;
    const uint32_t l = (uint32_t)(((uint64_t)k + s + 1) % n);
    st[(size_t)blockIdx.y * n + k] = (double)M[(size_t)k * n + l];
}

template <PairOp OP, class T, int RI, int RJ, int TT = PAIR_TILE, bool LPRE = false, int PW = 0>
hipError_t launch_pairs_t(const PairArgs& a, hipStream_t st) {
    static_assert(PAIR_TILE % TT == 0, "the row stride is padded to PAIR_TILE");
    const uint32_t tiles = (a.n + TT - 1) / TT;
    dim3 grid(tiles, tiles);
    constexpr int NT = (TT / RI) * (TT / RJ);
    hipLaunchKernelGGL((k_pairs<OP, T, RI, RJ, TT, LPRE, PW>), grid, dim3(NT), 0, st, (const T*)a.U, (const T*)a.Cn, a.rows, a.ldn,
                       (const T*)a.len, a.K, a.n, a.S, a.totals, a.alpha, (int)a.normalized, a.out, (int)a.out_fp64,
                       a.band, a.out2 ? 1 : 0, a.out2, a.flags, a.flag_tau);
    return hipGetLastError();
}

// Two-accumulator ops (generalized, VAW) keep 4 x 4 pairs per thread on 64 x 64
// tiles: their FP32 + FP64 accumulators for 4 x 8 pairs do not fit in registers.
template <class T>
hipError_t launch_pairs_prec(PairOp op, const PairArgs& a, hipStream_t st) {
    switch (op) {
    case PairOp::weighted: return launch_pairs_t<PairOp::weighted, T, 8, 8>(a, st);
    case PairOp::unweighted_plain: return launch_pairs_t<PairOp::unweighted_plain, T, 4, 8>(a, st);
    // FP32 generalized powers (pair_term's PW): alpha = 1/2 by square roots, outside [0, 1]
    // with the exponent kept exact, otherwise the hardware exp2(a log2 s)
    case PairOp::generalized:
        if (sizeof(T) == 4 && a.alpha == 0.5) return launch_pairs_t<PairOp::generalized, T, 4, 4, 64, false, 2>(a, st);
        return sizeof(T) == 4 && !(a.alpha >= 0.0 && a.alpha <= 1.0)
                   ? launch_pairs_t<PairOp::generalized, T, 4, 4, 64, false, 1>(a, st)
                   : launch_pairs_t<PairOp::generalized, T, 4, 4, 64>(a, st);
    case PairOp::vaw_weighted:
        // FP64 only: in FP32, rounding l u and l v separately doubles the worst-case error of
        // l |u - v| (2^-23 (u + v) l), past what the near-duplicate bound allows
        return a.len_nonneg && sizeof(T) == 8 ? launch_pairs_t<PairOp::vaw_weighted, T, 4, 4, 64, true>(a, st)
                                              : launch_pairs_t<PairOp::vaw_weighted, T, 4, 4, 64>(a, st);
    case PairOp::vaw_generalized:
        if (sizeof(T) == 4 && a.alpha == 0.5)
            return launch_pairs_t<PairOp::vaw_generalized, T, 4, 4, 64, false, 2>(a, st);
        return sizeof(T) == 4 && !(a.alpha >= 0.0 && a.alpha <= 1.0)
                   ? launch_pairs_t<PairOp::vaw_generalized, T, 4, 4, 64, false, 1>(a, st)
                   : launch_pairs_t<PairOp::vaw_generalized, T, 4, 4, 64>(a, st);
    case PairOp::vaw_unweighted:
        return a.len_nonneg ? launch_pairs_t<PairOp::vaw_unweighted, T, 4, 4, 64, true>(a, st)
                            : launch_pairs_t<PairOp::vaw_unweighted, T, 4, 4, 64>(a, st);
    }
    return hipErrorInvalidValue;
}

}  // namespace

__global__ void __launch_bounds__(256) k_fill(double* __restrict__ p, size_t n, double v) {
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) p[i] = v;
}

__global__ void __launch_bounds__(256) k_fill_u32(uint32_t* __restrict__ p, size_t n, uint32_t v) {
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) p[i] = v;
}

// ------------------------------------------------------------------ near-duplicate refinement

// d (row i, column j) is below the refinement bound: tau, or tau * (S_i + S_j)
// when the distance is not normalised (its denominator is S_i + S_j).  NaN
// (an empty sample) is never below it.
__device__ __forceinline__ bool close_pair(float d, const double* __restrict__ S, uint32_t i, uint32_t j, float tau) {
    return S ? (double)d < (double)tau * (S[i] + S[j]) : d < tau;
}

// One workgroup per row i walks out[i][j], j > i, of the selected pairs:
// flags[0] += pairs below the bound, flags[1 + s] = 1 for both samples of each.
__global__ void __launch_bounds__(256) k_flag_close(uint32_t* __restrict__ flags, const float* __restrict__ out,
                                                    uint32_t n, const double* __restrict__ S, float tau,
                                                    GapBand band) {
    __shared__ uint32_t part[4];
    for (uint32_t i = blockIdx.x; i < n; i += gridDim.x) {
        uint32_t c = 0;
        const float* row = out + (size_t)i * n;
        for (uint32_t j = i + 1 + threadIdx.x; j < n; j += 256)
            if (band.has(j - i) && close_pair(row[j], S, i, j, tau)) {
                ++c;
                flags[1 + j] = 1u;
            }
#pragma unroll
        for (int d = 32; d >= 1; d >>= 1) c += __shfl_xor(c, d, 64);
        if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = c;
        __syncthreads();
        if (threadIdx.x == 0) {
            const uint32_t t = part[0] + part[1] + part[2] + part[3];
            if (t) {
                atomicAdd(flags, t);
                flags[1 + i] = 1u;
            }
        }
        __syncthreads();
    }
}

// Sub-table samples a < b (original samples idx[a] < idx[b]) whose FP32
// distance is below the bound take the FP64 distance sub[a][b] of the
// sub-table's matrix (ns x ns), rounded once to FP32, in both triangles; with
// cross, only the pairs a < na <= b.
__global__ void __launch_bounds__(256) k_refine_scatter(float* __restrict__ out, uint32_t n,
                                                        const double* __restrict__ sub, uint32_t ns,
                                                        const uint32_t* __restrict__ idx, uint32_t na, int cross,
                                                        const double* __restrict__ S, float tau) {
    const uint32_t a = blockIdx.y;
    const uint32_t b = blockIdx.x * 256 + threadIdx.x;
    if (b >= ns || b <= a || (cross && (a >= na || b < na))) return;
    const uint32_t i = idx[a], j = idx[b];
    if (!close_pair(out[(size_t)i * n + j], S, i, j, tau)) return;
    const float v = (float)sub[(size_t)a * ns + b];
    out[(size_t)i * n + j] = v;
    out[(size_t)j * n + i] = v;
}

// ------------------------------------------------------------------ launchers

hipError_t launch_fill_u32(uint32_t* p, size_t n, uint32_t v, hipStream_t st) {
    if (!n) return hipSuccess;
    hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)std::min<size_t>(grid1(n, 256), 65536)), dim3(256), 0, st, p, n, v);
    return hipGetLastError();
}

hipError_t launch_flag_close(uint32_t* flags, const float* out, uint32_t n, const double* S, float tau,
                             const GapBand& band, hipStream_t st) {
    const hipError_t e = launch_fill_u32(flags, (size_t)n + 1, 0u, st);
    if (e != hipSuccess || n < 2) return e;
    hipLaunchKernelGGL(k_flag_close, dim3(std::min<uint32_t>(n, 65535)), dim3(256), 0, st, flags, out, n, S, tau, band);
    return hipGetLastError();
}

GapBand stripe_band(uint32_t n, uint32_t s0, uint32_t s1) {
    GapBand b;
    if (s1 <= s0 || n < 2) {   // nothing selected
        b.a0 = 1; b.a1 = 0; b.b0 = 1; b.b1 = 0;
        return b;
    }
    b.a0 = s0 + 1;
    b.a1 = s1;
    b.b0 = n > s1 ? n - s1 : 1;
    b.b1 = n >= s0 + 1 ? n - 1 - s0 : 0;
    return b;
}

hipError_t launch_refine_scatter(float* out, uint32_t n, const double* sub, uint32_t ns, const uint32_t* idx,
                                 uint32_t na, bool cross, const double* S, float tau, hipStream_t st) {
    if (ns < 2) return hipSuccess;
    if (ns > 65535) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_refine_scatter, dim3(grid1(ns, 256), ns), dim3(256), 0, st, out, n, sub, ns, idx, na,
                       (int)cross, S, tau);
    return hipGetLastError();
}

// sub[a * ns + b] = out[idx[a] * n + idx[b]]
__global__ void __launch_bounds__(256) k_gather_square(float* __restrict__ sub, const float* __restrict__ out,
                                                       uint32_t n, const uint32_t* __restrict__ idx, uint32_t ns) {
    const uint32_t a = blockIdx.y, b = blockIdx.x * 256 + threadIdx.x;
    if (a >= ns || b >= ns) return;
    sub[(size_t)a * ns + b] = out[(size_t)idx[a] * n + idx[b]];
}

hipError_t launch_gather_square(float* sub, const float* out, uint32_t n, const uint32_t* idx, uint32_t ns,
                                hipStream_t st) {
    if (!ns) return hipSuccess;
    if (ns > 65535) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_gather_square, dim3(grid1(ns, 256), ns), dim3(256), 0, st, sub, out, n, idx, ns);
    return hipGetLastError();
}

hipError_t launch_fill(double* p, size_t n, double v, hipStream_t st) {
    if (!n) return hipSuccess;
    hipLaunchKernelGGL(k_fill, dim3((unsigned)std::min<size_t>(grid1(n, 256), 65536)), dim3(256), 0, st, p, n, v);
    return hipGetLastError();
}

hipError_t launch_embed_leaves(double* E, size_t ld, uint32_t col0, const int32_t* leaf_k, const int32_t* leaf_row,
                               uint32_t n_leaves, const uint32_t* indptr, const uint32_t* indices, const double* data,
                               const double* totals, bool normalize, hipStream_t st) {
    if (!n_leaves) return hipSuccess;
    if (ld % 2) return hipErrorInvalidValue;   // rows of an even number of columns (16-byte stores)
    hipLaunchKernelGGL(k_embed_leaves, dim3(n_leaves), dim3(EMB_THREADS), 0, st, E, ld, col0, leaf_k, leaf_row, indptr,
                       indices, data, totals, (int)normalize);
    return hipGetLastError();
}

hipError_t launch_embed_level(double* E, size_t ldn, const int32_t* nodes, uint32_t n_nodes, const int32_t* child_ptr,
                              const int32_t* child_idx, hipStream_t st) {
    if (!n_nodes) return hipSuccess;
    if (ldn % 2) return hipErrorInvalidValue;   // two samples per thread
    const uint32_t cb = grid1(ldn / 2, EMB_THREADS);
    const uint32_t gy = std::min<uint32_t>(n_nodes, 8192);
    hipLaunchKernelGGL(k_embed_level, dim3(cb, gy), dim3(EMB_THREADS), 0, st, E, ldn, nodes, n_nodes, child_ptr,
                       child_idx);
    return hipGetLastError();
}

hipError_t launch_embed_paths(double* E, size_t ldn, const HeavyPaths& hp, uint32_t p0, uint32_t p1, hipStream_t st) {
    if (p1 <= p0) return hipSuccess;
    if (ldn % 2) return hipErrorInvalidValue;   // two samples per thread
    const uint32_t cb = grid1(ldn / 2, EMB_THREADS);
    const uint32_t gy = std::min<uint32_t>(p1 - p0, 8192);
    hipLaunchKernelGGL(k_embed_paths, dim3(cb, gy), dim3(EMB_THREADS), 0, st, E, ldn, hp.nodes, hp.light, hp.off + p0,
                       p1 - p0, hp.heavy, hp.child_ptr, hp.child_idx);
    return hipGetLastError();
}

hipError_t launch_bits_paths(uint32_t* B, uint32_t ldb, const HeavyPaths& hp, uint32_t p0, uint32_t p1,
                             hipStream_t st) {
    if (p1 <= p0) return hipSuccess;
    const uint32_t gy = std::min<uint32_t>(p1 - p0, 65535);
    hipLaunchKernelGGL(k_bits_paths, dim3(grid1(ldb, 256), gy), dim3(256), 0, st, B, ldb, hp.nodes, hp.light,
                       hp.off + p0, p1 - p0, hp.heavy, hp.child_ptr, hp.child_idx);
    return hipGetLastError();
}

hipError_t launch_bits_leaves(uint32_t* B, uint32_t ldb, const int32_t* leaf_k, const int32_t* leaf_row,
                              uint32_t n_leaves, const uint32_t* indptr, const uint32_t* indices, const double* data,
                              bool structural, hipStream_t st) {
    if (!n_leaves) return hipSuccess;
    if (!ldb || ldb > BITS_MAX_WORDS) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_bits_leaves, dim3(n_leaves), dim3(256), ldb * sizeof(uint32_t), st, B, ldb, leaf_k, leaf_row,
                       indptr, indices, data, (int)structural);
    return hipGetLastError();
}

hipError_t launch_bits_level(uint32_t* B, uint32_t ldb, const int32_t* nodes, uint32_t n_nodes,
                             const int32_t* child_ptr, const int32_t* child_idx, hipStream_t st) {
    if (!n_nodes) return hipSuccess;
    const uint64_t blocks = ((uint64_t)n_nodes * ldb + 255) / 256;
    if (blocks > 0x7fffffffull) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_bits_level, dim3((uint32_t)blocks), dim3(256), 0, st, B, ldb, nodes, n_nodes, child_ptr,
                       child_idx);
    return hipGetLastError();
}

hipError_t launch_count_bits(uint32_t* cnt, const uint32_t* B, uint32_t ldb, const int32_t* knodes, uint32_t K,
                             hipStream_t st) {
    if (!K) return hipSuccess;
    hipLaunchKernelGGL(k_count_bits, dim3(K), dim3(256), 0, st, cnt, B, ldb, knodes);
    return hipGetLastError();
}

hipError_t launch_bits_smask(uint32_t* smask, const uint32_t* B, uint32_t ldb, const int32_t* nodes, uint32_t K,
                             uint32_t npan, hipStream_t st) {
    if (!K || !npan) return hipSuccess;
    if (npan > ldb) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_bits_smask, dim3((K + 63) / 64, (npan + 63) / 64), dim3(256), 0, st, smask, B, ldb, nodes, K,
                       npan);
    return hipGetLastError();
}

hipError_t launch_presence_T_bits(int8_t* Pt, size_t ldp, const uint32_t* B, uint32_t ldb, const int32_t* col_node,
                                  uint32_t Kp, uint32_t n_pad, uint32_t n, hipStream_t st) {
    if (Kp % 64 || n_pad % 64 || (size_t)ldb * 32 < n_pad) return hipErrorInvalidValue;
    if (!Kp || !n_pad) return hipSuccess;
    hipLaunchKernelGGL(k_presence_T_bits, dim3(Kp / 64, n_pad / 64), dim3(256), 0, st, Pt, ldp, B, ldb, col_node, n);
    return hipGetLastError();
}

uint32_t fixed_rows_per_segment(uint32_t K) { return std::max<uint32_t>(64, (K / 64 + 31) / 32 * 32); }

uint32_t gather_segments(uint32_t K, size_t ldn) {
    if (!K) return 0;
    const uint32_t cb = grid1(ldn, EMB_THREADS);
    const uint32_t want = std::max<uint32_t>(1, 4096 / cb);       // >= 4096 workgroups
    const uint32_t most = (K + 63) / 64;                           // >= 64 rows per segment
    return std::min(want, most);
}

template <bool G, bool S, bool M>
static void gather_sums_t(dim3 grid, hipStream_t st, float* U, size_t ldu, uint32_t col0, double* Spart,
                          uint32_t* masks, uint32_t Kw, uint32_t* smask, const double* E, size_t lde,
                          const int32_t* knodes, const double* klen, uint32_t K, uint32_t rps, uint32_t n_valid,
                          const uint32_t* skip_u) {
    // two samples per thread where the layout allows 16-byte loads (SSU_GATHER2=0: one)
    static const bool two_ok = !(getenv("SSU_GATHER2") && atoi(getenv("SSU_GATHER2")) == 0);
    if (two_ok && lde % 2 == 0 && col0 % 2 == 0 && ldu % 2 == 0) {
        const dim3 g2(grid1(lde / 2, EMB_THREADS), grid.y);
        hipLaunchKernelGGL((k_gather_sums2<G, S, M>), g2, dim3(EMB_THREADS), 0, st, U, ldu, col0, Spart, masks, Kw,
                           smask, E, lde, knodes, klen, K, rps, n_valid, skip_u);
        return;
    }
    hipLaunchKernelGGL((k_gather_sums<G, S, M>), grid, dim3(EMB_THREADS), 0, st, U, ldu, col0, Spart, masks, Kw,
                       smask, E, lde, knodes, klen, K, rps, n_valid, skip_u);
}

hipError_t launch_gather_sums(float* U, size_t ldu, double* S, double* Spart, uint32_t* masks, uint32_t* smask,
                              const double* E, size_t lde, uint32_t col0, const int32_t* knodes, const double* klen,
                              uint32_t K, uint32_t n, hipStream_t st, const uint32_t* skip_u, uint32_t rps_fixed) {
    if (masks && (lde % 64 || col0 % 64)) return hipErrorInvalidValue;
    if (rps_fixed % 32) return hipErrorInvalidValue;
    if (smask && !masks) return hipErrorInvalidValue;
    if (col0 >= n) return hipSuccess;
    const uint32_t cnt = (uint32_t)std::min<size_t>(lde, n - col0);   // real samples in this block
    if (!K) {
        if (S) hipLaunchKernelGGL(k_sum_segments, dim3(grid1(cnt, EMB_THREADS)), dim3(EMB_THREADS), 0, st, S + col0,
                                  Spart, lde, 0u, cnt);
        return hipGetLastError();
    }
    const size_t ldn = lde;
    const uint32_t cb = grid1(ldn, EMB_THREADS);
    const uint32_t segs0 = gather_segments(K, ldn);
    const uint32_t rps = rps_fixed ? rps_fixed
                                   : (((K + segs0 - 1) / segs0) + 31) / 32 * 32;   // mask words never straddle segments
    const uint32_t segs = (K + rps - 1) / rps;
    const uint32_t Kw = (K + 31) / 32;
    const dim3 grid(cb, segs);
    const int key = (U ? 4 : 0) | (S ? 2 : 0) | (masks ? 1 : 0);
    switch (key) {
    case 7: gather_sums_t<true, true, true>(grid, st, U, ldu, col0, Spart, masks, Kw, smask, E, lde, knodes, klen, K, rps, cnt, skip_u); break;
    case 6: gather_sums_t<true, true, false>(grid, st, U, ldu, col0, Spart, masks, Kw, smask, E, lde, knodes, klen, K, rps, cnt, skip_u); break;
    case 5: gather_sums_t<true, false, true>(grid, st, U, ldu, col0, Spart, masks, Kw, smask, E, lde, knodes, klen, K, rps, cnt, skip_u); break;
    case 4: gather_sums_t<true, false, false>(grid, st, U, ldu, col0, Spart, masks, Kw, smask, E, lde, knodes, klen, K, rps, cnt, skip_u); break;
    case 3: gather_sums_t<false, true, true>(grid, st, U, ldu, col0, Spart, masks, Kw, smask, E, lde, knodes, klen, K, rps, cnt, skip_u); break;
    case 2: gather_sums_t<false, true, false>(grid, st, U, ldu, col0, Spart, masks, Kw, smask, E, lde, knodes, klen, K, rps, cnt, skip_u); break;
    case 1: gather_sums_t<false, false, true>(grid, st, U, ldu, col0, Spart, masks, Kw, smask, E, lde, knodes, klen, K, rps, cnt, skip_u); break;
    case 0: return hipSuccess;
    default: return hipErrorInvalidValue;
    }
    if (S) hipLaunchKernelGGL(k_sum_segments, dim3(grid1(cnt, EMB_THREADS)), dim3(EMB_THREADS), 0, st, S + col0, Spart,
                              lde, segs, cnt);
    return hipGetLastError();
}

hipError_t launch_faith(double* out, const double* E, size_t ldn, const double* len, uint32_t m, uint32_t n,
                        hipStream_t st) {
    if (!n) return hipSuccess;
    hipLaunchKernelGGL(k_faith, dim3(grid1(n, 64)), dim3(64), 0, st, out, E, ldn, len, m, n);
    return hipGetLastError();
}

hipError_t launch_presence_T(int8_t* Pt, size_t ldp, const double* E, size_t lde, uint32_t col0, const int32_t* col_node,
                             uint32_t Kp, uint32_t n, hipStream_t st) {
    if (Kp % 64 || lde % 64 || col0 % 64) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_presence_T, dim3(Kp / 64, (uint32_t)(lde / 64)), dim3(256), 0, st, Pt, ldp, E, lde, col0,
                       col_node, n);
    return hipGetLastError();
}

hipError_t launch_gram_i8(double* G, const int8_t* Pt, size_t ldp, const int8_t* digits, const GramChunk* chunks,
                          uint32_t n_chunks, uint32_t n_pad, const uint32_t* order, uint32_t n_tiles, hipStream_t st) {
    if (n_pad % GRAM_BM || ldp % GRAM_BK) return hipErrorInvalidValue;
    const uint32_t tiles = n_pad / GRAM_BM;
    if (tiles > 65535 || n_tiles > tiles * (tiles + 1) / 2 || !order) return hipErrorInvalidValue;
    if (!n_tiles) return hipSuccess;
    const uint32_t per_xcd = (n_tiles + N_XCDS - 1) / N_XCDS;
    hipLaunchKernelGGL(k_gram_i8r, dim3(per_xcd * N_XCDS), dim3(GRAM_THREADS), 0, st, G, tiles, Pt, ldp, digits,
                       chunks, n_chunks, (uint32_t)(ldp / GRAM_BK), order, n_tiles, per_xcd);
    return hipGetLastError();
}

template <class TOut>
hipError_t launch_unweighted_combine(TOut* out, const double* G, uint32_t n_pad, uint32_t n, bool normalized,
                                     hipStream_t st, uint32_t row0, uint32_t row1) {
    row1 = std::min(row1, n);
    if (row0 >= row1) return hipSuccess;
    hipLaunchKernelGGL((k_unweighted_combine<TOut>), dim3(grid1(n, 64), grid1(row1 - row0, 4)), dim3(256), 0, st, out,
                       G, n_pad / GRAM_BM, n, (int)normalized, row0, row1);
    return hipGetLastError();
}
template <class TOut>
hipError_t launch_unweighted_combine_band(TOut* out, const double* G, uint32_t n_pad, uint32_t n, bool normalized,
                                          hipStream_t st, uint32_t row0, uint32_t row1) {
    row1 = std::min(row1, n);
    if (row0 >= row1) return hipSuccess;
    if (row0 % 64 || (row1 % 64 && row1 != n) || n_pad < n) return hipErrorInvalidValue;
    const uint32_t nb = (n + 63) / 64, rb = (row1 - row0 + 63) / 64;
    hipLaunchKernelGGL((k_unweighted_combine_band<TOut>), dim3(nb, rb), dim3(256), 0, st, out, G, n_pad / GRAM_BM, n,
                       (int)normalized, row0, row1);
    return hipGetLastError();
}
template hipError_t launch_unweighted_combine_band<float>(float*, const double*, uint32_t, uint32_t, bool,
                                                          hipStream_t, uint32_t, uint32_t);
template hipError_t launch_unweighted_combine_band<double>(double*, const double*, uint32_t, uint32_t, bool,
                                                           hipStream_t, uint32_t, uint32_t);
template hipError_t launch_unweighted_combine<float>(float*, const double*, uint32_t, uint32_t, bool, hipStream_t,
                                                     uint32_t, uint32_t);
template hipError_t launch_unweighted_combine<double>(double*, const double*, uint32_t, uint32_t, bool, hipStream_t,
                                                      uint32_t, uint32_t);

hipError_t launch_pairs(PairOp op, bool fp64, const PairArgs& a, hipStream_t st) {
    if (a.out2 && op == PairOp::weighted) return hipErrorInvalidValue;   // (one accumulator: no denominator)
    return fp64 ? launch_pairs_prec<double>(op, a, st) : launch_pairs_prec<float>(op, a, st);
}

template <class TIn, class TOut>
static void pairs_sparse_t(uint32_t grid, hipStream_t st, const SparseArgs& a, uint32_t per_xcd) {
    hipLaunchKernelGGL((k_pairs_sparse<TIn, TOut>), dim3(grid), dim3(64), 0, st, (const TIn*)a.U, a.rows, a.ldn, a.len,
                       a.masks, a.Kw, a.n, a.S, (int)a.normalized, (TOut*)a.out, a.visits, a.order, a.n_blocks,
                       per_xcd, a.dmask, (int)(a.dmask != nullptr));
}

std::vector<uint32_t> upper_block_order(uint32_t nb, uint32_t super) {
    std::vector<uint32_t> order;
    order.reserve((size_t)nb * (nb + 1) / 2);
    for (uint32_t si = 0; si < nb; si += super)
        for (uint32_t sj = si; sj < nb; sj += super)
            for (uint32_t bi = si; bi < std::min(nb, si + super); ++bi)
                for (uint32_t bj = std::max(bi, sj); bj < std::min(nb, sj + super); ++bj) order.push_back(bi << 16 | bj);
    return order;
}

std::vector<uint32_t> rect_block_order(uint32_t n) {
    constexpr uint32_t SI = 32, SJ = 16;   // 32 x 16 blocks of 32 x 64 pairs: ~ the waves resident on one XCD
    const uint32_t nbi = (n + S3_BI - 1) / S3_BI, nbj = (n + S3_BJ - 1) / S3_BJ;
    std::vector<uint32_t> order;
    for (uint32_t si = 0; si < nbi; si += SI)
        for (uint32_t sj = si / 2; sj < nbj; sj += SJ)
            for (uint32_t bi = si; bi < std::min(nbi, si + SI); ++bi)
                for (uint32_t bj = sj; bj < std::min(nbj, sj + SJ); ++bj)
                    if (bi <= 2 * bj + 1) order.push_back(bi << 16 | bj);
    return order;
}

template <class TOut>
static void pairs_sparse32_t(uint32_t grid, hipStream_t st, const SparseArgs& a, uint32_t per_xcd) {
    hipLaunchKernelGGL((k_pairs_sparse32<TOut>), dim3(grid), dim3(64), 0, st, (const float*)a.U, a.ldn, a.len, a.masks,
                       a.Kw, a.n, a.S, (int)a.normalized, (TOut*)a.out, a.visits, a.order, a.n_blocks, per_xcd,
                       a.dmask, (int)(a.dmask != nullptr));
}

static bool sparse_args_ok(const SparseNodesArgs& a) {
    // posting slots are addressed with 32-bit offsets (bytes included)
    const bool common = a.dmask && a.smask && a.pc && a.pidx && a.poff && a.len && a.rows && a.rcount && a.roff &&
                        a.rbsum && a.npan == (a.n + 31) / 32 && a.slots < (1u << 30);
    if (!common) return false;
    if (a.gram) return a.n_pad >= a.n && a.n_pad % GRAM_BM == 0;   // unweighted: values, sums and out unused
    return a.pcnt && a.obsum && a.thr && a.pval && a.U && (a.S || a.vaw) && a.out && a.out_fp64 == a.values_fp64 &&
           a.ldu >= (size_t)a.npan * 32 && (!a.vaw || (a.Cv && a.pcv));
}

hipError_t launch_sparse_lists(const SparseNodesArgs& a, hipStream_t st) {
    if (!a.K || !a.n) return hipSuccess;
    if (!sparse_args_ok(a)) return hipErrorInvalidValue;
    const uint32_t gk = (a.K + 63) / 64;
    // weighted: dense nodes by occupancy, now, and the posting offsets as the scan of the sparse
    // nodes' counts (at most `slots` by the plan's structural bound); unweighted: all K nodes are
    // sparse (dmask zero) with offsets fixed since plan creation
    if (!a.gram) {
        const uint32_t nb = (uint32_t)((a.K + SCAN_BLOCK - 1) / SCAN_BLOCK);
        hipLaunchKernelGGL(k_node_split, dim3(gk), dim3(64), 0, st, a.dmask, a.pcnt, a.smask, a.K, a.npan, a.thr,
                           a.fixed_dense);
        hipLaunchKernelGGL(k_scan_sums, dim3(nb), dim3(SCAN_T), 0, st, a.obsum, a.pcnt, (size_t)a.K);
        hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(SCAN_T), 0, st, a.obsum, nb);
        hipLaunchKernelGGL(k_scan_apply, dim3(nb), dim3(SCAN_T), 0, st, a.poff, a.pcnt, a.obsum, (size_t)a.K, nb);
    }
    const bool narrow = a.n <= 65536;   // sample indices fit 16 bits: 6 bytes per posting instead of 8
    if (a.values_fp64 && !a.src_fp64) return hipErrorInvalidValue;   // FP64 values come from FP64 rows
    const int key = (narrow ? 1 : 0) | (a.values_fp64 ? 2 : 0) | (a.src_fp64 ? 4 : 0);
    const void* U = a.gram ? nullptr : a.U;
    void* pval = a.gram ? nullptr : a.pval;
#define SSU_POSTINGS(IX, TV, TU)                                                                               \
    hipLaunchKernelGGL((k_postings<IX, TV, TU>), dim3(gk), dim3(64 * PST_WAVES), 0, st, (IX*)a.pidx, (TV*)pval, a.pc, \
                       a.smask, a.dmask, (const TU*)U, a.urows, a.ldu, a.K, a.npan, (const uint32_t*)a.poff, a.slots)
    switch (key) {
    case 0: SSU_POSTINGS(uint32_t, float, float); break;
    case 1: SSU_POSTINGS(uint16_t, float, float); break;
    case 4: SSU_POSTINGS(uint32_t, float, double); break;
    case 5: SSU_POSTINGS(uint16_t, float, double); break;
    case 6: SSU_POSTINGS(uint32_t, double, double); break;
    default: SSU_POSTINGS(uint16_t, double, double); break;
    }
#undef SSU_POSTINGS
    if (a.vaw) {   // the counts, from the count rows, in the same slots (indices already written)
#define SSU_COUNTS(IX, TV, TU)                                                                                  \
    hipLaunchKernelGGL((k_postings<IX, TV, TU>), dim3(gk), dim3(64 * PST_WAVES), 0, st, (IX*)nullptr, (TV*)a.pcv,  \
                       a.pc, a.smask, a.dmask, (const TU*)a.Cv, a.urows, a.ldu, a.K, a.npan, (const uint32_t*)a.poff, \
                       a.slots)
        switch (key) {
        case 0: case 1: SSU_COUNTS(uint32_t, float, float); break;
        case 4: case 5: SSU_COUNTS(uint32_t, float, double); break;
        default: SSU_COUNTS(uint32_t, double, double); break;
        }
#undef SSU_COUNTS
    }
    // row lists: counts per (sample, node chunk), their exclusive scan, then the entries
    const uint32_t nkc = row_list_chunks(a.n, a.K), kcw = row_list_width(a.n, a.K);
    const size_t m = (size_t)a.n * nkc;
    const uint32_t nsb = (uint32_t)((m + SCAN_BLOCK - 1) / SCAN_BLOCK);
    hipLaunchKernelGGL(k_row_lists, dim3(a.npan, nkc), dim3(64), 0, st, nullptr, a.rcount, nullptr, a.smask, a.dmask,
                       a.pc, a.K, a.n, nkc, kcw, (const uint32_t*)a.poff, a.slots);
    hipLaunchKernelGGL(k_scan_sums, dim3(nsb), dim3(SCAN_T), 0, st, a.rbsum, a.rcount, m);
    hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(SCAN_T), 0, st, a.rbsum, nsb);
    hipLaunchKernelGGL(k_scan_apply, dim3(nsb), dim3(SCAN_T), 0, st, a.roff, a.rcount, a.rbsum, m, nsb);
    hipLaunchKernelGGL(k_row_lists, dim3(a.npan, nkc), dim3(64), 0, st, a.rows, nullptr, a.roff, a.smask, a.dmask,
                       a.pc, a.K, a.n, nkc, kcw, (const uint32_t*)a.poff, a.slots);
    return hipGetLastError();
}

// weighted: the partial sums go out in the value type (FP32 or FP64 output)
template <class IX, bool GRAM, class TV, bool GEN = false, int VAW = 0>
static void pairs_rows_t(uint32_t blocks, hipStream_t st, const SparseNodesArgs& a, uint32_t nchunks, uint32_t nkc,
                         uint32_t a0, uint32_t na, uint32_t mode) {
    hipLaunchKernelGGL((k_pairs_rows<TV, IX, GRAM, TV, GEN, VAW>), dim3(blocks), dim3(64 * SR_WAVES), 0, st,
                       GRAM ? nullptr : (TV*)a.out, a.walked, a.rows, a.roff, a.pc, (const IX*)a.pidx,
                       GRAM ? nullptr : (const TV*)a.pval, a.len, GRAM ? nullptr : a.S, a.K, a.n, a.npan,
                       (const uint32_t*)a.poff, a.slots, nchunks, nkc, a.gram, a.n_pad / GRAM_BM, a.ssp,
                       (TV*)a.out2, a.pspmax, a.alpha, a.band, a0, na, mode, (const TV*)a.pcv, a.tot, a.H, a.D,
                       a.kidx);
}

// VAW: node bounds, row bounds, one-sided sums, then the walk
template <class IX, class TV, int VAW>
static void vaw_rows_t(uint32_t blocks, hipStream_t st, const SparseNodesArgs& a, uint32_t nchunks, uint32_t nkc,
                       uint32_t a0, uint32_t na, uint32_t mode) {
    hipLaunchKernelGGL((k_vaw_node_bound<IX, TV, VAW == 2, VAW == 3>), dim3(a.K), dim3(64), 0, st,
                       a.nbound, (const IX*)a.pidx, (const TV*)a.pval, (const TV*)a.pcv, (const uint32_t*)a.poff,
                       a.tot, a.len, a.K, a.slots, a.alpha);
    hipLaunchKernelGGL(k_row_weights, dim3(a.n), dim3(64), 0, st, a.ssp, a.rows, a.roff, nkc, a.nbound,
                       (const uint32_t*)a.poff, a.K, a.n, a.slots);
    hipLaunchKernelGGL((k_vaw_onesided<TV, VAW>), dim3(a.n, (a.D + 64 * VH_COLS - 1) / (64 * VH_COLS)), dim3(64), 0,
                       st, a.H, a.D, a.tau, a.rows, a.roff, nkc, (const uint32_t*)a.poff, (const TV*)a.pval,
                       (const TV*)a.pcv, a.len, a.tot, a.K, a.n, a.slots, a.alpha);
    pairs_rows_t<IX, false, TV, false, VAW>(blocks, st, a, nchunks, nkc, a0, na, mode);
}
template <int VAW>
static void vaw_rows_v(bool narrow, bool fp64, uint32_t blocks, hipStream_t st, const SparseNodesArgs& a,
                       uint32_t nchunks, uint32_t nkc, uint32_t a0, uint32_t na, uint32_t mode) {
    if (fp64) {
        if (narrow) vaw_rows_t<uint16_t, double, VAW>(blocks, st, a, nchunks, nkc, a0, na, mode);
        else vaw_rows_t<uint32_t, double, VAW>(blocks, st, a, nchunks, nkc, a0, na, mode);
    } else {
        if (narrow) vaw_rows_t<uint16_t, float, VAW>(blocks, st, a, nchunks, nkc, a0, na, mode);
        else vaw_rows_t<uint32_t, float, VAW>(blocks, st, a, nchunks, nkc, a0, na, mode);
    }
}

hipError_t launch_sparse_rows(const SparseNodesArgs& a, hipStream_t st, uint32_t row0, uint32_t row1, int parts) {
    if (!a.K || !a.n) return hipSuccess;
    if (!sparse_args_ok(a)) return hipErrorInvalidValue;
    row1 = std::min(row1, a.n);
    if (row0 >= row1) return hipSuccess;
    const uint32_t mode = (parts & ROWS_OFF_DIAGONAL ? SR_WALK : 0u) | (parts & ROWS_DIAGONAL ? SR_DIAG : 0u);
    if (!mode || (mode == SR_DIAG && !a.gram)) return hipErrorInvalidValue;   // (the diagonal is the Gram's)
    const uint32_t na = row1 - row0;
    const uint32_t nkc = row_list_chunks(a.n, a.K);
    const uint32_t cw = a.gen || a.vaw ? SR_CW / 2 : SR_CW;
    const uint32_t nchunks = (a.n + cw - 1) / cw;
    const uint64_t blocks = mode == SR_DIAG ? (uint64_t)na : (uint64_t)na * nchunks;
    if (blocks > 0x7fffffffull) return hipErrorInvalidValue;
    const bool narrow = a.n <= 65536;
    const bool whole = row0 == 0 && row1 == a.n;   // the per-sample sums are made for all rows at once
    if (a.vaw) {
        if (a.gram || a.gen || a.vaw < 1 || a.vaw > 3 || !a.ssp || !a.out2 || !a.pcv || !a.tot || !a.tau || !a.H ||
            !a.nbound || !a.D || a.D > a.n || (!a.kidx && a.D != a.n) || !whole ||
            (a.vaw == 3 && !(a.alpha >= 0.0 && a.alpha <= 1.0)))
            return hipErrorInvalidValue;
        const uint32_t nb = (uint32_t)blocks;
        if (a.vaw == 1) vaw_rows_v<1>(narrow, a.values_fp64, nb, st, a, nchunks, nkc, row0, na, mode);
        else if (a.vaw == 2) vaw_rows_v<2>(narrow, a.values_fp64, nb, st, a, nchunks, nkc, row0, na, mode);
        else vaw_rows_v<3>(narrow, a.values_fp64, nb, st, a, nchunks, nkc, row0, na, mode);
        return hipGetLastError();
    }
    if (a.gen && a.values_fp64) {   // generalized, FP64 values
        if (a.gram || !a.ssp || !a.out2 || !a.pspmax || !whole) return hipErrorInvalidValue;
        hipLaunchKernelGGL((k_sparse_sums<double, true>), dim3(a.n), dim3(64), 0, st, a.ssp, a.rows, a.roff, nkc,
                           (const uint32_t*)a.poff, (const double*)a.pval, a.len, a.K, a.n, a.slots, a.alpha);
        hipLaunchKernelGGL(k_max, dim3(1), dim3(256), 0, st, a.pspmax, a.ssp, a.n);
        if (narrow) pairs_rows_t<uint16_t, false, double, true>((uint32_t)blocks, st, a, nchunks, nkc, row0, na, mode);
        else pairs_rows_t<uint32_t, false, double, true>((uint32_t)blocks, st, a, nchunks, nkc, row0, na, mode);
    } else if (a.gen) {   // generalized, FP32 values
        if (a.gram || !a.ssp || !a.out2 || !a.pspmax || !whole) return hipErrorInvalidValue;
        hipLaunchKernelGGL((k_sparse_sums<float, true>), dim3(a.n), dim3(64), 0, st, a.ssp, a.rows, a.roff, nkc,
                           (const uint32_t*)a.poff, (const float*)a.pval, a.len, a.K, a.n, a.slots, a.alpha);
        hipLaunchKernelGGL(k_max, dim3(1), dim3(256), 0, st, a.pspmax, a.ssp, a.n);
        if (narrow) pairs_rows_t<uint16_t, false, float, true>((uint32_t)blocks, st, a, nchunks, nkc, row0, na, mode);
        else pairs_rows_t<uint32_t, false, float, true>((uint32_t)blocks, st, a, nchunks, nkc, row0, na, mode);
    } else if (a.gram) {
        if (narrow) pairs_rows_t<uint16_t, true, float>((uint32_t)blocks, st, a, nchunks, nkc, row0, na, mode);
        else pairs_rows_t<uint32_t, true, float>((uint32_t)blocks, st, a, nchunks, nkc, row0, na, mode);
    } else if (a.values_fp64) {
        if (narrow) pairs_rows_t<uint16_t, false, double>((uint32_t)blocks, st, a, nchunks, nkc, row0, na, mode);
        else pairs_rows_t<uint32_t, false, double>((uint32_t)blocks, st, a, nchunks, nkc, row0, na, mode);
    } else {
        if (!a.ssp || !whole) return hipErrorInvalidValue;
        hipLaunchKernelGGL((k_sparse_sums<float, false>), dim3(a.n), dim3(64), 0, st, a.ssp, a.rows, a.roff, nkc,
                           (const uint32_t*)a.poff, (const float*)a.pval, a.len, a.K, a.n, a.slots, 0.0f);
        if (narrow) pairs_rows_t<uint16_t, false, float>((uint32_t)blocks, st, a, nchunks, nkc, row0, na, mode);
        else pairs_rows_t<uint32_t, false, float>((uint32_t)blocks, st, a, nchunks, nkc, row0, na, mode);
    }
    return hipGetLastError();
}

hipError_t launch_sparse_nodes(const SparseNodesArgs& a, hipStream_t st) {
    if (a.gram) return hipErrorInvalidValue;   // the Gram tiles must be written between the two halves
    const hipError_t e = launch_sparse_lists(a, st);
    return e != hipSuccess ? e : launch_sparse_rows(a, st, 0, a.n, ROWS_OFF_DIAGONAL | ROWS_DIAGONAL);
}

hipError_t launch_count_present(uint32_t* cnt, const double* E, size_t lde, uint32_t col0, const int32_t* knodes,
                                uint32_t K, uint32_t n, hipStream_t st) {
    if (!K || col0 >= n) return hipSuccess;
    const uint32_t n_valid = (uint32_t)std::min<size_t>(lde, n - col0);
    hipLaunchKernelGGL(k_count_present, dim3(K), dim3(256), 0, st, cnt, E, lde, knodes, n_valid);
    return hipGetLastError();
}

uint32_t sparse_rows_blocks(uint32_t n) { return n * ((n + SR_CW / 2 - 1) / (SR_CW / 2)); }   // (half-width chunks)
uint32_t sparse_rows_launched(uint32_t n, bool gen) {
    const uint32_t cw = gen ? SR_CW / 2 : SR_CW;
    return n * ((n + cw - 1) / cw);
}
uint32_t row_list_width(uint32_t n, uint32_t K) {
    uint64_t w = RL_KC;
    auto chunks = [&](uint64_t width) { return std::max<uint64_t>(1, (K + width - 1) / width); };
    while (chunks(w) > RL_MAX_CHUNKS || (uint64_t)std::max<uint32_t>(n, 1) * chunks(w) > RL_MAX_ENTRIES) w *= 2;
    return (uint32_t)std::min<uint64_t>(w, 1ull << 31);
}
uint32_t row_list_chunks(uint32_t n, uint32_t K) {
    const uint64_t w = row_list_width(n, K);
    return (uint32_t)std::max<uint64_t>(1, (K + w - 1) / w);
}
size_t offset_scan_blocks(uint32_t K) { return ((size_t)K + SCAN_BLOCK - 1) / SCAN_BLOCK; }
size_t row_list_scan_blocks(uint32_t n, uint32_t K) {
    return ((size_t)n * row_list_chunks(n, K) + SCAN_BLOCK - 1) / SCAN_BLOCK;
}

hipError_t launch_pairs_gen32(const SparseArgs& a, float alpha, void* out2, hipStream_t st) {
    if (a.ldn % S3_BJ || a.out_fp64) return hipErrorInvalidValue;
    const uint32_t nbj = (a.n + S3_BJ - 1) / S3_BJ;
    if (!nbj) return hipSuccess;
    if (a.dmask && !out2) return hipErrorInvalidValue;
    if (!a.order) return hipErrorInvalidValue;
    if (!a.n_blocks) return hipSuccess;   // a stripe selection without a block here
    const uint32_t per_xcd = (a.n_blocks + N_XCDS - 1) / N_XCDS;
    if (alpha == 0.5f)   // square roots in place of log2 / exp2 (gen_term's HALF)
        hipLaunchKernelGGL((k_pairs_gen32<float, true>), dim3(per_xcd * N_XCDS), dim3(64), 0, st, (const float*)a.U,
                           a.ldn, a.len, a.masks, a.Kw, a.n, (float*)a.out, (float*)out2, a.visits, a.order,
                           a.n_blocks, per_xcd, a.dmask, (int)(a.dmask != nullptr), alpha);
    else
        hipLaunchKernelGGL((k_pairs_gen32<float>), dim3(per_xcd * N_XCDS), dim3(64), 0, st, (const float*)a.U, a.ldn,
                           a.len, a.masks, a.Kw, a.n, (float*)a.out, (float*)out2, a.visits, a.order, a.n_blocks,
                           per_xcd, a.dmask, (int)(a.dmask != nullptr), alpha);
    return hipGetLastError();
}

hipError_t launch_pairs_gen64(const SparseArgs& a, double alpha, void* out2, hipStream_t st) {
    if (a.ldn % S3_BJ || !a.out_fp64 || !a.rows) return hipErrorInvalidValue;
    const uint32_t nbj = (a.n + S3_BJ - 1) / S3_BJ;
    if (!nbj) return hipSuccess;
    if (a.dmask && !out2) return hipErrorInvalidValue;
    if (!a.order) return hipErrorInvalidValue;
    if (!a.n_blocks) return hipSuccess;   // a stripe selection without a block here
    const uint32_t per_xcd = (a.n_blocks + N_XCDS - 1) / N_XCDS;
    hipLaunchKernelGGL(k_pairs_gen64, dim3(per_xcd * N_XCDS), dim3(64), 0, st, (const double*)a.U, a.rows, a.ldn,
                       a.len, a.masks, a.Kw, a.n, (double*)a.out, (double*)out2, a.visits, a.order, a.n_blocks,
                       per_xcd, a.dmask, (int)(a.dmask != nullptr), alpha);
    return hipGetLastError();
}

hipError_t launch_pairs_sparse32(const SparseArgs& a, hipStream_t st) {
    if (a.ldn % S3_BJ) return hipErrorInvalidValue;
    const uint32_t nbi = (a.n + S3_BI - 1) / S3_BI, nbj = (a.n + S3_BJ - 1) / S3_BJ;
    if (!nbi) return hipSuccess;
    if ((size_t)nbj * S3_BJ > a.ldn || nbi > 65535 || !a.order) return hipErrorInvalidValue;
    if (!a.n_blocks) return hipSuccess;   // a stripe selection without a block here
    const uint32_t per_xcd = (a.n_blocks + N_XCDS - 1) / N_XCDS;
    if (a.out_fp64) pairs_sparse32_t<double>(per_xcd * N_XCDS, st, a, per_xcd);
    else pairs_sparse32_t<float>(per_xcd * N_XCDS, st, a, per_xcd);
    return hipGetLastError();
}

hipError_t launch_pairs_sparse(bool in_fp64, const SparseArgs& a, hipStream_t st) {
    if (a.ldn % SP_B) return hipErrorInvalidValue;
    const uint32_t nb = (a.n + SP_B - 1) / SP_B;
    if (!nb) return hipSuccess;
    if ((size_t)nb * SP_B > a.ldn || nb > 65535 || a.n_blocks > nb * (nb + 1) / 2 || !a.order)
        return hipErrorInvalidValue;
    if (!a.n_blocks) return hipSuccess;   // a stripe selection without a block here
    const uint32_t per_xcd = (a.n_blocks + N_XCDS - 1) / N_XCDS;
    const uint32_t grid = per_xcd * N_XCDS;
    if (in_fp64) {
        if (a.out_fp64) pairs_sparse_t<double, double>(grid, st, a, per_xcd);
        else pairs_sparse_t<double, float>(grid, st, a, per_xcd);
    } else {
        if (a.out_fp64) pairs_sparse_t<float, double>(grid, st, a, per_xcd);
        else pairs_sparse_t<float, float>(grid, st, a, per_xcd);
    }
    return hipGetLastError();
}

template <class TIn>
hipError_t launch_matrix_to_stripes(double* stripes, const TIn* M, uint32_t n, uint32_t start, uint32_t stop,
                                    hipStream_t st) {
    if (stop <= start) return hipSuccess;
    hipLaunchKernelGGL((k_matrix_to_stripes<TIn>), dim3(grid1(n, 256), stop - start), dim3(256), 0, st, stripes, M, n,
                       start);
    return hipGetLastError();
}
template hipError_t launch_matrix_to_stripes<float>(double*, const float*, uint32_t, uint32_t, uint32_t, hipStream_t);
template hipError_t launch_matrix_to_stripes<double>(double*, const double*, uint32_t, uint32_t, uint32_t, hipStream_t);

}  // namespace ssu
