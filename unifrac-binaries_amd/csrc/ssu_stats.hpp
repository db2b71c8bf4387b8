// Ordination and significance on a distance matrix, on the device:
//   * Gower centring, B = -1/2 (I - 11'/n) D∘D (I - 11'/n);
//   * PCoA by randomized FSVD (Halko, Martinsson, Shkolnisky & Tygert, "An
//     algorithm for the principal component analysis of large data sets",
//     SIAM J. Sci. Comput. 2011; the variant scikit-bio uses for its "fsvd"
//     method: k = dims + 2 Gaussian probes, QR of [B G, B^3 G, ...], SVD of
//     B Q; two power levels here where scikit-bio uses one);
//   * PERMANOVA pseudo-F (Anderson 2001) with a permutation p-value.
// The reference reaches these through its third-party skbb library
// (src/skbio_alt.cpp:144-233); the n x n products and the permutation sums
// run here as HIP kernels, the (n x 3k) QR and the small SVD on the host.
#pragma once

#include <cstdint>
#include <string>

namespace ssu {

// centred (n x n, row major) from mat; element types double or float
bool center_matrix(const void* mat, bool in_fp64, uint32_t n, double* centred, std::string* err);

// The same centring of an FP64 matrix already in device memory into another
// (or the same) device buffer, enqueued on `stream` (nullptr: the null stream)
// without allocating: its row means live in a fixed device array, so n is at
// most NOALLOC_MAX_N.  The first call on a device looks the array up.
constexpr uint32_t NOALLOC_MAX_N = 65536;
bool center_matrix_noalloc(const double* d_mat, uint32_t n, double* d_centred, void* stream, std::string* err);

// eigenvalues[dims], samples[n * dims] (sample-major: eigenvector * sqrt(eigenvalue)),
// proportion[dims] = eigenvalue / trace(B).  Dimensions past n are zero.
bool pcoa(const void* mat, bool in_fp64, uint32_t n, uint32_t dims, uint64_t seed, double* eigenvalues,
          double* samples, double* proportion, std::string* err);

// grouping[n] in 0..groups-1; pseudo-F of the upper triangle and
// p = (#{permutations with F >= F_observed} + 1) / (n_perm + 1).  One group or
// n groups: NaN for both (the statistic is undefined).
bool permanova(const void* mat, bool in_fp64, uint32_t n, const uint32_t* grouping, uint32_t n_perm, uint64_t seed,
               double* fstat, double* pvalue, std::string* err);

}  // namespace ssu
