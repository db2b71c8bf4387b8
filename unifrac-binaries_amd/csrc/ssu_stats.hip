// PCoA (Gower centring + randomized FSVD) and PERMANOVA.  The n x n work
// (centring, the products B X, the permutation sums) runs in the kernels
// below; the thin QR of the n x 3k probe block and the SVD of the small
// factor run on the host in FP64.
#include "ssu_stats.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <limits>
#include <mutex>
#include <vector>

namespace ssu {

namespace {

#define ST_HIP(x)                                                                                    \
    do {                                                                                             \
        hipError_t e_ = (x);                                                                         \
        if (e_ != hipSuccess) {                                                                      \
            if (err) *err = std::string(#x) + ": " + hipGetErrorString(e_);                          \
            return false;                                                                            \
        }                                                                                            \
    } while (0)

struct Buf {
    void* p = nullptr;
    ~Buf() {
        if (p) (void)hipFree(p);
    }
    template <class T> T* as() const { return (T*)p; }
};

bool alloc(Buf& b, size_t bytes, std::string* err) {
    if (hipMalloc(&b.p, std::max<size_t>(bytes, 16)) != hipSuccess) {
        if (err) *err = "hipMalloc(" + std::to_string(bytes) + ") failed";
        b.p = nullptr;
        return false;
    }
    return true;
}

constexpr uint32_t ST_T = 256;

// Sum over the workgroup, the same order every time.
__device__ __forceinline__ double block_sum(double v, double* part) {
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) v += __shfl_down(v, d, 64);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = v;
    __syncthreads();
    double s = 0.0;
    for (uint32_t w = 0; w < ST_T / 64; ++w) s += part[w];
    __syncthreads();
    return s;
}

// rmean[i] = mean_j (-d_ij^2 / 2); upper[i] = sum_{j > i} d_ij^2
template <class T>
__global__ void __launch_bounds__(ST_T) k_row_stats(double* __restrict__ rmean, double* __restrict__ upper,
                                                    const T* __restrict__ D, uint32_t n) {
    __shared__ double part[ST_T / 64];
    const uint32_t i = blockIdx.x;
    const T* row = D + (size_t)i * n;
    double a = 0.0, u = 0.0;
    for (uint32_t j = threadIdx.x; j < n; j += ST_T) {
        const double d = (double)row[j];
        a += -0.5 * d * d;
        if (j > i) u += d * d;
    }
    a = block_sum(a, part);
    u = block_sum(u, part);
    if (threadIdx.x == 0) {
        rmean[i] = a / n;
        if (upper) upper[i] = u;
    }
}

// B_ij = -d_ij^2 / 2 - rmean_i - rmean_j + grand, in place
template <class T>
__global__ void __launch_bounds__(ST_T) k_center_apply(T* __restrict__ B, const double* __restrict__ rmean,
                                                       double grand, uint32_t n) {
    const size_t q = (size_t)blockIdx.x * ST_T + threadIdx.x;
    if (q >= (size_t)n * n) return;
    const uint32_t i = (uint32_t)(q / n), j = (uint32_t)(q % n);
    const double d = (double)B[q];
    B[q] = (T)(-0.5 * d * d - rmean[i] - rmean[j] + grand);
}

// Y[i][c0 + c] = sum_j B_ij X[j][c0 + c] for c < cw <= 32 (X, Y row major with m columns)
template <class T>
__global__ void __launch_bounds__(ST_T) k_symm_mult(double* __restrict__ Y, const T* __restrict__ B,
                                                    const double* __restrict__ X, uint32_t n, uint32_t m,
                                                    uint32_t c0, uint32_t cw) {
    __shared__ double part[32][ST_T / 64];
    const uint32_t i = blockIdx.x;
    const T* row = B + (size_t)i * n;
    double acc[32];
#pragma unroll
    for (int c = 0; c < 32; ++c) acc[c] = 0.0;
    for (uint32_t j = threadIdx.x; j < n; j += ST_T) {
        const double b = (double)row[j];
        const double* x = X + (size_t)j * m + c0;
#pragma unroll
        for (uint32_t c = 0; c < 32; ++c)
            if (c < cw) acc[c] += b * x[c];
    }
    const uint32_t lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (uint32_t c = 0; c < 32; ++c) {
        double v = acc[c];
#pragma unroll
        for (int d = 32; d > 0; d >>= 1) v += __shfl_down(v, d, 64);
        if (lane == 0) part[c][w] = v;
    }
    __syncthreads();
    if (threadIdx.x < cw) {
        double s = 0.0;
        for (uint32_t q = 0; q < ST_T / 64; ++q) s += part[threadIdx.x][q];
        Y[(size_t)i * m + c0 + threadIdx.x] = s;
    }
}

// PERMANOVA within-group sums: partial[p][rb] = sum over rows i of row block rb
// and j > i with L_p(i) == L_p(j) of d_ij^2 / size(L_p(i)).  A workgroup
// owns 64 rows x 64 label rows (permutations); each thread 4 x 4 of them.
constexpr uint32_t PM_R = 64, PM_P = 64, PM_J = 64;
template <class T>
__global__ void __launch_bounds__(ST_T) k_permanova(double* __restrict__ partial, const T* __restrict__ D,
                                                    const uint32_t* __restrict__ L, const double* __restrict__ inv_size,
                                                    uint32_t n, uint32_t n_lab) {
    __shared__ double dsq[PM_R][PM_J + 1];
    __shared__ uint32_t lab[PM_P][PM_J + 1];
    __shared__ double red[16][PM_P];
    const uint32_t rb = blockIdx.x, pb = blockIdx.y;
    const uint32_t i0 = rb * PM_R, p0 = pb * PM_P;
    const uint32_t tr = (threadIdx.x >> 4) * 4, tp = (threadIdx.x & 15) * 4;
    uint32_t li[4][4];
    double acc[4][4];
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const uint32_t i = i0 + tr + r, p = p0 + tp + q;
            li[r][q] = (i < n && p < n_lab) ? L[(size_t)p * n + i] : 0xffffffffu;
            acc[r][q] = 0.0;
        }
    for (uint32_t j0 = i0; j0 < n; j0 += PM_J) {
        for (uint32_t e = threadIdx.x; e < PM_R * PM_J; e += ST_T) {
            const uint32_t r = e / PM_J, c = e % PM_J;
            const uint32_t i = i0 + r, j = j0 + c;
            double d = 0.0;
            if (i < n && j < n) d = (double)D[(size_t)i * n + j];
            dsq[r][c] = d * d;
            const uint32_t p = p0 + r;   // PM_P == PM_R
            lab[r][c] = (p < n_lab && j < n) ? L[(size_t)p * n + j] : 0xfffffffeu;
        }
        __syncthreads();
        for (uint32_t c = 0; c < PM_J; ++c) {
            const uint32_t j = j0 + c;
            uint32_t lj[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) lj[q] = lab[tp + q][c];
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const double d2 = (j > i0 + tr + r) ? dsq[tr + r][c] : 0.0;
#pragma unroll
                for (int q = 0; q < 4; ++q) acc[r][q] += (lj[q] == li[r][q]) ? d2 : 0.0;
            }
        }
        __syncthreads();
    }
    // per permutation: sum the rows of this thread (weighted), then the 16 row groups in order
    double own[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        own[q] = 0.0;
#pragma unroll
        for (int r = 0; r < 4; ++r)
            if (li[r][q] < 0xfffffffeu) own[q] += acc[r][q] * inv_size[li[r][q]];
        red[threadIdx.x >> 4][tp + q] = own[q];
    }
    __syncthreads();
    if (threadIdx.x < PM_P) {
        double s = 0.0;
        for (uint32_t g = 0; g < 16; ++g) s += red[g][threadIdx.x];
        const uint32_t p = p0 + threadIdx.x;
        if (p < n_lab) partial[(size_t)p * gridDim.x + rb] = s;
    }
}

// ----------------------------------------------------------- host helpers

uint64_t mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// standard normal from two counter-based uniforms (Box-Muller)
double gaussian(uint64_t seed, uint64_t idx) {
    const double u1 = ((mix64(seed ^ mix64(2 * idx)) >> 11) + 1) * 0x1.0p-53;   // (0, 1]
    const double u2 = (mix64(seed ^ mix64(2 * idx + 1)) >> 11) * 0x1.0p-53;
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
}

// Power levels of the range finder.  scikit-bio's fsvd uses one; with one, the
// reference's own tolerances on its 57-sample case (src/test_ska.cpp:247-420)
// fail for some probe draws, with two they hold (measured over 40 draws).
constexpr uint32_t PCOA_LEVELS = 2;

// Thin QR by Householder reflections: A is n x m column major (overwritten);
// returns Q, n x q column major, q = min(n, m), with orthonormal columns.
std::vector<double> thin_q(std::vector<double>& A, size_t n, size_t m, size_t& q) {
    q = std::min(n, m);
    std::vector<std::vector<double>> V(q);
    for (size_t k = 0; k < q; ++k) {
        double* col = &A[k * n];
        double norm = 0.0;
        for (size_t i = k; i < n; ++i) norm += col[i] * col[i];
        norm = std::sqrt(norm);
        std::vector<double> v(n - k, 0.0);
        if (norm > 0.0) {
            const double alpha = col[k] >= 0.0 ? -norm : norm;
            for (size_t i = k; i < n; ++i) v[i - k] = col[i];
            v[0] -= alpha;
            double vn = 0.0;
            for (double x : v) vn += x * x;
            vn = std::sqrt(vn);
            if (vn > 0.0)
                for (double& x : v) x /= vn;
            for (size_t c = k; c < m; ++c) {   // A[k:, c] -= 2 v (v . A[k:, c])
                double* a = &A[c * n];
                double dot = 0.0;
                for (size_t i = k; i < n; ++i) dot += v[i - k] * a[i];
                dot *= 2.0;
                for (size_t i = k; i < n; ++i) a[i] -= dot * v[i - k];
            }
        }
        V[k] = std::move(v);
    }
    std::vector<double> Q(n * q, 0.0);
    for (size_t c = 0; c < q; ++c) Q[c * n + c] = 1.0;
    for (size_t k = q; k-- > 0;) {   // Q = H_0 ... H_{q-1} I[:, :q]
        const std::vector<double>& v = V[k];
        for (size_t c = 0; c < q; ++c) {
            double* a = &Q[c * n];
            double dot = 0.0;
            for (size_t i = k; i < n; ++i) dot += v[i - k] * a[i];
            dot *= 2.0;
            for (size_t i = k; i < n; ++i) a[i] -= dot * v[i - k];
        }
    }
    return Q;
}

// R factor of a thin QR (A n x m column major, overwritten), m x m column major
std::vector<double> thin_r(std::vector<double>& A, size_t n, size_t m) {
    std::vector<double> R(m * m, 0.0);
    const size_t q = std::min(n, m);
    for (size_t k = 0; k < q; ++k) {
        double* col = &A[k * n];
        double norm = 0.0;
        for (size_t i = k; i < n; ++i) norm += col[i] * col[i];
        norm = std::sqrt(norm);
        if (norm > 0.0) {
            const double alpha = col[k] >= 0.0 ? -norm : norm;
            std::vector<double> v(col + k, col + n);
            v[0] -= alpha;
            double vn = 0.0;
            for (double x : v) vn += x * x;
            vn = std::sqrt(vn);
            if (vn > 0.0)
                for (double& x : v) x /= vn;
            for (size_t c = k; c < m; ++c) {
                double* a = &A[c * n];
                double dot = 0.0;
                for (size_t i = k; i < n; ++i) dot += v[i - k] * a[i];
                dot *= 2.0;
                for (size_t i = k; i < n; ++i) a[i] -= dot * v[i - k];
            }
        }
        for (size_t c = k; c < m; ++c) R[c * m + k] = A[c * n + k];
    }
    return R;
}

// One-sided Jacobi (Hestenes) on the m x m matrix R (column major): singular
// values S (descending) and right singular vectors V (column major, same order).
void jacobi_svd(std::vector<double> R, size_t m, std::vector<double>& S, std::vector<double>& V) {
    V.assign(m * m, 0.0);
    for (size_t c = 0; c < m; ++c) V[c * m + c] = 1.0;
    for (int sweep = 0; sweep < 60; ++sweep) {
        double off = 0.0;
        for (size_t p = 0; p + 1 < m; ++p)
            for (size_t q = p + 1; q < m; ++q) {
                double a = 0.0, b = 0.0, g = 0.0;
                for (size_t i = 0; i < m; ++i) {
                    a += R[p * m + i] * R[p * m + i];
                    b += R[q * m + i] * R[q * m + i];
                    g += R[p * m + i] * R[q * m + i];
                }
                if (g == 0.0 || std::fabs(g) <= 1e-15 * std::sqrt(a * b)) continue;
                off = std::max(off, std::fabs(g) / std::sqrt(a * b));
                const double zeta = (b - a) / (2.0 * g);
                const double t = (zeta >= 0.0 ? 1.0 : -1.0) / (std::fabs(zeta) + std::sqrt(1.0 + zeta * zeta));
                const double c = 1.0 / std::sqrt(1.0 + t * t), s = c * t;
                for (size_t i = 0; i < m; ++i) {
                    const double x = R[p * m + i], y = R[q * m + i];
                    R[p * m + i] = c * x - s * y;
                    R[q * m + i] = s * x + c * y;
                    const double vx = V[p * m + i], vy = V[q * m + i];
                    V[p * m + i] = c * vx - s * vy;
                    V[q * m + i] = s * vx + c * vy;
                }
            }
        if (off <= 1e-15) break;
    }
    S.assign(m, 0.0);
    for (size_t c = 0; c < m; ++c) {
        double s = 0.0;
        for (size_t i = 0; i < m; ++i) s += R[c * m + i] * R[c * m + i];
        S[c] = std::sqrt(s);
    }
    std::vector<size_t> order(m);
    for (size_t c = 0; c < m; ++c) order[c] = c;
    std::stable_sort(order.begin(), order.end(), [&](size_t x, size_t y) { return S[x] > S[y]; });
    std::vector<double> S2(m), V2(m * m);
    for (size_t c = 0; c < m; ++c) {
        S2[c] = S[order[c]];
        std::copy(&V[order[c] * m], &V[order[c] * m] + m, &V2[c * m]);
    }
    S = std::move(S2);
    V = std::move(V2);
}

template <class T>
bool upload_matrix(Buf& d, const void* mat, uint32_t n, std::string* err) {
    const size_t bytes = (size_t)n * n * sizeof(T);
    if (!alloc(d, bytes, err)) return false;
    ST_HIP(hipMemcpy(d.p, mat, bytes, hipMemcpyHostToDevice));
    return true;
}

template <class T>
bool center_on_device(Buf& dB, const void* mat, uint32_t n, std::vector<double>& rmean, double& grand,
                      std::vector<double>& upper, std::string* err) {
    if (!upload_matrix<T>(dB, mat, n, err)) return false;
    Buf dr, du;
    if (!alloc(dr, n * sizeof(double), err) || !alloc(du, n * sizeof(double), err)) return false;
    hipLaunchKernelGGL((k_row_stats<T>), dim3(n), dim3(ST_T), 0, 0, dr.as<double>(), du.as<double>(),
                       dB.as<const T>(), n);
    ST_HIP(hipGetLastError());
    rmean.resize(n);
    upper.resize(n);
    ST_HIP(hipMemcpy(rmean.data(), dr.p, n * sizeof(double), hipMemcpyDeviceToHost));
    ST_HIP(hipMemcpy(upper.data(), du.p, n * sizeof(double), hipMemcpyDeviceToHost));
    grand = 0.0;
    for (double r : rmean) grand += r;
    grand /= n;
    const size_t nn = (size_t)n * n;
    hipLaunchKernelGGL((k_center_apply<T>), dim3((uint32_t)((nn + ST_T - 1) / ST_T)), dim3(ST_T), 0, 0, dB.as<T>(),
                       dr.as<const double>(), grand, n);
    ST_HIP(hipGetLastError());
    ST_HIP(hipDeviceSynchronize());
    return true;
}

// Y = B X, X and Y n x m row major (device)
template <class T>
bool symm_mult(double* Y, const Buf& dB, const double* X, uint32_t n, uint32_t m, std::string* err) {
    for (uint32_t c0 = 0; c0 < m; c0 += 32) {
        hipLaunchKernelGGL((k_symm_mult<T>), dim3(n), dim3(ST_T), 0, 0, Y, dB.as<const T>(), X, n, m, c0,
                           std::min<uint32_t>(32, m - c0));
        ST_HIP(hipGetLastError());
    }
    return true;
}

template <class T>
bool pcoa_t(const void* mat, uint32_t n, uint32_t dims, uint64_t seed, double* eig, double* samples,
            double* proportion, std::string* err) {
    std::fill(eig, eig + dims, 0.0);
    std::fill(samples, samples + (size_t)n * dims, 0.0);
    std::fill(proportion, proportion + dims, 0.0);
    if (n == 0 || dims == 0) return true;
    Buf dB;
    std::vector<double> rmean, upper;
    double grand = 0.0;
    if (!center_on_device<T>(dB, mat, n, rmean, grand, upper, err)) return false;
    double trace = 0.0;
    for (uint32_t i = 0; i < n; ++i) {
        const double d = (double)((const T*)mat)[(size_t)i * n + i];
        trace += -0.5 * d * d - 2.0 * rmean[i] + grand;
    }
    // probes G (n x k), H = [B G | B^3 G | B^5 G] (n x 3k)
    const uint32_t k = dims + 2, m = (PCOA_LEVELS + 1) * k;
    std::vector<double> G((size_t)n * k);
    for (size_t q = 0; q < G.size(); ++q) G[q] = gaussian(seed, q);
    Buf dX, dY, dZ;
    if (!alloc(dX, G.size() * 8, err) || !alloc(dY, G.size() * 8, err) || !alloc(dZ, G.size() * 8, err))
        return false;
    ST_HIP(hipMemcpy(dX.p, G.data(), G.size() * 8, hipMemcpyHostToDevice));
    std::vector<double> H((size_t)n * m);   // column major
    std::vector<double> h((size_t)n * k);
    for (uint32_t level = 0; level <= PCOA_LEVELS; ++level) {
        // level 0: Y = B G; then Y = B (B Y_prev)
        if (level == 0) {
            if (!symm_mult<T>(dY.as<double>(), dB, dX.as<double>(), n, k, err)) return false;
        } else {
            if (!symm_mult<T>(dZ.as<double>(), dB, dY.as<double>(), n, k, err) ||
                !symm_mult<T>(dY.as<double>(), dB, dZ.as<double>(), n, k, err))
                return false;
        }
        ST_HIP(hipMemcpy(h.data(), dY.p, h.size() * 8, hipMemcpyDeviceToHost));
        for (uint32_t i = 0; i < n; ++i)
            for (uint32_t c = 0; c < k; ++c) H[(size_t)(level * k + c) * n + i] = h[(size_t)i * k + c];
    }
    size_t q = 0;
    const std::vector<double> Q = thin_q(H, n, m, q);
    // T = B Q (n x q)
    std::vector<double> Qr((size_t)n * q);
    for (uint32_t i = 0; i < n; ++i)
        for (size_t c = 0; c < q; ++c) Qr[(size_t)i * q + c] = Q[c * n + i];
    Buf dQ, dTq;
    if (!alloc(dQ, Qr.size() * 8, err) || !alloc(dTq, Qr.size() * 8, err)) return false;
    ST_HIP(hipMemcpy(dQ.p, Qr.data(), Qr.size() * 8, hipMemcpyHostToDevice));
    if (!symm_mult<T>(dTq.as<double>(), dB, dQ.as<double>(), n, (uint32_t)q, err)) return false;
    std::vector<double> tr((size_t)n * q);
    ST_HIP(hipMemcpy(tr.data(), dTq.p, tr.size() * 8, hipMemcpyDeviceToHost));
    std::vector<double> Tc((size_t)n * q);
    for (uint32_t i = 0; i < n; ++i)
        for (size_t c = 0; c < q; ++c) Tc[c * n + i] = tr[(size_t)i * q + c];
    // SVD of T through its R factor: T = Q2 R, R = U S V' -> right singular vectors V
    const std::vector<double> R = thin_r(Tc, n, q);
    std::vector<double> S, V;
    jacobi_svd(R, q, S, V);
    const uint32_t used = (uint32_t)std::min<size_t>(dims, q);
    for (uint32_t d = 0; d < used; ++d) {
        eig[d] = S[d];
        proportion[d] = trace != 0.0 ? S[d] / trace : 0.0;
        const double sq = std::sqrt(std::max(S[d], 0.0));
        for (uint32_t i = 0; i < n; ++i) {   // eigenvector d = Q V[:, d]
            double u = 0.0;
            for (size_t c = 0; c < q; ++c) u += Q[c * n + i] * V[d * q + c];
            samples[(size_t)i * dims + d] = u * sq;
        }
    }
    return true;
}

template <class T>
bool permanova_t(const void* mat, uint32_t n, const uint32_t* grouping, uint32_t n_perm, uint64_t seed,
                 double* fstat, double* pvalue, std::string* err) {
    *fstat = *pvalue = std::numeric_limits<double>::quiet_NaN();
    if (n < 2) return true;
    uint32_t groups = 0;
    for (uint32_t i = 0; i < n; ++i) groups = std::max(groups, grouping[i] + 1);
    std::vector<double> size(groups, 0.0);
    for (uint32_t i = 0; i < n; ++i) size[grouping[i]] += 1.0;
    uint32_t present = 0;
    for (double s : size) present += s > 0.0;
    if (present < 2 || present >= n) return true;   // pseudo-F undefined
    std::vector<double> inv(groups, 0.0);
    for (uint32_t g = 0; g < groups; ++g) inv[g] = size[g] > 0.0 ? 1.0 / size[g] : 0.0;
    // label rows: the grouping, then n_perm Fisher-Yates shuffles of it
    const uint32_t n_lab = n_perm + 1;
    std::vector<uint32_t> L((size_t)n_lab * n);
    std::copy(grouping, grouping + n, L.begin());
    for (uint32_t p = 1; p < n_lab; ++p) {
        uint32_t* row = &L[(size_t)p * n];
        std::copy(grouping, grouping + n, row);
        const uint64_t key = mix64(seed ^ mix64(p));
        for (uint32_t i = n - 1; i > 0; --i) {
            const uint32_t j = (uint32_t)(mix64(key + i) % (i + 1));
            std::swap(row[i], row[j]);
        }
    }
    Buf dD, dL, dinv, dpart;
    if (!upload_matrix<T>(dD, mat, n, err) || !alloc(dL, L.size() * 4, err) || !alloc(dinv, groups * 8, err))
        return false;
    ST_HIP(hipMemcpy(dL.p, L.data(), L.size() * 4, hipMemcpyHostToDevice));
    ST_HIP(hipMemcpy(dinv.p, inv.data(), groups * 8, hipMemcpyHostToDevice));
    const uint32_t rbs = (n + PM_R - 1) / PM_R, pbs = (n_lab + PM_P - 1) / PM_P;
    if (!alloc(dpart, (size_t)n_lab * rbs * 8, err)) return false;
    hipLaunchKernelGGL((k_permanova<T>), dim3(rbs, pbs), dim3(ST_T), 0, 0, dpart.as<double>(), dD.as<const T>(),
                       dL.as<const uint32_t>(), dinv.as<const double>(), n, n_lab);
    ST_HIP(hipGetLastError());
    std::vector<double> part((size_t)n_lab * rbs);
    ST_HIP(hipMemcpy(part.data(), dpart.p, part.size() * 8, hipMemcpyDeviceToHost));
    // s_T over the upper triangle
    Buf dr, du;
    if (!alloc(dr, n * 8, err) || !alloc(du, n * 8, err)) return false;
    hipLaunchKernelGGL((k_row_stats<T>), dim3(n), dim3(ST_T), 0, 0, dr.as<double>(), du.as<double>(),
                       dD.as<const T>(), n);
    ST_HIP(hipGetLastError());
    std::vector<double> upper(n);
    ST_HIP(hipMemcpy(upper.data(), du.p, n * 8, hipMemcpyDeviceToHost));
    double sT = 0.0;
    for (double u : upper) sT += u;
    sT /= n;
    auto f_of = [&](uint32_t p) {
        double sW = 0.0;
        for (uint32_t b = 0; b < rbs; ++b) sW += part[(size_t)p * rbs + b];
        return ((sT - sW) / (present - 1)) / (sW / (n - present));
    };
    const double f0 = f_of(0);
    uint32_t ge = 0;
    for (uint32_t p = 1; p < n_lab; ++p) ge += f_of(p) >= f0;
    *fstat = f0;
    *pvalue = n_perm ? (double)(ge + 1) / (double)(n_perm + 1) : std::numeric_limits<double>::quiet_NaN();
    return true;
}

}  // namespace

// ---------------------------------------------------------------- allocation-free centring
// The device scratch of center_matrix_noalloc: the row means of one call at
// [0, n) and its grand mean at [NOALLOC_MAX_N].  Each GPU holds its own copy
// for the life of the process; every call on a GPU overwrites it, so calls
// that overlap on one GPU race.
static __device__ __align__(256) double scratch[NOALLOC_MAX_N + 32];

namespace {

constexpr int SCRATCH_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
double* g_scratch_dev[SCRATCH_MAX_DEVICES] = {};   // per device: the address of its copy of `scratch`

// the current device's copy of `scratch` (looked up once per device)
bool scratch_of_device(double** out, std::string* err) {
    int dev = 0;
    ST_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= SCRATCH_MAX_DEVICES) {
        if (err) *err = "device index " + std::to_string(dev) + " beyond the scratch table";
        return false;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void* ptr = nullptr;
        ST_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        g_scratch_dev[dev] = (double*)ptr;
    }
    *out = g_scratch_dev[dev];
    return true;
}

// rm[n_max] = sum_i rm[i] / n, the row means added in index order (as center_on_device does on the host)
__global__ void k_grand_mean(double* __restrict__ rm, uint32_t n) {
    if (blockIdx.x || threadIdx.x) return;
    double g = 0.0;
    for (uint32_t i = 0; i < n; ++i) g += rm[i];
    rm[NOALLOC_MAX_N] = g / n;
}

// C_ij = -D_ij^2 / 2 - rmean_i - rmean_j + grand (the arithmetic of k_center_apply), C may be D
__global__ void __launch_bounds__(ST_T) k_center_copy(double* C, const double* D, const double* __restrict__ rm,
                                                      uint32_t n) {
    const size_t q = (size_t)blockIdx.x * ST_T + threadIdx.x;
    if (q >= (size_t)n * n) return;
    const uint32_t i = (uint32_t)(q / n), j = (uint32_t)(q % n);
    const double d = D[q];
    C[q] = -0.5 * d * d - rm[i] - rm[j] + rm[NOALLOC_MAX_N];
}

}  // namespace

bool center_matrix_noalloc(const double* d_mat, uint32_t n, double* d_centred, void* stream, std::string* err) {
    if (n > NOALLOC_MAX_N) {
        if (err) *err = "ssu_center_matrix_fp64_noalloc: n = " + std::to_string(n) + " exceeds its maximum of " +
                        std::to_string(NOALLOC_MAX_N) + " samples";
        return false;
    }
    if (n == 0) return true;
    double* rm = nullptr;
    if (!scratch_of_device(&rm, err)) return false;
    hipStream_t st = (hipStream_t)stream;
    hipLaunchKernelGGL((k_row_stats<double>), dim3(n), dim3(ST_T), 0, st, rm, nullptr, d_mat, n);
    ST_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_grand_mean, dim3(1), dim3(64), 0, st, rm, n);
    ST_HIP(hipGetLastError());
    const size_t nn = (size_t)n * n;
    hipLaunchKernelGGL(k_center_copy, dim3((uint32_t)((nn + ST_T - 1) / ST_T)), dim3(ST_T), 0, st, d_centred, d_mat,
                       rm, n);
    ST_HIP(hipGetLastError());
    return true;
}

bool center_matrix(const void* mat, bool in_fp64, uint32_t n, double* centred, std::string* err) {
    if (n == 0) return true;
    Buf dB;
    std::vector<double> rmean, upper;
    double grand = 0.0;
    const size_t nn = (size_t)n * n;
    if (in_fp64) {
        if (!center_on_device<double>(dB, mat, n, rmean, grand, upper, err)) return false;
        ST_HIP(hipMemcpy(centred, dB.p, nn * 8, hipMemcpyDeviceToHost));
    } else {
        if (!center_on_device<float>(dB, mat, n, rmean, grand, upper, err)) return false;
        std::vector<float> tmp(nn);
        ST_HIP(hipMemcpy(tmp.data(), dB.p, nn * 4, hipMemcpyDeviceToHost));
        for (size_t q = 0; q < nn; ++q) centred[q] = tmp[q];
    }
    return true;
}

bool pcoa(const void* mat, bool in_fp64, uint32_t n, uint32_t dims, uint64_t seed, double* eigenvalues,
          double* samples, double* proportion, std::string* err) {
    return in_fp64 ? pcoa_t<double>(mat, n, dims, seed, eigenvalues, samples, proportion, err)
                   : pcoa_t<float>(mat, n, dims, seed, eigenvalues, samples, proportion, err);
}

bool permanova(const void* mat, bool in_fp64, uint32_t n, const uint32_t* grouping, uint32_t n_perm, uint64_t seed,
               double* fstat, double* pvalue, std::string* err) {
    return in_fp64 ? permanova_t<double>(mat, n, grouping, n_perm, seed, fstat, pvalue, err)
                   : permanova_t<float>(mat, n, grouping, n_perm, seed, fstat, pvalue, err);
}

}  // namespace ssu
