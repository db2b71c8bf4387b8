// Partial-result files and their merge (the C ABI of src/api.hpp:857-991).
//
// File layout (src/api.hpp:816-856 documents it; src/api.cpp:1972-2055 writes it):
//   uint32 header[8] = {0x088ABA02, n_samples, n_stripes, stripe_start,
//                       stripe_total, is_upper_triangle, ids_bytes, ids_bytes_lz4}
//   LZ4 block of the sample ids, each NUL-terminated, ids_bytes long decoded
//   per stripe: uint32 size, LZ4 block of n_samples doubles
//   uint32 footer = 0x088ABA02
// All integers little-endian, as written by the x86-64 reference.
#include <fcntl.h>
#include <unistd.h>

#include <algorithm>
#include <atomic>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "ssu.h"
#include "ssu_lz4.hpp"
#include "ssu_results.hpp"

namespace {

constexpr uint32_t MAGIC_V2 = PARTIAL_MAGIC_V2;  // src/api.hpp:23

struct Fd {
    int fd = -1;
    explicit Fd(int f) : fd(f) {}
    ~Fd() {
        if (fd >= 0) close(fd);
    }
    Fd(const Fd&) = delete;
    Fd& operator=(const Fd&) = delete;
};

bool write_all(int fd, const void* p, size_t n) {
    const char* c = (const char*)p;
    while (n) {
        const ssize_t w = write(fd, c, n);
        if (w <= 0) return false;
        c += w;
        n -= (size_t)w;
    }
    return true;
}

bool read_all(int fd, void* p, size_t n) {
    char* c = (char*)p;
    while (n) {
        const ssize_t r = read(fd, c, n);
        if (r <= 0) return false;
        c += r;
        n -= (size_t)r;
    }
    return true;
}

// Fields common to partial_mat_t and partial_dyn_mat_t, parsed from the header
// and the sample-id block; the file offset is left at the first stripe.
struct Header {
    uint32_t n_samples = 0, n_stripes = 0, stripe_start = 0, stripe_total = 0;
    bool upper = false;
    std::vector<std::string> ids;
};

// src/api.cpp:2071-2137 (read_partial_header_fd): same checks, same codes
IOStatus read_header(int fd, Header& h) {
    uint32_t w[8];
    if (!read_all(fd, w, sizeof(w))) return magic_incompatible;
    if (w[0] != MAGIC_V2) return magic_incompatible;
    h.n_samples = w[1];
    h.n_stripes = w[2];
    h.stripe_start = w[3];
    h.stripe_total = w[4];
    h.upper = w[5] != 0;
    if (h.n_samples == 0 || h.n_stripes == 0 || h.stripe_total == 0 || !h.upper) return bad_header;
    if (h.stripe_total >= h.n_samples || h.n_stripes > h.stripe_total || h.stripe_start >= h.stripe_total ||
        (uint64_t)h.stripe_start + h.n_stripes > h.stripe_total)
        return bad_header;
    const uint32_t raw = w[6], packed = w[7];
    if (raw == 0 || packed == 0) return bad_header;
    std::vector<uint8_t> cbuf(packed), buf(raw);
    if (!read_all(fd, cbuf.data(), packed)) return magic_incompatible;
    if (ssu::lz4_decompress(cbuf.data(), packed, buf.data(), raw) != (long)raw) return magic_incompatible;
    h.ids.clear();
    h.ids.reserve(h.n_samples);
    size_t pos = 0;
    for (uint32_t i = 0; i < h.n_samples; ++i) {
        const void* nul = memchr(buf.data() + pos, 0, raw - pos);
        if (!nul) return magic_incompatible;
        const size_t len = (const uint8_t*)nul - (buf.data() + pos);
        h.ids.emplace_back((const char*)buf.data() + pos, len);
        pos += len + 1;
    }
    return read_okay;
}

// one stripe record at the current offset: size, LZ4 block -> n doubles
IOStatus read_stripe(int fd, uint32_t n_samples, double** out) {
    uint32_t size = 0;
    if (!read_all(fd, &size, sizeof(size))) return magic_incompatible;
    const size_t raw = sizeof(double) * (size_t)n_samples;
    if (size == 0 || size > ssu::lz4_bound(raw)) return magic_incompatible;
    std::vector<uint8_t> cbuf(size);
    if (!read_all(fd, cbuf.data(), size)) return magic_incompatible;
    double* v = (double*)malloc(raw);
    if (!v) return bad_header;
    if (ssu::lz4_decompress(cbuf.data(), size, (uint8_t*)v, raw) != (long)raw) {
        free(v);
        return magic_incompatible;
    }
    *out = v;
    return read_okay;
}

template <class TPMat>
void fill_common(TPMat* r, const Header& h) {
    r->n_samples = h.n_samples;
    r->sample_ids = ssu::dup_ids(h.ids);
    r->stripe_start = h.stripe_start;
    r->stripe_stop = h.stripe_start + h.n_stripes;
    r->stripe_total = h.stripe_total;
    r->is_upper_triangle = h.upper;
}

// src/api.cpp:2322-2394 (check_partial); returns codes instead of exiting
template <class TPMat>
MergeStatus check_partials(const TPMat* const* mats, int n_partials, bool verbose) {
    if (n_partials <= 0 || !mats || !mats[0]) {
        if (verbose) fprintf(stderr, "Zero or less partials.\n");
        return incomplete_stripe_set;
    }
    const uint32_t n = mats[0]->n_samples;
    const uint32_t total = mats[0]->stripe_total;
    std::vector<char> seen(total, 0);
    uint64_t count = 0;
    for (int i = 0; i < n_partials; ++i) {
        const TPMat* p = mats[i];
        if (p->n_samples != n) {
            if (verbose) fprintf(stderr, "Wrong number of samples in %i, %i!=%i\n", i, (int)p->n_samples, (int)n);
            return partials_mismatch;
        }
        if (p->stripe_total != total) {
            if (verbose) fprintf(stderr, "Wrong number of stripes in %i, %i!=%i\n", i, (int)total, (int)p->stripe_total);
            return partials_mismatch;
        }
        if (p->is_upper_triangle != mats[0]->is_upper_triangle) {
            if (verbose)
                fprintf(stderr, "Wrong number of is_upper_triangle in %i, %i!=%i\n", i,
                        (int)mats[0]->is_upper_triangle, (int)p->is_upper_triangle);
            return square_mismatch;
        }
        for (uint32_t j = 0; j < n; ++j)
            if (strcmp(mats[0]->sample_ids[j], p->sample_ids[j]) != 0) {
                if (verbose)
                    fprintf(stderr, "Wrong number of sample id %i in %i, %s!=%s\n", (int)j, i, mats[0]->sample_ids[j],
                            p->sample_ids[j]);
                return sample_id_consistency;
            }
        if (p->stripe_stop > total || p->stripe_start > p->stripe_stop) {
            if (verbose) fprintf(stderr, "Stripe range of %i outside the matrix\n", i);
            return partials_mismatch;
        }
        for (uint32_t s = p->stripe_start; s < p->stripe_stop; ++s) {
            if (seen[s]) {
                if (verbose) fprintf(stderr, "Overlap in %i vs %i\n", i, (int)s);
                return stripes_overlap;
            }
            seen[s] = 1;
            ++count;
        }
    }
    if (count != total) {
        if (verbose) fprintf(stderr, "Insufficient number of stripes found, %i!=%i\n", (int)count, (int)total);
        return incomplete_stripe_set;
    }
    return merge_okay;
}

// Stripe s holds pair (k, (k+s+1) mod n) at position k (src/unifrac_task_impl.hpp:367);
// some pairs are held twice (the last stripe of either parity).  Upper-triangle
// entry (i, j), d = j - i, takes stripe d-1 at i when d-1 < n_stripes and
// stripe n-d-1 at j otherwise (src/unifrac.cpp:228-340).  So a wrapped
// position (k > l) is used only when its pair has no unwrapped copy.
//
// Stripes are decompressed and placed in blocks of SB on several threads: each
// stripe holds its own pairs, so the blocks write disjoint entries.  Each
// partial's stripe offsets are learned first from the size words (pread, one
// thread per partial); stripes a caller already loaded are used as they are,
// and every loaded stripe is freed once placed, as the one-thread merge did.
template <class TReal>
bool stripes_into_matrix(partial_dyn_mat_t** mats, int n_partials, uint32_t n, uint32_t total, TReal* M) {
    for (uint64_t i = 0; i < n; ++i) M[i * n + i] = 0;
    std::vector<int> owner(total, -1);
    for (int p = 0; p < n_partials; ++p)
        for (uint32_t s = mats[p]->stripe_start; s < mats[p]->stripe_stop; ++s) owner[s] = p;
    std::vector<int> fds(n_partials, -1);
    struct CloseAll {
        std::vector<int>& f;
        ~CloseAll() {
            for (int x : f)
                if (x >= 0) close(x);
        }
    } close_all{fds};
    for (int p = 0; p < n_partials; ++p)
        if ((fds[p] = open(mats[p]->filename, O_RDONLY)) < 0) return false;
    // offsets of every stripe record, walking forward from the first
    std::vector<uint8_t> offsets_ok(n_partials, 1);
    {
        std::vector<std::thread> th;
        for (int p = 0; p < n_partials; ++p)
            th.emplace_back([&, p] {
                partial_dyn_mat_t* pm = mats[p];
                const uint32_t ns = pm->stripe_stop - pm->stripe_start;
                for (uint32_t c = 0; c + 1 < ns; ++c) {
                    if (pm->offsets[c + 1]) continue;
                    uint32_t size = 0;
                    if (pread(fds[p], &size, sizeof(size), (off_t)pm->offsets[c]) != (ssize_t)sizeof(size)) {
                        offsets_ok[p] = 0;
                        return;
                    }
                    pm->offsets[c + 1] = pm->offsets[c] + sizeof(uint32_t) + size;
                }
            });
        for (auto& t : th) t.join();
    }
    for (uint8_t ok : offsets_ok)
        if (!ok) return false;
    constexpr uint32_t SB = 64;   // stripes placed together
    constexpr uint32_t KB = 64;   // samples per tile
    const size_t raw = sizeof(double) * (size_t)n;
    const uint32_t blocks = (total + SB - 1) / SB;
    const unsigned threads =
        std::max(1u, std::min({16u, std::thread::hardware_concurrency(), blocks}));
    std::atomic<uint32_t> next{0};
    std::atomic<bool> failed{false};
    auto work = [&] {
        std::vector<double> buf((size_t)SB * n);
        std::vector<uint8_t> cbuf;
        std::vector<const double*> rows(SB);
        for (uint32_t blk; !failed.load(std::memory_order_relaxed) && (blk = next.fetch_add(1)) < blocks;) {
            const uint32_t s0 = blk * SB, s1 = std::min(total, s0 + SB);
            for (uint32_t s = s0; s < s1; ++s) {
                partial_dyn_mat_t* pm = mats[owner[s]];
                const uint32_t idx = s - pm->stripe_start;
                if (pm->stripes[idx]) {
                    rows[s - s0] = pm->stripes[idx];
                    continue;
                }
                const int fd = fds[owner[s]];
                uint32_t size = 0;
                double* v = buf.data() + (size_t)(s - s0) * n;
                if (pread(fd, &size, sizeof(size), (off_t)pm->offsets[idx]) != (ssize_t)sizeof(size) || size == 0 ||
                    size > ssu::lz4_bound(raw)) {
                    failed = true;
                    return;
                }
                cbuf.resize(size);
                if (pread(fd, cbuf.data(), size, (off_t)(pm->offsets[idx] + sizeof(uint32_t))) != (ssize_t)size ||
                    ssu::lz4_decompress(cbuf.data(), size, (uint8_t*)v, raw) != (long)raw) {
                    failed = true;
                    return;
                }
                rows[s - s0] = v;
            }
            for (uint32_t k0 = 0; k0 < n; k0 += KB) {
                const uint32_t k1 = std::min(n, k0 + KB);
                for (uint32_t s = s0; s < s1; ++s) {
                    const double* v = rows[s - s0];
                    const uint32_t d = s + 1;
                    const bool wrapped_dup = n - d - 1 < total;   // the pair's distance n-d is also stored unwrapped
                    for (uint32_t k = k0; k < k1; ++k) {
                        const uint32_t l = (uint32_t)(((uint64_t)k + d) % n);
                        if (k > l && wrapped_dup) continue;
                        const TReal x = (TReal)v[k];
                        M[(uint64_t)k * n + l] = x;
                        M[(uint64_t)l * n + k] = x;
                    }
                }
            }
            for (uint32_t s = s0; s < s1; ++s) {
                partial_dyn_mat_t* pm = mats[owner[s]];
                const uint32_t idx = s - pm->stripe_start;
                free(pm->stripes[idx]);
                pm->stripes[idx] = nullptr;
            }
        }
    };
    std::vector<std::thread> pool;
    for (unsigned t = 1; t < threads; ++t) pool.emplace_back(work);
    work();
    for (auto& t : pool) t.join();
    return !failed;
}

// src/api.cpp:2448-2475 (merge_partial_to_matrix_T)
template <class TReal, class TMat>
MergeStatus merge_partials(partial_dyn_mat_t** mats, int n_partials, const char* mmap_dir, TMat** result) {
    const MergeStatus err = check_partials(mats, n_partials, false);
    if (err != merge_okay) return err;
    const uint32_t n = mats[0]->n_samples;
    std::vector<std::string> ids(mats[0]->sample_ids, mats[0]->sample_ids + n);
    if (!ssu::init_full<TReal, TMat>(*result, ids, mmap_dir)) {
        ssu::destroy_full<TReal, TMat>(result);
        return incomplete_stripe_set;
    }
    if (!stripes_into_matrix<TReal>(mats, n_partials, n, mats[0]->stripe_total, (*result)->matrix)) {
        ssu::set_error("merge: a stripe could not be read back from its partial file");
        ssu::destroy_full<TReal, TMat>(result);
        return incomplete_stripe_set;
    }
    return merge_okay;
}

}  // namespace

// src/api.cpp:1972-2055
IOStatus write_partial(const char* filename, const partial_mat_t* result) {
    if (!filename || !result) return write_error;
    Fd f(open(filename, O_WRONLY | O_CREAT | O_TRUNC, S_IRUSR | S_IWUSR));
    if (f.fd < 0) return write_error;
    const uint32_t n_stripes = result->stripe_stop - result->stripe_start;
    std::string ids;
    for (uint32_t i = 0; i < result->n_samples; ++i) {
        ids.append(result->sample_ids[i]);
        ids.push_back('\0');
    }
    std::vector<uint8_t> cbuf(ssu::lz4_bound(std::max(ids.size(), sizeof(double) * (size_t)result->n_samples)) + 4);
    const size_t packed = ssu::lz4_compress((const uint8_t*)ids.data(), ids.size(), cbuf.data(), cbuf.size());
    if (packed == 0) return open_error;
    const uint32_t header[8] = {MAGIC_V2, result->n_samples, n_stripes, result->stripe_start,
                                result->stripe_total, (uint32_t)(result->is_upper_triangle ? 1 : 0),
                                (uint32_t)ids.size(), (uint32_t)packed};
    if (!write_all(f.fd, header, sizeof(header)) || !write_all(f.fd, cbuf.data(), packed)) return write_error;
    const size_t raw = sizeof(double) * (size_t)result->n_samples;
    for (uint32_t s = 0; s < n_stripes; ++s) {
        const size_t c = ssu::lz4_compress((const uint8_t*)result->stripes[s], raw, cbuf.data() + 4, cbuf.size() - 4);
        if (c == 0) return open_error;
        const uint32_t c32 = (uint32_t)c;
        memcpy(cbuf.data(), &c32, 4);
        if (!write_all(f.fd, cbuf.data(), c + 4)) return write_error;
    }
    const uint32_t footer = MAGIC_V2;
    if (!write_all(f.fd, &footer, sizeof(footer))) return open_error;
    return write_okay;
}

// src/api.cpp:2247-2282
IOStatus read_partial(const char* filename, partial_mat_t** result_out) {
    if (!result_out) return read_error;
    *result_out = nullptr;
    if (!filename) return open_error;
    Fd f(open(filename, O_RDONLY));
    if (f.fd < 0) return open_error;
    Header h;
    IOStatus st = read_header(f.fd, h);
    if (st != read_okay) return st;
    partial_mat_t* r = (partial_mat_t*)calloc(1, sizeof(partial_mat_t));
    fill_common(r, h);
    r->stripes = (double**)calloc(h.n_stripes, sizeof(double*));
    for (uint32_t s = 0; s < h.n_stripes && st == read_okay; ++s) st = read_stripe(f.fd, h.n_samples, &r->stripes[s]);
    if (st == read_okay) {
        uint32_t footer = 0;
        if (!read_all(f.fd, &footer, sizeof(footer)) || footer != MAGIC_V2) st = magic_incompatible;
    }
    if (st != read_okay) {
        destroy_partial_mat(&r);
        return st;
    }
    *result_out = r;
    return read_okay;
}

// src/api.cpp:2284-2307: header and sample ids now, stripes on demand
IOStatus read_partial_header(const char* filename, partial_dyn_mat_t** result_out) {
    if (!result_out) return read_error;
    *result_out = nullptr;
    if (!filename) return open_error;
    Fd f(open(filename, O_RDONLY));
    if (f.fd < 0) return open_error;
    Header h;
    const IOStatus st = read_header(f.fd, h);
    if (st != read_okay) return st;
    partial_dyn_mat_t* r = (partial_dyn_mat_t*)calloc(1, sizeof(partial_dyn_mat_t));
    fill_common(r, h);
    r->stripes = (double**)calloc(h.n_stripes, sizeof(double*));
    r->offsets = (uint64_t*)calloc(h.n_stripes, sizeof(uint64_t));
    r->offsets[0] = (uint64_t)lseek(f.fd, 0, SEEK_CUR);
    r->filename = strdup(filename);
    *result_out = r;
    return read_okay;
}

// src/api.cpp:2184-2245, 2309-2320: offsets are learned walking forward from
// the nearest known one; an already loaded stripe is not read again
IOStatus read_partial_one_stripe(partial_dyn_mat_t* result, uint32_t stripe_idx) {
    if (!result || stripe_idx >= result->stripe_stop - result->stripe_start) return bad_header;
    if (result->stripes[stripe_idx]) return read_okay;
    Fd f(open(result->filename, O_RDONLY));
    if (f.fd < 0) return open_error;
    uint32_t cur = stripe_idx;
    while (result->offsets[cur] == 0) --cur;   // offsets[0] is always known
    for (; cur < stripe_idx; ++cur) {
        if (lseek(f.fd, (off_t)result->offsets[cur], SEEK_SET) != (off_t)result->offsets[cur]) return bad_header;
        uint32_t size = 0;
        if (!read_all(f.fd, &size, sizeof(size))) return magic_incompatible;
        result->offsets[cur + 1] = result->offsets[cur] + sizeof(uint32_t) + size;
    }
    if (lseek(f.fd, (off_t)result->offsets[stripe_idx], SEEK_SET) != (off_t)result->offsets[stripe_idx]) return bad_header;
    return read_stripe(f.fd, result->n_samples, &result->stripes[stripe_idx]);
}

MergeStatus validate_partial(const partial_dyn_mat_t* const* partial_mats, int n_partials) {
    return check_partials(partial_mats, n_partials, true);
}

MergeStatus merge_partial_to_matrix(partial_dyn_mat_t** partial_mats, int n_partials, mat_full_fp64_t** result) {
    return merge_partials<double, mat_full_fp64_t>(partial_mats, n_partials, nullptr, result);
}

MergeStatus merge_partial_to_matrix_fp32(partial_dyn_mat_t** partial_mats, int n_partials, mat_full_fp32_t** result) {
    return merge_partials<float, mat_full_fp32_t>(partial_mats, n_partials, nullptr, result);
}

MergeStatus merge_partial_to_mmap_matrix(partial_dyn_mat_t** partial_mats, int n_partials, const char* mmap_dir,
                                         mat_full_fp64_t** result) {
    return merge_partials<double, mat_full_fp64_t>(partial_mats, n_partials, mmap_dir, result);
}

MergeStatus merge_partial_to_mmap_matrix_fp32(partial_dyn_mat_t** partial_mats, int n_partials, const char* mmap_dir,
                                              mat_full_fp32_t** result) {
    return merge_partials<float, mat_full_fp32_t>(partial_mats, n_partials, mmap_dir, result);
}
