// Host orchestration of the device pipeline: schedule building, uploads,
// buffer management and the per-method kernel sequence.
#include "ssu_engine.hpp"
#include "ssu_kernels.hpp"
#include "ssu_tune.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <functional>
#include <map>
#include <mutex>
#include <thread>

namespace ssu {

namespace {

#define SSU_HIP(call)                                                                        \
    do {                                                                                     \
        hipError_t e_ = (call);                                                              \
        if (e_ != hipSuccess) {                                                              \
            if (err) *err = std::string(#call) + " failed: " + hipGetErrorString(e_);        \
            return false;                                                                    \
        }                                                                                    \
    } while (0)

size_t round_up(size_t x, size_t m) { return (x + m - 1) / m * m; }

// Gram path occupancy split: smallest table it is used for, and the device
// memory its posting lists, row lists, prefix counts and sample bits may take
constexpr uint32_t GSPLIT_MIN_N = 1024;
constexpr double GSPLIT_BUDGET = 8.0 * (1u << 30);

struct DevMem {
    void* p = nullptr;
    size_t bytes = 0;
    DevMem() = default;
    DevMem(const DevMem&) = delete;
    DevMem& operator=(const DevMem&) = delete;
    ~DevMem() { release(); }
    void take(DevMem& o) {   // o's memory, o left empty
        release();
        p = o.p;
        bytes = o.bytes;
        o.p = nullptr;
        o.bytes = 0;
    }
    void release() {
        if (p) pool_free(p, bytes);
        p = nullptr;
        bytes = 0;
    }
    bool ensure(size_t b, std::string* err) {
        if (b <= bytes && p) return true;
        release();
        p = pool_alloc(std::max<size_t>(b, 16), &bytes, err);
        return p != nullptr;
    }
    template <class T> T* as() const { return (T*)p; }
};

template <class T>
bool upload(DevMem& d, const std::vector<T>& h, hipStream_t st, std::string* err) {
    if (!d.ensure(h.size() * sizeof(T), err)) return false;
    if (!h.empty()) SSU_HIP(hipMemcpyAsync(d.p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, st));
    return true;
}

// (through the pinned staging area: returns once the host buffer has been read)
bool upload_raw(DevMem& d, const void* h, size_t bytes, hipStream_t st, std::string* err) {
    if (!d.ensure(bytes, err)) return false;
    return copy_to_device(d.p, h, bytes, st, err);
}

}  // namespace

// Device memory cache.  hipMalloc and hipFree of the multi-gigabyte buffers a
// plan holds cost far more than the plan's kernels (they map and unmap pages),
// so freed blocks stay with the process and later plans of similar sizes reuse
// them: sizes are rounded to 2 MiB, a request takes the smallest cached block
// of its device that is large enough and at most a quarter larger, and the
// cache keeps at most SSU_POOL_GIB GiB (default: 90% of the first device's
// memory, so that the plans of a 65,536-sample matrix stay cached between
// calls; 0 turns caching off), releasing its largest blocks beyond that.  A
// failed hipMalloc empties the cache and tries once more.
// ssu_release_device_memory() empties it.
namespace {
struct Pool {
    std::mutex mu;
    std::multimap<size_t, std::pair<int, void*>> free_blocks;   // size -> (device, block)
    size_t cached = 0;
    size_t cap = 0;
    bool cap_set = false;
};
Pool& pool() {
    static Pool* p = new Pool();   // never destroyed: the runtime may already be gone at exit
    return *p;
}
size_t pool_cap(Pool& P) {
    if (!P.cap_set) {
        double gib = 96.0;
        size_t free_b = 0, total_b = 0;
        if (hipMemGetInfo(&free_b, &total_b) == hipSuccess && total_b > 0) gib = 0.9 * (double)total_b / (1ull << 30);
        else (void)hipGetLastError();
        if (const char* env = getenv("SSU_POOL_GIB")) gib = atof(env);
        P.cap = gib > 0 ? (size_t)(gib * (double)(1ull << 30)) : 0;
        P.cap_set = true;
    }
    return P.cap;
}
void pool_trim_locked(Pool& P, size_t keep) {
    while (P.cached > keep && !P.free_blocks.empty()) {
        auto it = std::prev(P.free_blocks.end());   // largest first
        int cur = 0;
        (void)hipGetDevice(&cur);
        if (it->second.first != cur) (void)hipSetDevice(it->second.first);
        (void)hipFree(it->second.second);
        if (it->second.first != cur) (void)hipSetDevice(cur);
        P.cached -= it->first;
        P.free_blocks.erase(it);
    }
}
}  // namespace

void* pool_alloc(size_t bytes, size_t* got, std::string* err) {
    constexpr size_t G = size_t(2) << 20;
    const size_t rb = (bytes + G - 1) / G * G;
    int dev = 0;
    (void)hipGetDevice(&dev);
    Pool& P = pool();
    {
        std::lock_guard<std::mutex> lock(P.mu);
        for (auto it = P.free_blocks.lower_bound(rb); it != P.free_blocks.end() && it->first <= rb + rb / 4; ++it) {
            if (it->second.first != dev) continue;
            void* q = it->second.second;
            *got = it->first;
            P.cached -= it->first;
            P.free_blocks.erase(it);
            return q;
        }
    }
    void* q = nullptr;
    hipError_t e = hipMalloc(&q, rb);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        {
            std::lock_guard<std::mutex> lock(P.mu);
            pool_trim_locked(P, 0);
        }
        e = hipMalloc(&q, rb);
    }
    if (e != hipSuccess) {
        (void)hipGetLastError();
        if (err) *err = "hipMalloc(" + std::to_string(rb) + " bytes) failed: " + hipGetErrorString(e);
        *got = 0;
        return nullptr;
    }
    *got = rb;
    return q;
}

void pool_free(void* p, size_t bytes) {
    if (!p) return;
    Pool& P = pool();
    std::lock_guard<std::mutex> lock(P.mu);
    const size_t cap = pool_cap(P);
    if (bytes > cap) { (void)hipFree(p); return; }
    int dev = 0;
    (void)hipGetDevice(&dev);
    P.free_blocks.emplace(bytes, std::make_pair(dev, p));
    P.cached += bytes;
    pool_trim_locked(P, cap);
}

void pool_release_all() {
    Pool& P = pool();
    std::lock_guard<std::mutex> lock(P.mu);
    pool_trim_locked(P, 0);
}

// Pinned staging for host copies: STAGE_BUFS chunks, allocated on first use
// and kept for the life of the process (pinned allocations are slow to make).
namespace {
constexpr size_t STAGE_CHUNK = size_t(64) << 20;
constexpr int STAGE_BUFS = 4;
struct Stage {
    std::mutex mu;
    void* buf[STAGE_BUFS] = {};
    hipEvent_t done[STAGE_BUFS] = {};
    bool failed = false;
};
Stage& stage() {
    static Stage* s = new Stage();   // never destroyed: the runtime may already be gone at exit
    return *s;
}
// the staging buffers, made on first use (false: pinned memory unavailable)
bool stage_ready(Stage& S) {
    if (S.failed) return false;
    if (S.buf[0]) return true;
    for (int b = 0; b < STAGE_BUFS; ++b)
        if (hipHostMalloc(&S.buf[b], STAGE_CHUNK, hipHostMallocDefault) != hipSuccess ||
            hipEventCreateWithFlags(&S.done[b], hipEventDisableTiming) != hipSuccess) {
            S.failed = true;   // plain copies from now on
            return false;
        }
    return true;
}

// Host work over a persistent pool of worker threads (started on first use, up
// to 16 with the caller): fn(part) for part = 0 .. parts() - 1, the calling
// thread taking the last part and waiting for the rest.
class CopyPool {
public:
    static CopyPool& get() {
        static CopyPool* p = new CopyPool();   // never joined: the workers wait for the process to end
        return *p;
    }
    unsigned parts() const { return workers_ + 1; }
    void run(const std::function<void(unsigned)>& fn) {
        if (workers_ == 0) { fn(0); return; }
        std::lock_guard<std::mutex> one(call_mu_);   // one job at a time on the pool
        std::unique_lock<std::mutex> lk(mu_);
        fn_ = &fn;
        pending_ = workers_;
        ++gen_;
        cv_.notify_all();
        lk.unlock();
        fn(workers_);
        lk.lock();
        done_.wait(lk, [&] { return pending_ == 0; });
        fn_ = nullptr;
    }
    void copy(char* dst, const char* src, size_t bytes) {
        if (workers_ == 0 || bytes < (size_t(4) << 20)) { std::memcpy(dst, src, bytes); return; }
        const size_t part = (bytes / parts() + 4095) / 4096 * 4096;
        run([&](unsigned w) {
            const size_t a = part * w;
            if (a < bytes) std::memcpy(dst + a, src + a, w + 1 == parts() ? bytes - a : std::min(part, bytes - a));
        });
    }
private:
    CopyPool() {
        const unsigned hc = std::thread::hardware_concurrency();
        workers_ = std::min(16u, std::max(hc, 1u)) - 1;
        for (unsigned w = 0; w < workers_; ++w) std::thread([this, w] { loop(w); }).detach();
    }
    void loop(unsigned w) {
        uint64_t seen = 0;
        std::unique_lock<std::mutex> lk(mu_);
        for (;;) {
            cv_.wait(lk, [&] { return gen_ != seen; });
            seen = gen_;
            const std::function<void(unsigned)>* fn = fn_;
            lk.unlock();
            (*fn)(w);
            lk.lock();
            if (--pending_ == 0) done_.notify_one();
        }
    }
    std::mutex call_mu_, mu_;
    std::condition_variable cv_, done_;
    unsigned workers_ = 0, pending_ = 0;
    uint64_t gen_ = 0;
    const std::function<void(unsigned)>* fn_ = nullptr;
};

// dst[0..bytes) = src[0..bytes) on the copy pool
void spread_copy(char* dst, const char* src, size_t bytes, unsigned) { CopyPool::get().copy(dst, src, bytes); }
}  // namespace

bool copy_ranges_to_host(char* host, const char* dev, const std::vector<CopyRange>& ranges, void* stream,
                         std::string* err, bool prefault) {
    hipStream_t st = (hipStream_t)stream;
    size_t bytes = 0;
    for (const CopyRange& r : ranges) bytes += r.len * r.rows;
    if (!bytes) return true;
    Stage& S = stage();
    std::unique_lock<std::mutex> lock(S.mu);
    if (bytes <= STAGE_CHUNK / 4 || !stage_ready(S)) {
        lock.unlock();
        for (const CopyRange& r : ranges) {
            for (void* e : r.ready) SSU_HIP(hipEventSynchronize((hipEvent_t)e));
            SSU_HIP(hipStreamSynchronize(st));
            if (!r.len || !r.rows) continue;
            if (r.rows > 1)
                SSU_HIP(hipMemcpy2D(host + r.off, r.pitch, dev + r.off, r.pitch, r.len, r.rows, hipMemcpyDeviceToHost));
            else
                SSU_HIP(hipMemcpy(host + r.off, dev + r.off, r.len, hipMemcpyDeviceToHost));
        }
        return true;
    }
    // pieces of at most STAGE_CHUNK bytes, each within one range: runs of a 1-D
    // range, or whole rows of a rectangle (a row wider than a staging buffer is
    // taken as 1-D runs)
    struct Piece { size_t off, len, rows, pitch; uint32_t range; };
    std::vector<Piece> pieces;
    for (uint32_t r = 0; r < ranges.size(); ++r) {
        const CopyRange& R = ranges[r];
        if (!R.len || !R.rows) continue;
        if (R.rows > 1 && R.len <= STAGE_CHUNK) {
            const size_t per = STAGE_CHUNK / R.len;
            for (size_t q = 0; q < R.rows; q += per)
                pieces.push_back({R.off + q * R.pitch, R.len, std::min(per, R.rows - q), R.pitch, r});
            continue;
        }
        for (size_t row = 0; row < R.rows; ++row)
            for (size_t o = 0; o < R.len; o += STAGE_CHUNK)
                pieces.push_back({R.off + row * R.pitch + o, std::min(STAGE_CHUNK, R.len - o), 1, 0, r});
    }
    static const bool timing = getenv("SSU_TIMING") && atoi(getenv("SSU_TIMING")) != 0;
    const auto t0 = std::chrono::steady_clock::now();
    double wait_ms = 0.0, host_ms = 0.0;
    auto since = [](std::chrono::steady_clock::time_point a) {
        return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - a).count();
    };
    // a piece's copy waits on the device for its range's events, then fills buffer c % STAGE_BUFS
    auto issue = [&](size_t c) -> hipError_t {
        const Piece& q = pieces[c];
        const int b = (int)(c % STAGE_BUFS);
        for (void* e : ranges[q.range].ready) {
            const hipError_t x = hipStreamWaitEvent(st, (hipEvent_t)e, 0);
            if (x != hipSuccess) return x;
        }
        hipError_t x = q.rows > 1 ? hipMemcpy2DAsync(S.buf[b], q.len, dev + q.off, q.pitch, q.len, q.rows,
                                                     hipMemcpyDeviceToHost, st)
                                  : hipMemcpyAsync(S.buf[b], dev + q.off, q.len, hipMemcpyDeviceToHost, st);
        return x != hipSuccess ? x : hipEventRecord(S.done[b], st);
    };
    const size_t np = pieces.size();
    for (size_t c = 0; c < std::min<size_t>(np, STAGE_BUFS); ++c) SSU_HIP(issue(c));
    // While the first piece is still on its way (its range computing), touch the
    // destination's pages on the pool: first-touch page faults of a freshly
    // allocated result would otherwise slow every host copy
    size_t touched = 0, span = 0;
    double touch_ms = 0.0;
    if (prefault && !ranges.empty()) {
        size_t lo = ranges.front().off;
        for (const CopyRange& r : ranges) lo = std::min(lo, r.off);
        for (const CopyRange& r : ranges)
            if (r.len && r.rows) span = std::max(span, r.off + (r.rows - 1) * r.pitch + r.len - lo);
        const auto tt = std::chrono::steady_clock::now();
        constexpr size_t TOUCH = size_t(256) << 20, PAGE = 4096;
        CopyPool& pool = CopyPool::get();
        while (touched < span && hipEventQuery(S.done[0]) == hipErrorNotReady) {
            const size_t len = std::min(TOUCH, span - touched);
            char* base = host + lo + touched;
            const size_t pages = (len + PAGE - 1) / PAGE, parts = pool.parts();
            pool.run([&](unsigned t) {
                for (size_t q = pages * t / parts; q < pages * (t + 1) / parts; ++q)
                    *reinterpret_cast<volatile char*>(base + q * PAGE) = 0;
            });
            touched += len;
        }
        touch_ms = since(tt);
    }
    for (size_t c = 0; c < np; ++c) {
        const int b = (int)(c % STAGE_BUFS);
        auto tw = std::chrono::steady_clock::now();
        SSU_HIP(hipEventSynchronize(S.done[b]));
        wait_ms += since(tw);
        tw = std::chrono::steady_clock::now();
        const Piece& q = pieces[c];
        if (q.rows > 1) {   // the rows, packed in the buffer, to their places
            CopyPool& pool = CopyPool::get();
            const size_t parts = pool.parts();
            const char* src = (const char*)S.buf[b];
            pool.run([&](unsigned t) {
                for (size_t r = q.rows * t / parts; r < q.rows * (t + 1) / parts; ++r)
                    memcpy(host + q.off + r * q.pitch, src + r * q.len, q.len);
            });
        } else {
            spread_copy(host + q.off, (const char*)S.buf[b], q.len, 0);
        }
        host_ms += since(tw);
        if (c + STAGE_BUFS < np) SSU_HIP(issue(c + STAGE_BUFS));   // into the buffer just emptied
    }
    if (timing)
        fprintf(stderr,
                "[ssu copy] %8.1f MB in %zu pieces, %7.2f ms (waiting %7.2f ms, host copies %7.2f ms, %.1f MB of "
                "pages touched first in %.2f ms)\n",
                bytes / 1e6, np, since(t0), wait_ms, host_ms, touched / 1e6, touch_ms);
    return true;
}

bool copy_to_host(void* host, const void* dev, size_t bytes, void* stream, std::string* err) {
    return copy_ranges_to_host((char*)host, (const char*)dev, {CopyRange{0, bytes, {}}}, stream, err);
}

bool copy_to_device(void* dev, const void* host, size_t bytes, void* stream, std::string* err) {
    if (!bytes) return true;
    hipStream_t st = (hipStream_t)stream;
    Stage& S = stage();
    std::unique_lock<std::mutex> lock(S.mu);
    if (bytes <= STAGE_CHUNK / 4 || !stage_ready(S)) {
        lock.unlock();
        SSU_HIP(hipMemcpyAsync(dev, host, bytes, hipMemcpyHostToDevice, st));
        SSU_HIP(hipStreamSynchronize(st));   // the caller's buffer may go once this returns
        return true;
    }
    const size_t nch = (bytes + STAGE_CHUNK - 1) / STAGE_CHUNK;
    bool used[STAGE_BUFS] = {};
    for (size_t c = 0; c < nch; ++c) {
        const int b = (int)(c % STAGE_BUFS);
        const size_t len = std::min(STAGE_CHUNK, bytes - c * STAGE_CHUNK);
        if (used[b]) SSU_HIP(hipEventSynchronize(S.done[b]));   // its previous chunk has left for the device
        spread_copy((char*)S.buf[b], (const char*)host + c * STAGE_CHUNK, len, 0);
        SSU_HIP(hipMemcpyAsync((char*)dev + c * STAGE_CHUNK, S.buf[b], len, hipMemcpyHostToDevice, st));
        SSU_HIP(hipEventRecord(S.done[b], st));
        used[b] = true;
    }
    for (int b = 0; b < STAGE_BUFS; ++b)
        if (used[b]) SSU_HIP(hipEventSynchronize(S.done[b]));
    return true;
}

bool device_available(std::string* err) {
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n == 0) {
        if (err) *err = std::string("no HIP device available (") + hipGetErrorString(e) + ")";
        return false;
    }
    return true;
}

struct PlanImpl {
    ComputeSpec spec;
    uint32_t n = 0, m = 0, n_obs = 0;
    PostorderNodes nodes;            // the plan's tree, kept for refinement sub-tables
    // near-duplicate refinement (FP32 weighted / generalized): pairs below the
    // bound are recomputed in FP64; refine_S: the bound scales with S_i + S_j
    bool refine_on = false, refine_S = false;
    // refine_vaw (unnormalised VAW weighted): the pair kernel marks the pairs, against
    // each pair's VAW denominator, and every pair of the marked samples is recomputed
    bool refine_vaw = false;
    float refine_tau = 0.0f;
    uint64_t refined = 0;
    std::vector<uint32_t> refined_samples;   // the samples of the last refinement (its pairs lie among them)
    size_t ldn = 0;
    size_t blk = 0;                  // sample columns per embedding block (ldn: one block)
    bool gram = false;
    bool sparse = false;             // weighted (or generalized) through the shared-node kernels
    bool gen = false;                // generalized on the shared-node path
    // VAW on the shared-node path: nodes held by at least split_thr samples of the
    // plan's table (vdense_mask) stay with the dense pair kernel (its rows vrows,
    // lengths vlen_*), the others go to the posting lists; the one-sided sums are
    // taken at the table's distinct totals tau_h (sample s at kidx_h[s])
    bool vsplit = false;
    uint32_t n_vdense = 0;
    std::vector<uint32_t> vdense_mask, kidx_h;
    std::vector<int32_t> vrows;
    std::vector<float> vlen_f;
    std::vector<double> vlen_d, tau_h;
    uint32_t split_thr = 0;          // FP32 shared-node path: nodes held by fewer samples use posting lists (0: off)
    double split_theta = 0.0;        // the occupancy split chosen (threshold / n; 0: none)
    uint32_t split_slots = 0;        // bound on their postings in any execution
    // bit q: processed node q is held by fewer than split_thr samples of the plan's
    // table, so it is never dense in any execution (its FP32 operand row is not needed)
    std::vector<uint32_t> urare;
    // Gram path occupancy split: sparse nodes (held by few samples of the plan's
    // table) are summed from posting lists, the Gram keeps the others
    uint32_t gsplit_slots = 0;       // posting slots of all sparse nodes (0: off); d_poff holds their offsets
    std::vector<int32_t> sp_node;    // the sparse nodes
    std::vector<double> sp_len;      // their lengths as the accumulation sees them
    std::vector<uint8_t> gram_keep;  // per processed node: a Gram column (empty: every node)
    uint64_t nnz = 0;
    // Gram path on presence bits (every value finite and >= 0, n <= 262,144):
    // no FP64 embedding; ldb words of 32 samples per node row
    bool bits = false;
    uint32_t ldb = 0;
    std::vector<uint32_t> block_order;  // pair blocks / Gram tiles in launch order
    // shared-node pair blocks of 32 x 64 samples (FP32 operands, and generalized on
    // either), else 64 x 64 (FP64 weighted)
    bool rect() const { return sparse && (!spec.fp64 || gen); }
    // stripe selection (set_stripes): the pairs computed, and the blocks / tiles
    // holding one (run_order, on the device in d_order_run; empty: block_order)
    GapBand band;
    bool band_all = true;
    std::vector<uint32_t> run_order;
    const std::vector<uint32_t>& order_used() const { return band_all ? block_order : run_order; }
    // Output bands (execute_to_host): the tiled kernel runs as OUT_BANDS launches
    // (the Gram path OUT_BANDS_GRAM) over consecutive row-panel ranges (the blocks
    // of block_order in each, in order; the first two and the last two bands
    // narrower, so that copies start early and little is left to copy after the
    // last band; SSU_OUT_BANDS=<k> sets the count, 0 turns bands off).  Band k
    // writes every pair of its rows with the columns from its first row on, and
    // the mirrors, so after it its rows from that column on and those columns of
    // the rows below are final (an L whose size follows the band's work), and
    // their copy to the host overlaps the bands below.  The bands alternate
    // between st and band_st2, so that one band's last blocks share the device
    // with the next band's first.  Unweighted: band k is the Gram tiles of tile
    // rows in the band (band 0 also every diagonal tile), the posting walk of its
    // rows (after band 0's tiles: the diagonal G(a, a) of every row) and the
    // banded combine of its rows, in order on st.
    static constexpr uint32_t OUT_BANDS = 10, OUT_BANDS_GRAM = 16;
    std::vector<uint32_t> band_row0, band_off;   // rows [band_row0[k], band_row0[k + 1]); blocks [band_off[k], ..)
    std::vector<hipEvent_t> band_ev;
    hipEvent_t band_go = nullptr;                 // the bands' inputs are ready (band_st2 waits for it)
    hipStream_t copy_st = nullptr, band_st2 = nullptr;
    bool banded_run = false;
    std::function<bool(std::string*)> after_enqueue;   // host work between enqueue and the final wait
    bool prepare_bands(std::string* err);
    bool bad_lengths = false;

    // host schedule
    std::vector<int32_t> leaf_k, leaf_row;
    std::vector<int32_t> level_nodes;
    std::vector<uint32_t> level_off;
    std::vector<int32_t> child_ptr, child_idx;
    // heavy-path schedule (internal nodes by rounds of heavy paths; used instead of
    // the height levels when it takes fewer launches)
    bool use_paths = false;
    std::vector<int32_t> path_nodes, path_light, heavy;
    std::vector<uint32_t> path_off, round_off;
    std::vector<int32_t> knodes;     // nodes entering the pair accumulation
    std::vector<double> klen_d;
    std::vector<float> klen_f;
    std::vector<double> len_all;     // every non-root node, postorder (Faith PD)
    // exact Gram
    std::vector<int32_t> col_node;
    std::vector<int8_t> digits;      // [4][Kp]
    std::vector<GramChunk> chunks;
    size_t Kp = 0, n_pad = 0;

    hipStream_t st = nullptr;
    hipEvent_t ev[PH_COUNT + 2] = {};  // ev[EV_SPLIT]: posting lists start, ev[EV_ROWS]: Gram path rows end
    float times[PH_COUNT] = {};
    std::vector<double> totals;

    DevMem d_indptr, d_indices, d_data, d_totals, d_leaf_k, d_leaf_row, d_level_nodes, d_child_ptr, d_child_idx,
        d_path_nodes, d_path_light, d_path_off, d_heavy;
    HeavyPaths hpaths() const {
        return {d_path_nodes.as<int32_t>(), d_path_light.as<int32_t>(), d_path_off.as<uint32_t>(),
                d_heavy.as<int32_t>(), d_child_ptr.as<int32_t>(), d_child_idx.as<int32_t>()};
    }
    void build_paths(const PostorderNodes& nodes);
    DevMem d_knodes, d_klen, d_klen_sum, d_len_all, d_E, d_C, d_U, d_Cu, d_S, d_Spart, d_masks, d_visits, d_order, d_Pt, d_digits, d_chunks, d_G, d_col, d_out,
        d_faith, d_stripes, d_smask, d_dmask, d_pc, d_pidx, d_pval, d_walked, d_rows, d_rcount, d_roff, d_rbsum,
        d_sp_node, d_sp_len, d_umasks, d_B, d_poff, d_pcnt, d_obsum, d_Ssp, d_out2, d_pspmax, d_urare, d_close,
        d_ridx, d_order_run, d_order_band, d_vdense, d_vrows, d_vlen, d_tau, d_kidx, d_H, d_pcv, d_nbound;

    hipStream_t own_st = nullptr;    // created with the plan; st may point at a caller's stream
    // per-execution inputs and output (nullptr: the plan's own table data / totals / result)
    const double* x_data = nullptr;
    const double* x_totals = nullptr;
    void* x_out = nullptr;
    bool quiet = false;              // no events, no host synchronisation (stream capture)

    double* x_S = nullptr;           // per-sample sums S (nullptr: the plan's own)
    const double* data_ptr() const { return x_data ? x_data : d_data.as<double>(); }
    double* S_ptr() const { return x_S ? x_S : d_S.as<double>(); }
    const double* totals_ptr() const { return x_totals ? x_totals : d_totals.as<double>(); }
    void* out_ptr() const { return x_out ? x_out : d_out.p; }
    hipError_t mark(int i) { return quiet ? hipSuccess : hipEventRecord(ev[i], st); }

    ~PlanImpl() {
        for (auto& e : ev)
            if (e) (void)hipEventDestroy(e);
        for (auto& e : band_ev)
            if (e) (void)hipEventDestroy(e);
        if (band_go) (void)hipEventDestroy(band_go);
        if (copy_st) (void)hipStreamDestroy(copy_st);
        if (band_st2) (void)hipStreamDestroy(band_st2);
        if (own_st) (void)hipStreamDestroy(own_st);
    }

    bool build_gram_columns();
    bool split_gram(std::string* err);
    bool split_weighted(std::string* err);
    bool split_vaw(std::string* err);
    bool structural_counts(std::vector<uint32_t>& cnt, std::string* err);
    // presence bits of every node into dst, lb words per row (structural: every
    // stored entry counts as present)
    bool embed_bits(DevMem& dst, uint32_t lb, bool structural, std::string* err);
    // node-major FP64 embedding of sample columns [c0, c0 + w) into dst (row stride w)
    bool embed(DevMem& dst, bool normalize, size_t c0, size_t w, std::string* err);
    bool refine(const uint32_t* flags, std::string* err);
};

// Exact Gram column order.  Every non-zero length is an integer times a power
// of two, |l| = M * 2^lsb with M odd.  Lengths are grouped into buckets that
// share one scale 2^lsb_min; inside a bucket each length is the integer
// M' = |l| / 2^lsb_min < 2^bits, cut into 7-bit digits (int8 MFMA operands,
// four per pass).  Buckets greedily absorb smaller lengths while the digit
// pass count does not grow, so FP32 lengths (24-bit mantissas) spanning many
// binades need only a handful of buckets.
bool PlanImpl::build_gram_columns() {
    struct Item { int32_t k; int e; int lsb; uint64_t mant; bool neg; };
    std::vector<Item> items;
    for (size_t q = 0; q < knodes.size(); ++q) {
        const double l = spec.fp64 ? klen_d[q] : (double)klen_f[q];
        if (l == 0.0 || (!gram_keep.empty() && !gram_keep[q])) continue;
        if (!std::isfinite(l)) return false;
        int e = 0;
        const double f = std::frexp(std::fabs(l), &e);        // |l| = f * 2^e, f in [0.5, 1)
        const uint64_t mant = (uint64_t)std::ldexp(f, 53);     // exact, in [2^52, 2^53)
        items.push_back({knodes[q], e, e - 53 + __builtin_ctzll(mant), mant, l < 0});
    }
    std::stable_sort(items.begin(), items.end(), [](const Item& x, const Item& y) { return x.e > y.e; });
    col_node.clear();
    chunks.clear();
    std::vector<std::array<int8_t, 4>> dg;
    constexpr size_t kMaxChunk = (size_t(1) << 17) - GRAM_BK;  // keeps every int32 digit sum below 127 * 2^17
    size_t b0 = 0;
    while (b0 < items.size()) {
        const int top = items[b0].e;  // every member is < 2^top
        int lsb = items[b0].lsb;
        size_t b1 = b0 + 1;
        const int groups0 = (top - lsb + 27) / 28;
        while (b1 < items.size()) {
            const int nl = std::min(lsb, items[b1].lsb);
            const int g = (top - nl + 27) / 28;
            if (g > std::max(groups0, 1) || top - nl > 56) break;
            lsb = nl;
            ++b1;
        }
        const int bits = top - lsb;
        const int n_groups = std::max(1, (bits + 27) / 28);
        for (int g = 0; g < n_groups; ++g) {
            const double scale = std::ldexp(1.0, lsb + 28 * g);
            for (size_t c0 = b0; c0 < b1; c0 += kMaxChunk) {
                const size_t c1 = std::min(b1, c0 + kMaxChunk);
                GramChunk c;
                c.k0 = (uint32_t)col_node.size();
                c.scale = scale;
                for (size_t q = c0; q < c1; ++q) {
                    const Item& it = items[q];
                    const int sh = it.e - 53 - lsb;  // M' = mant * 2^sh, exact
                    const uint64_t v = sh >= 0 ? (it.mant << sh) : (it.mant >> (-sh));
                    std::array<int8_t, 4> d{};
                    for (int r = 0; r < 4; ++r) {
                        const int bit = 7 * (4 * g + r);
                        const int dig = bit < 64 ? (int)((v >> bit) & 127u) : 0;
                        d[r] = (int8_t)(it.neg ? -dig : dig);
                    }
                    col_node.push_back(it.k);
                    dg.push_back(d);
                }
                while (col_node.size() % GRAM_BK) { col_node.push_back(-1); dg.push_back({0, 0, 0, 0}); }
                c.k1 = (uint32_t)col_node.size();
                chunks.push_back(c);
            }
        }
        b0 = b1;
    }
    Kp = col_node.size();
    digits.assign(4 * std::max<size_t>(Kp, 1), 0);
    for (size_t c = 0; c < Kp; ++c)
        for (int r = 0; r < 4; ++r) digits[(size_t)r * Kp + c] = dg[c][r];
    return true;
}

// A node's structural occupancy: the number of samples with a stored table
// entry below it (presence bits of an all-present copy of the table, or the
// FP64 embedding of an all-ones copy, counted per node).  Any values on the
// same structure (the plan's own, or rarefied replicates drawn from it) make
// the node present in a subset of those samples, so bounds taken from it hold
// for every execution.
bool PlanImpl::structural_counts(std::vector<uint32_t>& cnt, std::string* err) {
    const uint32_t K = (uint32_t)knodes.size();
    cnt.assign(K, 0);
    if (!K) return true;
    DevMem ones, cnt_d, bits_tmp;
    if (!cnt_d.ensure(K * sizeof(uint32_t), err)) return false;
    const uint32_t lb = (uint32_t)(ldn / 32);
    if (lb <= BITS_MAX_WORDS) {
        DevMem& B = bits ? d_B : bits_tmp;   // the Gram path keeps its bit rows for the executions
        if (!embed_bits(B, lb, true, err)) return false;
        SSU_HIP(launch_count_bits(cnt_d.as<uint32_t>(), B.as<uint32_t>(), lb, d_knodes.as<int32_t>(), K, st));
    } else {
        if (!ones.ensure(std::max<uint64_t>(nnz, 1) * sizeof(double), err)) return false;
        SSU_HIP(launch_fill(ones.as<double>(), nnz, 1.0, st));
        SSU_HIP(hipMemsetAsync(cnt_d.p, 0, K * sizeof(uint32_t), st));
        x_data = ones.as<double>();
        for (size_t c0 = 0; c0 < ldn && c0 < n; c0 += blk) {
            const size_t w = std::min(blk, ldn - c0);
            if (!embed(d_E, false, c0, w, err)) { x_data = nullptr; return false; }
            SSU_HIP(launch_count_present(cnt_d.as<uint32_t>(), d_E.as<double>(), w, (uint32_t)c0,
                                         d_knodes.as<int32_t>(), K, n, st));
        }
        x_data = nullptr;
    }
    SSU_HIP(hipMemcpyAsync(cnt.data(), cnt_d.p, K * sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    SSU_HIP(hipStreamSynchronize(st));
    return true;
}

// Gram path occupancy split.  Nodes held by fewer than theta * n samples (by
// structural occupancy, and of positive length) leave the Gram for the posting
// lists, each with exactly its structural occupancy as posting slots; nodes
// held by no sample leave it altogether.  SSU_GRAM_SPLIT sets theta (0: off),
// SSU_GRAM_SPLIT_MIN_N the smallest table split (GSPLIT_MIN_N).
bool PlanImpl::split_gram(std::string* err) {
    const char* env_theta = getenv("SSU_GRAM_SPLIT");
    const double theta = env_theta ? atof(env_theta) : 1.0;
    uint32_t min_n = GSPLIT_MIN_N;
    if (const char* env = getenv("SSU_GRAM_SPLIT_MIN_N")) min_n = (uint32_t)atol(env);
    const uint32_t K = (uint32_t)knodes.size();
    if (!(theta > 0.0) || K == 0 || n < std::max<uint32_t>(min_n, 2)) return true;
    std::vector<uint32_t> cnt;
    if (!structural_counts(cnt, err)) return false;
    // the threshold: SSU_GRAM_SPLIT * n, or the cost model's choice on these occupancies
    uint32_t thr0 = env_theta ? (uint32_t)std::max(1.0, std::ceil(theta * n))
                              : choose_split(SplitKind::gram, n, cnt, 1ull << 30).thr;
    if (thr0 == 0) return true;

    const size_t npan = (n + 31) / 32, panels = ldn / 32;
    std::vector<uint32_t> off;
    auto classify = [&](uint32_t thr) {
        gram_keep.assign(K, 1);
        sp_node.clear();
        sp_len.clear();
        off.assign(1, 0u);
        uint64_t total = 0;
        for (uint32_t q = 0; q < K; ++q) {
            const double l = spec.fp64 ? klen_d[q] : (double)klen_f[q];
            if (cnt[q] == 0) { gram_keep[q] = 0; continue; }
            if (cnt[q] < thr && l > 0.0 && std::isfinite(l)) {
                gram_keep[q] = 0;
                sp_node.push_back(knodes[q]);
                sp_len.push_back(l);
                total += cnt[q];
                off.push_back((uint32_t)std::min<uint64_t>(total, 0xffffffffu));
            }
        }
        return total;
    };
    // postings (4-byte slots) and row lists (8-byte slots), prefix counts, sample
    // bits and the row lists' per-chunk counts and offsets; fewer than 2^30 slots
    // (32-bit slot offsets in the kernels)
    auto fits = [&](uint64_t slots) {
        const double ks = (double)sp_node.size();
        const double rl = (double)n * row_list_chunks(n, (uint32_t)sp_node.size()) * 8.0;   // row-list counts, offsets
        return slots * 12.0 + ks * (double)(npan + 1 + panels) * 4.0 + rl <= GSPLIT_BUDGET && slots < (1u << 30);
    };
    uint32_t thr = thr0;
    uint64_t slots = classify(thr);
    while (!fits(slots) && thr > 1) slots = classify(thr /= 2);
    split_theta = (double)thr / n;
    if (!fits(slots)) slots = classify(0);   // no sparse nodes
    if (!build_gram_columns()) return false;
    gsplit_slots = sp_node.empty() ? 0 : (uint32_t)slots;
    if (!sp_node.empty()) {
        const std::vector<uint32_t> zero((sp_node.size() + 31) / 32, 0u);
        if (!upload(d_sp_node, sp_node, st, err) || !upload(d_sp_len, sp_len, st, err) ||
            !upload(d_dmask, zero, st, err) || !upload(d_poff, off, st, err))
            return false;
        SSU_HIP(hipStreamSynchronize(st));   // `zero` and `off` are read before they go out of scope
    }
    return true;
}

// Shared-node weighted path: nodes held by fewer than theta * n samples are summed
// from posting lists instead of the 32 x 64 tiles (decided per execution from
// the actual occupancy).  Their postings at any execution number at most the
// sum over nodes of min(structural occupancy, thr - 1); that bound sizes the
// posting and row-list arrays and must stay below 2^30 (32-bit slot offsets in
// the kernels): theta is lowered to fit.  SSU_SPARSE_SPLIT sets theta (0: off).
bool PlanImpl::split_weighted(std::string* err) {
    const char* env_theta = getenv("SSU_SPARSE_SPLIT");
    const double theta = env_theta ? atof(env_theta) : 1.0;
    bool nonneg = true;
    for (size_t q = 0; q < knodes.size(); ++q) nonneg = nonneg && (spec.fp64 ? klen_d[q] >= 0.0 : klen_f[q] >= 0.0f);
    const uint32_t K = (uint32_t)knodes.size();
    if (!nonneg || !(theta > 0.0) || K == 0 || n < 2) return true;
    std::vector<uint32_t> cnt;
    if (!structural_counts(cnt, err)) return false;
    // the threshold: SSU_SPARSE_SPLIT * n, or the cost model's choice on these
    // occupancies (generalized terms cost several times the weighted ones on the
    // tiles, which moves the balance)
    const SplitKind kind = gen ? (spec.fp64 ? SplitKind::generalized64 : SplitKind::generalized32)
                               : spec.fp64 ? SplitKind::weighted64 : SplitKind::weighted32;
    uint32_t thr = env_theta ? (uint32_t)std::max(1.0, std::ceil(theta * n)) : choose_split(kind, n, cnt, 1ull << 30).thr;
    if (thr == 0) return true;
    auto bound = [&](uint32_t t) {
        uint64_t b = 0;
        for (uint32_t c : cnt) b += std::min<uint64_t>(c, t - 1);
        return b;
    };
    uint64_t slots = bound(thr);
    while (slots >= (1u << 30) && thr > 1) slots = bound(thr /= 2);
    if (slots >= (1u << 30)) return true;
    split_thr = thr;
    split_slots = (uint32_t)std::max<uint64_t>(slots, 1);
    urare.assign((K + 31) / 32, 0u);
    for (uint32_t q = 0; q < K; ++q)
        if (cnt[q] < thr) urare[q >> 5] |= 1u << (q & 31);
    return upload(d_urare, urare, st, err);   // (a member: the host copy outlives the upload)
}

// VAW on the shared-node path: nodes held by fewer than theta * n samples of the
// plan's table (structural occupancy: no execution holds more) are summed from
// posting lists, the others by the dense pair kernel, decided once here.  The
// one-sided sums are taken at the table's distinct totals (an execution with
// totals of its own takes them at every sample's).  The [n][D] sums must fit
// VAW_H_BUDGET; SSU_VAW_SPLIT sets theta (0: off).
bool PlanImpl::split_vaw(std::string* err) {
    constexpr size_t VAW_H_BUDGET = (size_t)16 << 30;
    const char* env_theta = getenv("SSU_VAW_SPLIT");
    const double theta = env_theta ? atof(env_theta) : 1.0;
    const uint32_t K = (uint32_t)knodes.size();
    vsplit = false;
    if (!(theta > 0.0) || K == 0 || n < 2 || (size_t)n * n * sizeof(double) > VAW_H_BUDGET) return true;
    // the distinct totals, increasing, and each sample's index among them
    tau_h.assign(totals.begin(), totals.begin() + n);
    std::sort(tau_h.begin(), tau_h.end());
    tau_h.erase(std::unique(tau_h.begin(), tau_h.end()), tau_h.end());
    kidx_h.resize(n);
    for (uint32_t s = 0; s < n; ++s)
        kidx_h[s] = (uint32_t)(std::lower_bound(tau_h.begin(), tau_h.end(), totals[s]) - tau_h.begin());
    std::vector<uint32_t> cnt;
    if (!structural_counts(cnt, err)) return false;
    const SplitKind kind = spec.fp64 ? SplitKind::vaw64 : SplitKind::vaw32;
    uint32_t thr = env_theta ? (uint32_t)std::max(1.0, std::ceil(theta * n))
                             : choose_split(kind, n, cnt, 1ull << 30, (uint32_t)tau_h.size()).thr;
    if (thr == 0) return true;
    auto bound = [&](uint32_t t) {
        uint64_t b = 0;
        for (uint32_t c : cnt) b += c < t ? c : 0;
        return b;
    };
    uint64_t slots = bound(thr);
    while (slots >= (1u << 30) && thr > 1) slots = bound(thr /= 2);
    if (slots >= (1u << 30)) return true;
    split_thr = thr;
    split_slots = (uint32_t)std::max<uint64_t>(slots, 1);
    vdense_mask.assign((K + 31) / 32, 0u);
    vrows.clear();
    vlen_f.clear();
    vlen_d.clear();
    for (uint32_t q = 0; q < K; ++q) {
        if (cnt[q] < thr) continue;
        vdense_mask[q >> 5] |= 1u << (q & 31);
        vrows.push_back(spec.fp64 ? knodes[q] : (int32_t)q);   // FP64: embedding rows; FP32: operand rows
        if (spec.fp64) vlen_d.push_back(klen_d[q]);
        else vlen_f.push_back(klen_f[q]);
    }
    n_vdense = (uint32_t)vrows.size();
    if (vrows.empty()) {   // (uploads of one padding row; the kernel takes n_vdense = 0 rows)
        vrows.push_back(0);
        if (spec.fp64) vlen_d.push_back(0.0);
        else vlen_f.push_back(0.0f);
    }
    vsplit = true;
    // (members: the host copies outlive the uploads)
    return upload(d_vdense, vdense_mask, st, err) && upload(d_vrows, vrows, st, err) &&
           (spec.fp64 ? upload(d_vlen, vlen_d, st, err) : upload(d_vlen, vlen_f, st, err)) &&
           upload(d_tau, tau_h, st, err) && upload(d_kidx, kidx_h, st, err);
}

bool PlanImpl::embed_bits(DevMem& dst, uint32_t lb, bool structural, std::string* err) {
    if (!dst.ensure(std::max<size_t>(m, 1) * lb * sizeof(uint32_t), err)) return false;
    uint32_t* B = dst.as<uint32_t>();
    SSU_HIP(launch_bits_leaves(B, lb, d_leaf_k.as<int32_t>(), d_leaf_row.as<int32_t>(), (uint32_t)leaf_k.size(),
                               d_indptr.as<uint32_t>(), d_indices.as<uint32_t>(), data_ptr(), structural, st));
    if (use_paths) {
        for (size_t r = 0; r + 1 < round_off.size(); ++r)
            SSU_HIP(launch_bits_paths(B, lb, hpaths(), round_off[r], round_off[r + 1], st));
        return true;
    }
    for (size_t h = 0; h + 1 < level_off.size(); ++h) {
        const uint32_t a = level_off[h], b = level_off[h + 1];
        SSU_HIP(launch_bits_level(B, lb, d_level_nodes.as<int32_t>() + a, b - a, d_child_ptr.as<int32_t>(),
                                  d_child_idx.as<int32_t>(), st));
    }
    return true;
}

bool PlanImpl::embed(DevMem& dst, bool normalize, size_t c0, size_t w, std::string* err) {
    if (!dst.ensure(std::max<size_t>(m, 1) * blk * sizeof(double), err)) return false;
    double* E = dst.as<double>();
    SSU_HIP(launch_embed_leaves(E, w, (uint32_t)c0, d_leaf_k.as<int32_t>(), d_leaf_row.as<int32_t>(),
                                (uint32_t)leaf_k.size(), d_indptr.as<uint32_t>(), d_indices.as<uint32_t>(),
                                data_ptr(), totals_ptr(), normalize, st));
    if (use_paths) {
        for (size_t r = 0; r + 1 < round_off.size(); ++r)
            SSU_HIP(launch_embed_paths(E, w, hpaths(), round_off[r], round_off[r + 1], st));
        return true;
    }
    for (size_t h = 0; h + 1 < level_off.size(); ++h) {
        const uint32_t a = level_off[h], b = level_off[h + 1];
        SSU_HIP(launch_embed_level(E, w, d_level_nodes.as<int32_t>() + a, b - a, d_child_ptr.as<int32_t>(),
                                   d_child_idx.as<int32_t>(), st));
    }
    return true;
}

// Heavy-path decomposition: every internal node continues the path of its
// largest child (the first of equal ones); a path runs from a leaf up to a node
// that is not its parent's heavy child (or a child of the root).  A path's
// round is one more than the latest round of the paths its other children head
// (0 when they are all leaves), so at most log2(leaves) + 1 rounds, against one
// launch per height level (a caterpillar tree of T tips has T - 1 levels).
void PlanImpl::build_paths(const PostorderNodes& nodes) {
    std::vector<int32_t> size(m, 1);
    heavy.assign(std::max<uint32_t>(m, 1), -1);
    for (uint32_t k = 0; k < m; ++k) {   // postorder: a node's subtree precedes it
        for (int32_t c = child_ptr[k]; c < child_ptr[k + 1]; ++c) {
            const int32_t ch = child_idx[c];
            size[k] += size[ch];
            if (heavy[k] < 0 || size[ch] > size[heavy[k]]) heavy[k] = ch;
        }
    }
    // round of the path headed at each node (-1: a leaf); paths in head order
    std::vector<int32_t> round(m, -1);
    std::vector<std::vector<int32_t>> by_round;
    std::vector<std::vector<uint32_t>> by_round_len;
    std::vector<int32_t> path;
    for (uint32_t h = 0; h < m; ++h) {
        const int32_t p = nodes.parent[h];
        const bool head = (uint32_t)p >= m || heavy[p] != (int32_t)h;
        if (!head || nodes.is_leaf[h] || child_ptr[h] == child_ptr[h + 1]) continue;
        path.clear();
        for (int32_t k = (int32_t)h; child_ptr[k] < child_ptr[k + 1]; k = heavy[k]) path.push_back(k);
        std::reverse(path.begin(), path.end());   // bottom to top
        int32_t r = 0;
        for (int32_t k : path)
            for (int32_t c = child_ptr[k]; c < child_ptr[k + 1]; ++c)
                if (child_idx[c] != heavy[k] && round[child_idx[c]] >= 0) r = std::max(r, round[child_idx[c]] + 1);
        round[h] = r;
        if ((size_t)r >= by_round.size()) { by_round.resize(r + 1); by_round_len.resize(r + 1); }
        by_round[r].insert(by_round[r].end(), path.begin(), path.end());
        by_round_len[r].push_back((uint32_t)path.size());
    }
    path_nodes.clear();
    path_light.clear();
    path_off.assign(1, 0u);
    round_off.assign(1, 0u);
    for (size_t r = 0; r < by_round.size(); ++r) {
        for (int32_t k : by_round[r]) {
            path_nodes.push_back(k);
            const bool two = child_ptr[k + 1] - child_ptr[k] == 2;
            const int32_t other = child_idx[child_ptr[k]] == heavy[k] ? child_idx[child_ptr[k] + 1] : child_idx[child_ptr[k]];
            path_light.push_back(two ? other : -1);
        }
        for (uint32_t len : by_round_len[r]) path_off.push_back(path_off.back() + len);
        round_off.push_back((uint32_t)path_off.size() - 1);
    }
}

struct TableUpload::Impl {
    std::thread th;
    DevMem ip, ix, d;
    bool ok = false;
    std::string err;
    const uint32_t* indptr = nullptr;
    uint32_t n_obs = 0;
};

std::shared_ptr<TableUpload> TableUpload::start(const uint32_t* indptr, const uint32_t* indices, const double* data,
                                                uint32_t n_obs) {
    auto u = std::make_shared<TableUpload>();
    u->impl.reset(new Impl());
    Impl* I = u->impl.get();
    I->indptr = indptr;
    I->n_obs = n_obs;
    I->th = std::thread([I, indptr, indices, data, n_obs] {
        const uint32_t base = n_obs ? indptr[0] : 0;
        if (n_obs && indptr[n_obs] < base) { I->err = "table indptr is not monotone"; return; }
        const uint64_t nz = n_obs ? (uint64_t)indptr[n_obs] - base : 0;
        hipStream_t st = nullptr;
        if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) { I->err = "hipStreamCreate failed"; return; }
        std::vector<uint32_t> ip(indptr, indptr + n_obs + 1);
        for (auto& x : ip) x -= base;
        I->ok = upload(I->ip, ip, st, &I->err) &&
                upload_raw(I->ix, indices + base, nz * sizeof(uint32_t), st, &I->err) &&
                upload_raw(I->d, data + base, nz * sizeof(double), st, &I->err) &&
                hipStreamSynchronize(st) == hipSuccess;   // `ip` is read until here
        (void)hipStreamDestroy(st);
    });
    return u;
}

TableUpload::~TableUpload() {
    if (impl && impl->th.joinable()) impl->th.join();
}

Plan::Plan() : p_(new PlanImpl) {}
Plan::~Plan() = default;

ComputeSpec spec_for_table(const Table& t, ComputeSpec s) {
    if (!s.vaw || s.fp64) return s;
    double t1 = 0.0, t2 = 0.0;   // the two largest totals
    for (double x : t.totals) {
        if (x > t1) { t2 = t1; t1 = x; }
        else if (x > t2) t2 = x;
    }
    if (!t.values_integral || !(t1 + t2 < 16777216.0)) s.fp64 = true;
    return s;
}

std::unique_ptr<Plan> Plan::create(const Table& table, const PostorderNodes& nodes, const ComputeSpec& spec,
                                   std::string* err, const std::shared_ptr<TableUpload>& pre) {
    if (!device_available(err)) return nullptr;
    // SSU_TIMING=1: wall-clock steps of the plan's creation on stderr
    const bool timing = getenv("SSU_TIMING") && atoi(getenv("SSU_TIMING")) != 0;
    auto tprev = std::chrono::steady_clock::now();
    auto step = [&](const char* what) {
        if (!timing) return;
        const auto now = std::chrono::steady_clock::now();
        fprintf(stderr, "[ssu plan] %-28s %8.2f ms\n", what, std::chrono::duration<double, std::milli>(now - tprev).count());
        tprev = now;
    };
    // the device kernels index with these values: reject malformed CSR up front
    for (uint32_t i = 0; i < table.n_obs; ++i)
        if (table.indptr[i + 1] < table.indptr[i]) { if (err) *err = "table indptr is not monotone"; return nullptr; }
    const uint64_t nz_all = table.nnz();
    const uint32_t nz_base = table.n_obs ? table.indptr[0] : 0;
    if (!table.index_ok) {   // (checked by the pass that formed the table's totals)
        if (err) *err = "table sample index out of range";
        return nullptr;
    }
    for (uint32_t k = 0; k < nodes.m; ++k)
        if (nodes.obs_row[k] >= (int32_t)table.n_obs || nodes.parent[k] < 0 || (uint32_t)nodes.parent[k] > nodes.m) {
            if (err) *err = "inconsistent postorder arrays";
            return nullptr;
        }
    std::unique_ptr<Plan> plan(new Plan());
    PlanImpl& P = *plan->p_;
    P.spec = spec;
    P.n = table.n_samples;
    P.m = nodes.m;
    P.n_obs = table.n_obs;
    P.ldn = round_up(std::max<uint32_t>(P.n, 1), PAIR_TILE);
    std::string e2;
    if (!err) err = &e2;
    hipError_t he = hipStreamCreateWithFlags(&P.own_st, hipStreamNonBlocking);
    if (he != hipSuccess) { *err = std::string("hipStreamCreate: ") + hipGetErrorString(he); return nullptr; }
    P.st = P.own_st;
    // the CSR goes up on a second thread while this one builds the schedules
    // (or is already on its way, started by the caller on the same arrays)
    P.nnz = nz_all;
    std::string up_err;
    bool up_ok = true;
    const bool pre_ok = pre && pre->impl && pre->impl->indptr == table.indptr && pre->impl->n_obs == table.n_obs;
    std::thread uploader([&] {
        if (pre_ok) {
            TableUpload::Impl& U = *pre->impl;
            if (U.th.joinable()) U.th.join();
            up_ok = U.ok;
            up_err = U.err;
            if (up_ok) {
                P.d_indptr.take(U.ip);
                P.d_indices.take(U.ix);
                P.d_data.take(U.d);
            }
            return;
        }
        std::vector<uint32_t> ip(table.indptr, table.indptr + table.n_obs + 1);
        for (auto& x : ip) x -= nz_base;
        up_ok = upload(P.d_indptr, ip, P.st, &up_err) &&
                upload_raw(P.d_indices, table.indices + nz_base, nz_all * sizeof(uint32_t), P.st, &up_err) &&
                upload_raw(P.d_data, table.data + nz_base, nz_all * sizeof(double), P.st, &up_err) &&
                hipStreamSynchronize(P.st) == hipSuccess;   // `ip` is read until here
    });
    struct Join {
        std::thread& t;
        ~Join() { if (t.joinable()) t.join(); }
    } join_uploader{uploader};
    // schedule: leaves, height levels and the accumulated nodes on a second
    // thread; children lists (postorder, left to right) and heavy paths here
    std::thread levels([&P, &nodes, &spec] {
        P.nodes = nodes;
        int32_t max_h = 0;
        for (uint32_t k = 0; k < P.m; ++k) {
            if (nodes.is_leaf[k]) { P.leaf_k.push_back((int32_t)k); P.leaf_row.push_back(nodes.obs_row[k]); }
            max_h = std::max(max_h, nodes.height[k]);
        }
        // internal nodes by height: a counting sort, postorder within a level
        std::vector<uint32_t> cnt(max_h + 2, 0);
        for (uint32_t k = 0; k < P.m; ++k)
            if (!nodes.is_leaf[k] && nodes.height[k] >= 1) ++cnt[nodes.height[k] + 1];
        for (int32_t h = 1; h <= max_h; ++h) cnt[h + 1] += cnt[h];
        P.level_nodes.assign(cnt[max_h + 1], 0);
        P.level_off.assign(std::max(max_h, 0) + 1, 0u);
        for (int32_t h = 1; h <= max_h; ++h) P.level_off[h] = cnt[h + 1];
        for (uint32_t k = 0; k < P.m; ++k)
            if (!nodes.is_leaf[k] && nodes.height[k] >= 1) P.level_nodes[cnt[nodes.height[k]]++] = (int32_t)k;
        P.len_all.assign(nodes.length.begin(), nodes.length.begin() + P.m);
        P.knodes.reserve(P.m);
        P.klen_d.reserve(P.m);
        P.klen_f.reserve(P.m);
        for (uint32_t k = 0; k < P.m; ++k) {
            if (spec.bypass_tips && nodes.is_leaf[k]) continue;  // src/unifrac_cmp.cpp:123-124
            P.knodes.push_back((int32_t)k);
            P.klen_d.push_back(nodes.length[k]);
            P.klen_f.push_back((float)nodes.length[k]);          // TFloat lengths, src/unifrac_cmp.cpp:127
        }
    });
    struct JoinLevels {
        std::thread& t;
        ~JoinLevels() { if (t.joinable()) t.join(); }
    } join_levels{levels};
    std::vector<int32_t> n_child(P.m + 1, 0);
    for (uint32_t k = 0; k < P.m; ++k)
        if ((uint32_t)nodes.parent[k] < P.m) n_child[nodes.parent[k]]++;
    P.child_ptr.assign(P.m + 1, 0);
    for (uint32_t k = 0; k < P.m; ++k) P.child_ptr[k + 1] = P.child_ptr[k] + n_child[k];
    P.child_idx.assign(std::max<int32_t>(P.child_ptr[P.m], 1), 0);
    std::vector<int32_t> fill(P.child_ptr.begin(), P.child_ptr.end() - 1);
    for (uint32_t k = 0; k < P.m; ++k)
        if ((uint32_t)nodes.parent[k] < P.m) P.child_idx[fill[nodes.parent[k]]++] = (int32_t)k;

    step("validate, children");
    P.build_paths(nodes);
    levels.join();
    // heavy paths when they take fewer launches than the height levels (SSU_EMBED_PATHS=0 / 1 forces)
    P.use_paths = P.round_off.size() < P.level_off.size();
    if (const char* env = getenv("SSU_EMBED_PATHS")) P.use_paths = atoi(env) != 0;

    step("schedule");
    const bool unw = spec.method == Method::unweighted || spec.method == Method::unweighted_unnormalized;
    P.gram = unw && !spec.vaw && spec.plain_kernel == 0;
    if (P.gram)   // non-finite lengths: VALU path
        for (size_t q = 0; q < P.knodes.size() && P.gram; ++q)
            P.gram = std::isfinite(spec.fp64 ? P.klen_d[q] : (double)P.klen_f[q]);
    P.n_pad = round_up(std::max<uint32_t>(P.n, 1), GRAM_BM);
    // every stored value finite and >= 0 (for the presence bits and the shared-node
    // kernels), found by the pass that formed the table's totals
    const bool vals_nonneg = table.values_nonneg;
    step("value checks");
    // Presence bits for the Gram path: with every value finite and >= 0 (and
    // positive totals), E[k][s] > 0 exactly when a stored value > 0 lies below
    // k, an OR up the tree.  SSU_GRAM_BITS=0 keeps the FP64 embedding.
    if (P.gram && P.ldn / 32 <= BITS_MAX_WORDS) {
        bool ok_vals = vals_nonneg;
        if (const char* env = getenv("SSU_GRAM_BITS")) ok_vals = ok_vals && atoi(env) != 0;
        if (ok_vals) {
            P.bits = true;
            P.ldb = (uint32_t)(P.ldn / 32);
        }
    }
    // The shared-node weighted kernel needs u >= 0 (so min(u, v) vanishes when
    // either sample lacks the node) and finite lengths (S_i + S_j - 2M).
    const bool weighted = spec.method == Method::weighted_normalized || spec.method == Method::weighted_unnormalized;
    // generalized (FP32, 0 <= alpha <= 1) shares the shared-node path: numerator and
    // denominator summed directly on the tiles, posting-list shares for the rare nodes
    P.gen = spec.method == Method::generalized && !spec.vaw && spec.plain_kernel == 0 &&
            spec.alpha >= 0.0 && spec.alpha <= 1.0 && !(getenv("SSU_GEN_SPARSE") && atoi(getenv("SSU_GEN_SPARSE")) == 0);
    P.sparse = (weighted && !spec.vaw && spec.plain_kernel == 0) || P.gen;
    if (P.sparse) {
        for (size_t q = 0; q < P.klen_d.size() && P.sparse; ++q)
            if (!std::isfinite(spec.fp64 ? P.klen_d[q] : (double)P.klen_f[q])) P.sparse = false;
        P.sparse = P.sparse && vals_nonneg;
    }
    P.gen = P.gen && P.sparse;
    // VAW (weighted, presence, or generalized with 0 <= alpha <= 1 terms) takes the
    // posting lists for its rare nodes when every value and length is finite and
    // >= 0 (split_vaw decides whether it pays)
    P.vsplit = spec.vaw && spec.plain_kernel == 0 && vals_nonneg &&
               (spec.method != Method::generalized || (spec.alpha >= 0.0 && spec.alpha <= 1.0));
    for (size_t q = 0; q < P.klen_d.size() && P.vsplit; ++q) {
        const double l = spec.fp64 ? P.klen_d[q] : (double)P.klen_f[q];
        P.vsplit = std::isfinite(l) && l >= 0.0;
    }
    // Near-duplicate refinement: an FP32 weighted or generalized distance is off
    // by at most 2^-24 of its denominator through the FP32 rounding of the
    // proportions (|u - v| from rounded values is off by at most 2^-24 (u + v)),
    // a relative error of 2^-24 / d, plus the accumulation's own, well below
    // 2e-7.  Pairs with d < 0.1 (d < 0.1 (S_i + S_j) unnormalised) are
    // recomputed in FP64, which keeps every entry within 1e-6.  The
    // unnormalised weighted distance needs S, computed on the shared-node path
    // (tables with negative values or non-finite lengths run the dense kernel
    // without it, and are not refined).  The unnormalised VAW distance is
    // bounded by its own VAW denominator, which only the pair kernel holds: it
    // marks those pairs itself (refine_vaw).  SSU_REFINE=0 turns refinement
    // off, SSU_REFINE_BELOW sets the bound.
    if (!spec.fp64 && spec.plain_kernel == 0) {
        P.refine_vaw = spec.method == Method::weighted_unnormalized && spec.vaw;
        P.refine_on = spec.method == Method::weighted_normalized || spec.method == Method::generalized ||
                      (spec.method == Method::weighted_unnormalized && (P.sparse || spec.vaw));
        P.refine_S = spec.method == Method::weighted_unnormalized && !spec.vaw;
        P.refine_tau = 0.1f;
        if (const char* env = getenv("SSU_REFINE_BELOW")) P.refine_tau = (float)atof(env);
        if (const char* env = getenv("SSU_REFINE")) P.refine_on = P.refine_on && atoi(env) != 0;
    }
    // Embedding blocks: the FP64 embedding of every node for every sample
    // (m x ldn doubles) is built in sample-column blocks when it would exceed
    // EMBED_BUDGET; each block feeds the FP32 operand, masks, sums or presence
    // bytes of its columns.  FP64 pair kernels read the embedding itself, so
    // they keep a single block.  SSU_EMBED_BLOCK_SAMPLES overrides the width.
    constexpr size_t EMBED_BUDGET = (size_t)48 << 30;
    const bool need_full = spec.fp64 && !P.gram;
    const size_t row_bytes = std::max<size_t>(P.m, 1) * sizeof(double) * (spec.vaw ? 2 : 1);
    P.blk = P.ldn;
    if (!need_full) {
        if (const char* env = getenv("SSU_EMBED_BLOCK_SAMPLES")) {
            const long v = atol(env);
            if (v > 0) P.blk = std::min(P.ldn, round_up((size_t)v, (size_t)PAIR_TILE));
        } else if (row_bytes * P.ldn > EMBED_BUDGET) {
            P.blk = std::max<size_t>(PAIR_TILE, EMBED_BUDGET / row_bytes / PAIR_TILE * PAIR_TILE);
        }
    }

    step("method setup");
    for (auto& e : P.ev)
        if (hipEventCreate(&e) != hipSuccess) { *err = "hipEventCreate failed"; return nullptr; }

    P.totals = table.totals;
    uploader.join();
    if (!up_ok) { *err = up_err; return nullptr; }
    bool ok = upload(P.d_totals, P.totals, P.st, err) && upload(P.d_leaf_k, P.leaf_k, P.st, err) &&
              upload(P.d_leaf_row, P.leaf_row, P.st, err) && upload(P.d_level_nodes, P.level_nodes, P.st, err) &&
              upload(P.d_child_ptr, P.child_ptr, P.st, err) && upload(P.d_child_idx, P.child_idx, P.st, err) &&
              upload(P.d_path_nodes, P.path_nodes, P.st, err) && upload(P.d_path_light, P.path_light, P.st, err) &&
              upload(P.d_path_off, P.path_off, P.st, err) && upload(P.d_heavy, P.heavy, P.st, err) &&
              upload(P.d_knodes, P.knodes, P.st, err) && upload(P.d_len_all, P.len_all, P.st, err);
    if (ok) ok = spec.fp64 ? upload(P.d_klen, P.klen_d, P.st, err) : upload(P.d_klen, P.klen_f, P.st, err);
    // lengths as the accumulation sees them, in FP64, for the per-sample denominators
    // (declared here: the asynchronous upload reads it until the synchronisation below)
    std::vector<double> ls(P.klen_f.size());
    if (ok) {
        for (size_t q = 0; q < ls.size(); ++q) ls[q] = spec.fp64 ? P.klen_d[q] : (double)P.klen_f[q];
        ok = upload(P.d_klen_sum, ls, P.st, err);
    }
    // 16 x 16 super-tiles of 64-sample blocks, 6 x 6 of 128-sample Gram tiles:
    // about the number of blocks (tiles) resident on one XCD
    if (ok && P.sparse) {
        P.block_order = P.rect() ? rect_block_order(P.n) : upper_block_order((P.n + 63) / 64, 16);
        ok = upload(P.d_order, P.block_order, P.st, err);
    }
    if (ok && P.gram) {
        P.block_order = upper_block_order((uint32_t)(P.n_pad / GRAM_BM), 6);
        ok = upload(P.d_order, P.block_order, P.st, err);
    }
    step("stream, events, uploads");
    if (ok && P.gram) ok = P.split_gram(err);   // builds the Gram columns (narrowed when it splits)
    if (ok && P.gram && P.col_node.empty() && !P.build_gram_columns()) ok = false;
    if (ok && P.sparse) ok = P.split_weighted(err);
    if (ok && P.vsplit) ok = P.split_vaw(err);
    step("occupancy split");
    if (ok && P.gram && P.Kp)
        ok = upload(P.d_col, P.col_node, P.st, err) && upload(P.d_digits, P.digits, P.st, err) &&
             upload(P.d_chunks, P.chunks, P.st, err);
    if (!ok) return nullptr;
    if (hipStreamSynchronize(P.st) != hipSuccess) { *err = "upload failed"; return nullptr; }
    step("gram columns, sync");
    return plan;
}

bool Plan::execute(std::string* err) { return execute_into(nullptr, nullptr, nullptr, false, err); }

bool Plan::execute_into(const double* d_data, const double* d_totals, void* d_out, bool quiet, std::string* err,
                        uint32_t* d_flags, double* d_sums) {
    PlanImpl& P = *p_;
    P.x_data = d_data;
    P.x_totals = d_totals;
    P.x_out = d_out;
    P.x_S = d_sums;
    P.quiet = quiet;
    const ComputeSpec& s = P.spec;
    const size_t nn = (size_t)P.n * P.n;
    const size_t osz = s.fp64 ? sizeof(double) : sizeof(float);
    if (!d_out && !P.d_out.ensure(std::max<size_t>(nn, 1) * osz, err)) return false;
    const uint32_t K = (uint32_t)P.knodes.size();
    const bool single = P.blk >= P.ldn;

    // what the pair stage needs from each embedding block
    PairOp op = PairOp::weighted;
    bool normalized = true;
    switch (s.method) {
    case Method::unweighted: op = s.vaw ? PairOp::vaw_unweighted : PairOp::unweighted_plain; break;
    case Method::unweighted_unnormalized: op = s.vaw ? PairOp::vaw_unweighted : PairOp::unweighted_plain; normalized = false; break;
    case Method::weighted_normalized: op = s.vaw ? PairOp::vaw_weighted : PairOp::weighted; break;
    case Method::weighted_unnormalized: op = s.vaw ? PairOp::vaw_weighted : PairOp::weighted; normalized = false; break;
    case Method::generalized: op = s.vaw ? PairOp::vaw_generalized : PairOp::generalized; break;
    }
    const bool dense = !P.gram && !P.sparse;
    // per-sample sums S: the shared-node kernel always, the dense kernel for weighted_normalized
    const bool sums = P.sparse || (dense && op == PairOp::weighted && normalized);
    const bool gather = !P.gram && !s.fp64;            // FP32 operand U (and Cu for VAW)
    const uint32_t Kw = (K + 31) / 32;
    if (sums) {
        if (!P.d_S.ensure(std::max<uint32_t>(P.n, 1) * sizeof(double), err)) return false;
        const size_t w_last = P.ldn - (P.ldn - 1) / P.blk * P.blk;
        size_t parts = std::max<size_t>({(size_t)gather_segments(K, P.blk) * P.blk,
                                         (size_t)gather_segments(K, w_last) * w_last, 1});
        if (s.fixed_segments && K)
            parts = std::max(parts, (size_t)((K + fixed_rows_per_segment(K) - 1) / fixed_rows_per_segment(K)) * P.blk);
        if (!P.d_Spart.ensure(parts * sizeof(double), err)) return false;
    }
    const bool shared = P.sparse || P.vsplit;   // sample masks and (with a split) posting lists
    if (shared && !P.d_masks.ensure(std::max<size_t>((P.ldn / 32) * Kw, 1) * sizeof(uint32_t), err)) return false;
    const bool split = shared && P.split_thr > 0 && K > 0;
    const bool vsplit = P.vsplit && split;
    // VAW one-sided sums: at the plan's distinct totals, or at every sample's when the
    // execution brings totals of its own
    const uint32_t vD = P.x_totals ? P.n : (uint32_t)P.tau_h.size();
    const uint32_t npan = (P.n + 31) / 32;
    if (split) {
        const size_t kk = std::max<size_t>(K, 1);
        if (!P.d_smask.ensure((P.ldn / 32) * kk * sizeof(uint32_t), err) ||
            !P.d_dmask.ensure(std::max<size_t>(Kw, 1) * sizeof(uint32_t), err) ||
            !P.d_pc.ensure((size_t)(npan + 1) * kk * sizeof(uint32_t), err) ||
            !P.d_pidx.ensure((size_t)P.split_slots * sizeof(uint32_t), err) ||
            !P.d_pval.ensure((size_t)P.split_slots * (s.fp64 ? sizeof(double) : sizeof(float)), err) ||
            !P.d_poff.ensure((kk + 1) * sizeof(uint32_t), err) ||
            !P.d_pcnt.ensure(kk * sizeof(uint32_t), err) ||
            !P.d_Ssp.ensure(std::max<uint32_t>(P.n, 1) * sizeof(double), err) ||
            ((P.gen || vsplit) && !P.d_out2.ensure(std::max<size_t>(nn, 1) * osz, err)) ||
            (P.gen && !P.d_pspmax.ensure(sizeof(double), err)) ||
            (vsplit && !P.d_pcv.ensure((size_t)P.split_slots * (s.fp64 ? sizeof(double) : sizeof(float)), err)) ||
            (vsplit && !P.d_H.ensure(std::max<size_t>((size_t)P.n * vD, 1) * sizeof(double), err)) ||
            (vsplit && !P.d_nbound.ensure(kk * sizeof(double), err)) ||
            !P.d_obsum.ensure((offset_scan_blocks(K) + 1) * sizeof(uint32_t), err) ||
            !P.d_walked.ensure((size_t)sparse_rows_blocks(P.n) * sizeof(uint32_t), err) ||
            !P.d_rows.ensure((size_t)P.split_slots * sizeof(unsigned long long), err) ||
            !P.d_rcount.ensure((size_t)P.n * row_list_chunks(P.n, K) * sizeof(uint32_t), err) ||
            !P.d_roff.ensure(((size_t)P.n * row_list_chunks(P.n, K) + 1) * sizeof(uint32_t), err) ||
            !P.d_rbsum.ensure((row_list_scan_blocks(P.n, K) + 1) * sizeof(uint32_t), err))
            return false;
    }
    // Gram path split: sample bits of the sparse nodes from each embedding block,
    // posting and row lists before the Gram, each row's share added to G after it
    const uint32_t Ks = (uint32_t)P.sp_node.size();
    const bool gsplit = P.gram && P.gsplit_slots > 0 && Ks > 0;
    if (gsplit) {
        const size_t ks = Ks, kws = (ks + 31) / 32, panels = P.ldn / 32;
        if (!P.d_umasks.ensure(panels * kws * sizeof(uint32_t), err) ||
            !P.d_smask.ensure(panels * ks * sizeof(uint32_t), err) ||
            !P.d_pc.ensure((size_t)(npan + 1) * ks * sizeof(uint32_t), err) ||
            !P.d_pidx.ensure((size_t)P.gsplit_slots * sizeof(uint32_t), err) ||
            !P.d_walked.ensure((size_t)sparse_rows_blocks(P.n) * sizeof(uint32_t), err) ||
            !P.d_rows.ensure((size_t)P.gsplit_slots * sizeof(unsigned long long), err) ||
            !P.d_rcount.ensure((size_t)P.n * row_list_chunks(P.n, Ks) * sizeof(uint32_t), err) ||
            !P.d_roff.ensure(((size_t)P.n * row_list_chunks(P.n, Ks) + 1) * sizeof(uint32_t), err) ||
            !P.d_rbsum.ensure((row_list_scan_blocks(P.n, Ks) + 1) * sizeof(uint32_t), err))
            return false;
    }
    if (P.gram && !P.d_G.ensure(P.n_pad * P.n_pad * sizeof(double), err)) return false;
    if (gather) {
        if (!P.d_U.ensure(std::max<size_t>(K, 1) * P.ldn * sizeof(float), err)) return false;
        if (s.vaw && !P.d_Cu.ensure(std::max<size_t>(K, 1) * P.ldn * sizeof(float), err)) return false;
    }
    if (P.gram && P.Kp && !P.d_Pt.ensure(P.n_pad * P.Kp, err)) return false;

    SSU_HIP(P.mark(0));
    if (P.bits) {   // Gram path on presence bits: one block of every sample
        if (!P.embed_bits(P.d_B, P.ldb, false, err)) return false;
        SSU_HIP(P.mark(1));
        if (P.Kp)
            SSU_HIP(launch_presence_T_bits(P.d_Pt.as<int8_t>(), P.Kp, P.d_B.as<uint32_t>(), P.ldb,
                                           P.d_col.as<int32_t>(), (uint32_t)P.Kp, (uint32_t)P.n_pad, P.n, P.st));
        if (gsplit)
            SSU_HIP(launch_bits_smask(P.d_smask.as<uint32_t>(), P.d_B.as<uint32_t>(), P.ldb,
                                      P.d_sp_node.as<int32_t>(), Ks, npan, P.st));
    }
    // one embedding block (E holds every sample): the postings of the FP32 split take their
    // values from E, and the gather skips the operand rows of nodes that are never dense
    const bool skip_rare = split && gather && single && !s.vaw && P.d_urare.p;
    for (size_t c0 = 0; c0 < P.ldn && !P.bits; c0 += P.blk) {
        const size_t w = std::min(P.blk, P.ldn - c0);
        if (!P.embed(P.d_E, s.normalize, c0, w, err)) return false;
        if (s.vaw && !P.embed(P.d_C, false, c0, w, err)) return false;  // raw counts, src/unifrac_cmp.cpp:250-251
        if (single) SSU_HIP(P.mark(1));
        if (P.gram) {
            if (P.Kp)
                SSU_HIP(launch_presence_T(P.d_Pt.as<int8_t>(), P.Kp, P.d_E.as<double>(), w, (uint32_t)c0,
                                          P.d_col.as<int32_t>(), (uint32_t)P.Kp, P.n, P.st));
            if (gsplit)
                SSU_HIP(launch_gather_sums(nullptr, 0, nullptr, nullptr, P.d_umasks.as<uint32_t>(),
                                           P.d_smask.as<uint32_t>(), P.d_E.as<double>(), w, (uint32_t)c0,
                                           P.d_sp_node.as<int32_t>(), nullptr, Ks, P.n, P.st));
        } else if (gather || sums || shared) {
            SSU_HIP(launch_gather_sums(gather ? P.d_U.as<float>() : nullptr, P.ldn, sums ? P.S_ptr() : nullptr,
                                       P.d_Spart.as<double>(), shared ? P.d_masks.as<uint32_t>() : nullptr,
                                       split ? P.d_smask.as<uint32_t>() : nullptr, P.d_E.as<double>(), w, (uint32_t)c0, P.d_knodes.as<int32_t>(),
                                       P.d_klen_sum.as<double>(), K, P.n, P.st,
                                       skip_rare ? P.d_urare.as<uint32_t>() : nullptr,
                                       s.fixed_segments ? fixed_rows_per_segment(K) : 0u));
            if (gather && s.vaw)
                SSU_HIP(launch_gather_sums(P.d_Cu.as<float>(), P.ldn, nullptr, nullptr, nullptr, nullptr, P.d_C.as<double>(), w,
                                           (uint32_t)c0, P.d_knodes.as<int32_t>(), nullptr, K, P.n, P.st));
        }
    }
    if (!single && !P.bits) SSU_HIP(P.mark(1));   // blocked: embed and prep share one phase
    if (split) {   // posting lists, row lists and each row's sparse-node share, into the output
        SSU_HIP(P.mark(EV_SPLIT));
        SparseNodesArgs b{};
        b.dmask = P.d_dmask.as<uint32_t>();
        b.smask = P.d_smask.as<uint32_t>();
        b.pc = P.d_pc.as<uint32_t>();
        b.pidx = P.d_pidx.p;
        b.pval = P.d_pval.p;
        b.U = s.fp64 || skip_rare ? P.d_E.p : P.d_U.p;          // the embedding rows themselves
        b.urows = s.fp64 || skip_rare ? P.d_knodes.as<int32_t>() : nullptr;
        b.values_fp64 = s.fp64;
        b.src_fp64 = s.fp64 || skip_rare;
        b.ldu = P.ldn;
        b.len = P.d_klen_sum.as<double>();
        b.S = P.S_ptr();
        b.out = P.out_ptr();
        b.out_fp64 = s.fp64;
        b.walked = P.d_walked.as<uint32_t>();
        b.rows = P.d_rows.as<unsigned long long>();
        b.rcount = P.d_rcount.as<uint32_t>();
        b.roff = P.d_roff.as<uint32_t>();
        b.rbsum = P.d_rbsum.as<uint32_t>();
        b.K = K;
        b.n = P.n;
        b.npan = npan;
        b.poff = P.d_poff.as<uint32_t>();
        b.pcnt = P.d_pcnt.as<uint32_t>();
        b.obsum = P.d_obsum.as<uint32_t>();
        b.slots = P.split_slots;
        b.thr = P.split_thr;
        b.ssp = P.d_Ssp.as<double>();
        b.gen = P.gen;
        b.alpha = s.alpha;
        b.out2 = P.gen || vsplit ? P.d_out2.p : nullptr;
        b.pspmax = P.gen ? P.d_pspmax.as<double>() : nullptr;
        b.band = P.band;
        if (vsplit) {
            const bool unw = s.method == Method::unweighted || s.method == Method::unweighted_unnormalized;
            b.vaw = s.method == Method::generalized ? 3 : unw ? 2 : 1;
            b.fixed_dense = P.d_vdense.as<uint32_t>();
            b.Cv = s.fp64 ? P.d_C.p : P.d_Cu.p;
            b.pcv = P.d_pcv.p;
            b.tot = P.totals_ptr();
            b.tau = P.x_totals ? P.x_totals : P.d_tau.as<double>();
            b.kidx = P.x_totals ? nullptr : P.d_kidx.as<uint32_t>();
            b.D = vD;
            b.H = P.d_H.as<double>();
            b.nbound = P.d_nbound.as<double>();
        }
        SSU_HIP(launch_sparse_nodes(b, P.st));
    }
    SparseNodesArgs gb{};
    if (gsplit) {   // posting lists and row lists of the Gram path's sparse nodes
        SSU_HIP(P.mark(EV_SPLIT));
        gb.dmask = P.d_dmask.as<uint32_t>();
        gb.smask = P.d_smask.as<uint32_t>();
        gb.pc = P.d_pc.as<uint32_t>();
        gb.pidx = P.d_pidx.p;
        gb.len = P.d_sp_len.as<double>();
        gb.walked = P.d_walked.as<uint32_t>();
        gb.rows = P.d_rows.as<unsigned long long>();
        gb.rcount = P.d_rcount.as<uint32_t>();
        gb.roff = P.d_roff.as<uint32_t>();
        gb.rbsum = P.d_rbsum.as<uint32_t>();
        gb.K = Ks;
        gb.n = P.n;
        gb.npan = npan;
        gb.poff = P.d_poff.as<uint32_t>();
        gb.slots = P.gsplit_slots;
        gb.gram = P.d_G.as<double>();
        gb.n_pad = (uint32_t)P.n_pad;
        gb.band = P.band;
        SSU_HIP(launch_sparse_lists(gb, P.st));
    }
    SSU_HIP(P.mark(2));

    const bool norm = s.method == Method::unweighted;
    auto combine = [&](uint32_t r0, uint32_t r1) {
        return s.fp64 ? launch_unweighted_combine<double>((double*)P.out_ptr(), P.d_G.as<double>(), P.n_pad, P.n,
                                                          norm, P.st, r0, r1)
                      : launch_unweighted_combine<float>((float*)P.out_ptr(), P.d_G.as<double>(), P.n_pad, P.n, norm,
                                                         P.st, r0, r1);
    };
    // banded: each band's pairs from its first row's column on, at (i, j) and (j, i) (an
    // L of final entries per band, for the overlapped copies)
    auto combine_band = [&](uint32_t r0, uint32_t r1) {
        return s.fp64 ? launch_unweighted_combine_band<double>((double*)P.out_ptr(), P.d_G.as<double>(), P.n_pad,
                                                               P.n, norm, P.st, r0, r1)
                      : launch_unweighted_combine_band<float>((float*)P.out_ptr(), P.d_G.as<double>(), P.n_pad, P.n,
                                                              norm, P.st, r0, r1);
    };
    if (P.gram && P.banded_run) {
        // bands in row order: the Gram tiles of the band's tile rows (band 0: every
        // diagonal tile too), then (band 0) the diagonal shares of every row, the
        // band's rows' posting walk and its rows of the output
        for (uint32_t k = 0; k + 1 < (uint32_t)P.band_off.size(); ++k) {
            const uint32_t r0 = P.band_row0[k], r1 = P.band_row0[k + 1];
            SSU_HIP(launch_gram_i8(P.d_G.as<double>(), P.d_Pt.as<int8_t>(), P.Kp, P.d_digits.as<int8_t>(),
                                   P.d_chunks.as<GramChunk>(), (uint32_t)P.chunks.size(), (uint32_t)P.n_pad,
                                   P.d_order_band.as<uint32_t>() + P.band_off[k], P.band_off[k + 1] - P.band_off[k],
                                   P.st));
            if (gsplit) {
                if (k == 0) SSU_HIP(launch_sparse_rows(gb, P.st, 0, P.n, ROWS_DIAGONAL));
                SSU_HIP(launch_sparse_rows(gb, P.st, r0, r1, ROWS_OFF_DIAGONAL));
            }
            SSU_HIP(combine_band(r0, r1));
            SSU_HIP(hipEventRecord(P.band_ev[k], P.st));
        }
        SSU_HIP(P.mark(3));
        if (gsplit) SSU_HIP(P.mark(EV_ROWS));
        SSU_HIP(P.mark(4));
    } else if (P.gram) {
        if (P.Kp) {
            SSU_HIP(launch_gram_i8(P.d_G.as<double>(), P.d_Pt.as<int8_t>(), P.Kp, P.d_digits.as<int8_t>(),
                                   P.d_chunks.as<GramChunk>(), (uint32_t)P.chunks.size(), (uint32_t)P.n_pad,
                                   P.band_all ? P.d_order.as<uint32_t>() : P.d_order_run.as<uint32_t>(),
                                   (uint32_t)P.order_used().size(), P.st));
        } else {
            SSU_HIP(launch_fill(P.d_G.as<double>(), P.n_pad * P.n_pad, 0.0, P.st));   // no memset node in captures
        }
        SSU_HIP(P.mark(3));
        if (gsplit) {
            SSU_HIP(launch_sparse_rows(gb, P.st));   // adds the sparse nodes' share to G
            SSU_HIP(P.mark(EV_ROWS));
        }
        SSU_HIP(combine(0, P.n));
        SSU_HIP(P.mark(4));
    } else if (P.sparse) {
        SparseArgs a{};
        a.U = s.fp64 ? P.d_E.p : P.d_U.p;
        a.rows = s.fp64 ? P.d_knodes.as<int32_t>() : nullptr;
        a.ldn = P.ldn;
        a.len = P.d_klen_sum.as<double>();
        a.masks = P.d_masks.as<uint32_t>();
        a.Kw = Kw;
        a.n = P.n;
        a.S = P.S_ptr();
        a.normalized = s.method == Method::weighted_normalized;
        a.out = P.out_ptr();
        a.out_fp64 = s.fp64;
        const size_t nb = (P.n + 63) / 64;
        if (!P.d_visits.ensure(std::max<size_t>(2 * nb * nb, 1) * sizeof(uint32_t), err)) return false;
        a.visits = P.d_visits.as<uint32_t>();
        a.order = P.band_all ? P.d_order.as<uint32_t>() : P.d_order_run.as<uint32_t>();
        a.n_blocks = (uint32_t)P.order_used().size();
        if (split) a.dmask = P.d_dmask.as<uint32_t>();   // dense nodes only; adds the posting-list share
        const uint32_t nbands = P.banded_run ? (uint32_t)P.band_ev.size() : 1;
        if (P.banded_run && nbands > 1) {
            SSU_HIP(hipEventRecord(P.band_go, P.st));
            SSU_HIP(hipStreamWaitEvent(P.band_st2, P.band_go, 0));
        }
        for (uint32_t k = 0; k < nbands; ++k) {
            hipStream_t bs = P.st;
            if (P.banded_run) {
                a.order = P.d_order_band.as<uint32_t>() + P.band_off[k];
                a.n_blocks = P.band_off[k + 1] - P.band_off[k];
                if (k & 1) bs = P.band_st2;
            }
            if (P.gen && s.fp64) SSU_HIP(launch_pairs_gen64(a, s.alpha, split ? P.d_out2.p : nullptr, bs));
            else if (P.gen) SSU_HIP(launch_pairs_gen32(a, (float)s.alpha, split ? P.d_out2.p : nullptr, bs));
            else if (s.fp64) SSU_HIP(launch_pairs_sparse(true, a, bs));
            else SSU_HIP(launch_pairs_sparse32(a, bs));
            if (P.banded_run) SSU_HIP(hipEventRecord(P.band_ev[k], bs));
        }
        if (P.banded_run && nbands > 1)   // the later work on st reads every band
            SSU_HIP(hipStreamWaitEvent(P.st, P.band_ev[(nbands - 2) | 1], 0));   // the last band on band_st2
        SSU_HIP(P.mark(3));
        SSU_HIP(P.mark(4));
    } else {
        PairArgs a{};
        a.n = P.n;
        a.K = K;
        a.ldn = P.ldn;
        a.alpha = s.alpha;
        a.totals = P.totals_ptr();
        a.out = P.out_ptr();
        a.out_fp64 = s.fp64;
        a.len = P.d_klen.p;
        a.normalized = normalized;
        a.S = sums ? P.S_ptr() : nullptr;
        a.band = P.band;
        if (s.fp64) {
            a.U = P.d_E.p;
            a.Cn = s.vaw ? P.d_C.p : nullptr;
            a.rows = P.d_knodes.as<int32_t>();
        } else {
            a.U = P.d_U.p;
            a.Cn = s.vaw ? P.d_Cu.p : nullptr;
            a.rows = nullptr;
        }
        if (vsplit) {   // the dense nodes only, plus the posting-list shares
            a.rows = P.d_vrows.as<int32_t>();
            a.len = P.d_vlen.p;
            a.K = P.n_vdense;
            a.out2 = P.d_out2.p;
            a.len_nonneg = true;   // (the split needs every length finite and >= 0)
        }
        if (P.refine_vaw) {   // the kernel marks the near-duplicates itself
            uint32_t* vf = d_flags;
            if (!vf && !quiet) {
                if (!P.d_close.ensure(((size_t)P.n + 1) * sizeof(uint32_t), err)) return false;
                vf = P.d_close.as<uint32_t>();
            }
            if (vf) {
                SSU_HIP(launch_fill_u32(vf, (size_t)P.n + 1, 0u, P.st));
                a.flags = vf;
                a.flag_tau = P.refine_tau;
            }
        }
        SSU_HIP(launch_pairs(op, s.fp64, a, P.st));
        SSU_HIP(P.mark(3));
        SSU_HIP(P.mark(4));
    }
    // marks for the near-duplicate refinement (the plan's own buffer unless the caller brings one)
    uint32_t* flags = d_flags;
    if (P.refine_on && !flags && !quiet) {
        if (!P.d_close.ensure(((size_t)P.n + 1) * sizeof(uint32_t), err)) return false;
        flags = P.d_close.as<uint32_t>();
    }
    if (P.refine_on && flags && !P.refine_vaw)
        SSU_HIP(launch_flag_close(flags, (const float*)P.out_ptr(), P.n, P.refine_S ? P.S_ptr() : nullptr,
                                  P.refine_tau, P.band, P.st));
    SSU_HIP(P.mark(5));
    if (P.quiet) return true;
    if (P.after_enqueue && !P.after_enqueue(err)) return false;   // overlapping copies (execute_to_host)
    SSU_HIP(hipStreamSynchronize(P.st));
    for (int i = 0; i < PH_TOTAL; ++i) SSU_HIP(hipEventElapsedTime(&P.times[i], P.ev[i], P.ev[i + 1]));
    SSU_HIP(hipEventElapsedTime(&P.times[PH_TOTAL], P.ev[0], P.ev[5]));
    P.times[PH_SPARSE] = 0.0f;
    if (split) {   // the structural bound holds by construction; a breach would have skipped the lists
        uint32_t total = 0;
        SSU_HIP(hipMemcpy(&total, P.d_poff.as<uint32_t>() + K, sizeof(uint32_t), hipMemcpyDeviceToHost));
        if (total > P.split_slots) {
            if (err) *err = "posting lists exceed the plan's structural bound";
            return false;
        }
    }
    if (split || gsplit) {   // prep ends where the posting-list phase starts
        SSU_HIP(hipEventElapsedTime(&P.times[PH_PREP], P.ev[1], P.ev[EV_SPLIT]));
        SSU_HIP(hipEventElapsedTime(&P.times[PH_SPARSE], P.ev[EV_SPLIT], P.ev[2]));
    }
    if (gsplit) {   // the rows kernel runs between the Gram and the combine
        float rows_ms = 0.0f;
        SSU_HIP(hipEventElapsedTime(&rows_ms, P.ev[3], P.ev[EV_ROWS]));
        SSU_HIP(hipEventElapsedTime(&P.times[PH_COMBINE], P.ev[EV_ROWS], P.ev[4]));
        P.times[PH_SPARSE] += rows_ms;
    }
    P.refined = 0;
    if (P.refine_on && flags && !P.refine(flags, err)) return false;
    return true;
}

bool Plan::refine(const double* d_data, const double* d_totals, void* d_out, const uint32_t* d_flags, double* d_sums,
                  std::string* err) {
    PlanImpl& P = *p_;
    if (!P.refine_on) return true;
    P.x_data = d_data;
    P.x_totals = d_totals;
    P.x_out = d_out;
    P.x_S = d_sums;
    P.quiet = false;
    P.refined = 0;
    if (!d_flags && !P.d_close.p) return true;   // never marked
    if (P.refine_S && !d_sums && !P.d_S.p) { if (err) *err = "refine: no per-sample sums"; return false; }
    return P.refine(d_flags ? d_flags : P.d_close.as<uint32_t>(), err);
}

bool Plan::refines() const { return p_->refine_on; }

// Bands of about equal row counts, in whole block rows (32 FP32 / 64 FP64
// samples; Gram: 128-sample tile rows, band 0 holding every diagonal tile).
bool PlanImpl::prepare_bands(std::string* err) {
    if (!band_ev.empty()) return true;
    const uint32_t ei = gram ? GRAM_BM : (rect() ? 32 : 64);
    const uint32_t nbi = (n + ei - 1) / ei;
    // the Gram path's bands are cheaper to compute than to copy: more, smaller ones
    // (16 bands measured 27.0 ms against 28.1 ms for 10 at 16,384 samples; the
    // tiled weighted kernel 137.9 against 137.6 ms)
    uint32_t want = gram ? OUT_BANDS_GRAM : OUT_BANDS;
    if (const char* env = getenv("SSU_OUT_BANDS")) want = std::max(1, atoi(env));
    const uint32_t nb = std::min<uint32_t>(want, nbi);
    // row shares 1 : 2 : 3 ... 3 : 2 : 1 (fewer than five bands: equal)
    std::vector<double> wsum(nb + 1, 0.0);
    for (uint32_t k = 0; k < nb; ++k) {
        const uint32_t edge = std::min(k, nb - 1 - k);
        wsum[k + 1] = wsum[k] + (nb < 5 ? 1.0 : std::min<double>(edge + 1, 3));
    }
    std::vector<uint32_t> cut(nb + 1);
    for (uint32_t k = 0; k <= nb; ++k) cut[k] = (uint32_t)std::llround(nbi * wsum[k] / wsum[nb]);
    for (uint32_t k = 1; k <= nb; ++k) cut[k] = std::max(cut[k], cut[k - 1]);   // (monotone; empty bands allowed)
    std::vector<uint32_t> ord;
    band_off.assign(1, 0u);
    band_row0.clear();
    for (uint32_t k = 0; k < nb; ++k) {
        for (uint32_t bij : block_order) {
            const uint32_t bi = bij >> 16, bj = bij & 0xffffu;
            const bool in_band = bi >= cut[k] && bi < cut[k + 1];
            if (gram ? (bi == bj ? k == 0 : in_band) : in_band) ord.push_back(bij);
        }
        band_off.push_back((uint32_t)ord.size());
        band_row0.push_back(std::min(n, cut[k] * ei));
    }
    band_row0.push_back(n);
    if (!upload(d_order_band, ord.empty() ? std::vector<uint32_t>{0u} : ord, st, err)) return false;
    SSU_HIP(hipStreamSynchronize(st));   // `ord` is read until here
    if (!copy_st) SSU_HIP(hipStreamCreateWithFlags(&copy_st, hipStreamNonBlocking));
    if (!band_st2) SSU_HIP(hipStreamCreateWithFlags(&band_st2, hipStreamNonBlocking));
    if (!band_go) SSU_HIP(hipEventCreateWithFlags(&band_go, hipEventDisableTiming));
    band_ev.assign(nb, nullptr);
    for (auto& e : band_ev) SSU_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    return true;
}

bool Plan::execute_to_host(void* host, std::string* err) {
    PlanImpl& P = *p_;
    // banded copies for the shared-node kernels over the full matrix (the other
    // paths write their output in one final pass)
    static const bool off = getenv("SSU_OUT_BANDS") && atoi(getenv("SSU_OUT_BANDS")) == 0;
    const bool gram_bands = P.gram && P.Kp;   // (no Gram columns: one fill, nothing to band)
    if (off || !(P.sparse || gram_bands) || !P.band_all || P.n < 4096 || P.block_order.empty())
        return execute(err) && copy_matrix(host, err);
    if (!P.prepare_bands(err)) return false;
    const size_t es = P.spec.fp64 ? sizeof(double) : sizeof(float);
    P.banded_run = true;
    P.after_enqueue = [&P, host, es](std::string* e) {
        std::vector<CopyRange> ranges;
        const size_t n = P.n, pitch = n * es;
        for (size_t k = 0; k < P.band_ev.size(); ++k) {
            const size_t r0 = P.band_row0[k], r1 = P.band_row0[k + 1];
            if (P.sparse || P.gram) {
                // band k writes every pair (i, j), i < j, of its rows i and the mirror
                // (j, i) (the tiled kernel's blocks; the Gram path's banded combine):
                // once it is done, its rows from column r0 on and the columns [r0, r1)
                // of the rows below are final, an L whose size follows the band's work
                // (the earlier columns of its rows came with earlier bands)
                ranges.push_back(CopyRange{(r0 * n + r0) * es, (n - r0) * es, {(void*)P.band_ev[k]}, r1 - r0, pitch});
                if (r1 < n)
                    ranges.push_back(CopyRange{(r1 * n + r0) * es, (r1 - r0) * es, {(void*)P.band_ev[k]}, n - r1, pitch});
                continue;
            }
            // band k's rows are final once it and the band before it are done
            CopyRange r{r0 * pitch, (r1 - r0) * pitch, {(void*)P.band_ev[k]}};
            if (k > 0) r.ready.push_back((void*)P.band_ev[k - 1]);
            ranges.push_back(r);
        }
        return copy_ranges_to_host((char*)host, P.d_out.as<char>(), ranges, P.copy_st, e, true);
    };
    const bool ok = execute(err);
    P.banded_run = false;
    P.after_enqueue = nullptr;
    if (!ok) return false;
    if (!P.refined) return true;
    // refinement rewrote entries after the bands were copied, all among the refined
    // samples: their square of entries again, or the whole matrix when that is
    // more than a quarter of it (FP32 output: FP64 plans are not refined)
    const std::vector<uint32_t>& F = P.refined_samples;
    const size_t ns = F.size(), nn = (size_t)P.n * P.n;
    if (P.spec.fp64 || ns > 65535 || ns * ns * 4 > nn) return copy_matrix(host, err);
    DevMem sub;
    std::vector<float> h(ns * ns);
    if (!upload(P.d_ridx, F, P.st, err) || !sub.ensure(std::max<size_t>(ns * ns, 1) * sizeof(float), err))
        return false;
    SSU_HIP(launch_gather_square(sub.as<float>(), (const float*)P.d_out.p, P.n, P.d_ridx.as<uint32_t>(),
                                 (uint32_t)ns, P.st));
    if (!copy_to_host(h.data(), sub.p, ns * ns * sizeof(float), P.st, err)) return false;
    SSU_HIP(hipStreamSynchronize(P.st));
    float* out = (float*)host;
    for (size_t a = 0; a < ns; ++a)
        for (size_t b = 0; b < ns; ++b) out[(size_t)F[a] * P.n + F[b]] = h[a * ns + b];
    return true;
}

bool Plan::set_stripes(uint32_t start, uint32_t stop, std::string* err) {
    PlanImpl& P = *p_;
    const uint32_t total = (P.n + 1) / 2;
    if (stop > total) {
        if (err) *err = "stripe range [" + std::to_string(start) + ", " + std::to_string(stop) + ") outside the " +
                        std::to_string(total) + " stripes";
        return false;
    }
    if (start == 0 && (stop == 0 || stop == total)) {
        P.band = GapBand();
        P.band_all = true;
        P.run_order.clear();
        return true;
    }
    P.band = stripe_band(P.n, start, stop);
    P.band_all = false;
    // the blocks / tiles of the kernel's order that hold a selected pair
    P.run_order.clear();
    uint32_t ei = 0, ej = 0;   // block edges (rows, columns)
    if (P.gram) ei = ej = GRAM_BM;
    else if (P.sparse && !P.rect()) ei = ej = 64;
    else if (P.sparse) { ei = 32; ej = 64; }
    for (uint32_t bij : P.block_order) {
        const uint32_t bi = bij >> 16, bj = bij & 0xffffu;
        const bool diag_tile = P.gram && bi == bj;   // diag(G) feeds every pair of the combine
        if (diag_tile || P.band.meets_tile(bi * ei, std::min(P.n, bi * ei + ei), bj * ej, std::min(P.n, bj * ej + ej)))
            P.run_order.push_back(bij);
    }
    if (!P.block_order.empty()) {
        if (!upload(P.d_order_run, P.run_order.empty() ? std::vector<uint32_t>{0u} : P.run_order, P.st, err)) return false;
        if (hipStreamSynchronize(P.st) != hipSuccess) { if (err) *err = "stripe order upload failed"; return false; }
    }
    return true;
}
uint64_t Plan::refined_pairs() const { return p_->refined; }

// The marked pairs, recomputed by the FP64 method on the sub-table of the
// samples involved: the same tree, the same per-sample values and totals, the
// dense FP64 pair kernel (each pair's terms summed over every node in order)
// and per-sample sums in segments fixed by the node count, so a pair's FP64
// distance does not depend on which other samples the sub-table holds -- a
// stripe range and the full matrix refine a pair to the same value.  Each pair
// is marked by its own FP32 distance (and S), tested again when its FP64 value
// is written.  The samples are taken in groups whose FP64
// plan fits in REFINE_BUDGET; with more than one group, every pair of groups
// runs as one sub-table and writes only its cross pairs.
bool PlanImpl::refine(const uint32_t* flags, std::string* err) {
    uint32_t count = 0;
    SSU_HIP(hipStreamSynchronize(st));
    SSU_HIP(hipMemcpy(&count, flags, sizeof(uint32_t), hipMemcpyDeviceToHost));
    if (!count) return true;
    std::vector<uint32_t> mark(n);
    SSU_HIP(hipMemcpy(mark.data(), flags + 1, (size_t)n * sizeof(uint32_t), hipMemcpyDeviceToHost));
    std::vector<uint32_t>& F = refined_samples;
    F.clear();
    for (uint32_t s = 0; s < n; ++s)
        if (mark[s]) F.push_back(s);
    // the table as this execution sees it
    std::vector<uint32_t> ip(n_obs + 1), ix(nnz);
    std::vector<double> vals(nnz), tot(n);
    SSU_HIP(hipMemcpy(ip.data(), d_indptr.p, ip.size() * sizeof(uint32_t), hipMemcpyDeviceToHost));
    if (nnz) {
        SSU_HIP(hipMemcpy(ix.data(), d_indices.p, nnz * sizeof(uint32_t), hipMemcpyDeviceToHost));
        SSU_HIP(hipMemcpy(vals.data(), data_ptr(), nnz * sizeof(double), hipMemcpyDeviceToHost));
    }
    if (n) SSU_HIP(hipMemcpy(tot.data(), totals_ptr(), (size_t)n * sizeof(double), hipMemcpyDeviceToHost));

    constexpr double REFINE_BUDGET = 24.0 * (1u << 30);
    // FP64 plan of g samples: the embedding m x g and the result g x g, 8 bytes each, twice over for scratch
    auto fits = [&](double samples) { return 16.0 * ((double)std::max<uint32_t>(m, 1) * samples + samples * samples) <= REFINE_BUDGET; };
    uint32_t g = 256;   // a sub-table holds up to 2 g samples
    while (g < 16384 && fits(4.0 * g)) g *= 2;
    std::vector<std::pair<uint32_t, uint32_t>> groups;   // [begin, end) in F
    const uint32_t nf = (uint32_t)F.size();
    if (nf <= 2 * g) groups.push_back({0, nf});
    else
        for (uint32_t b = 0; b < nf; b += g) groups.push_back({b, std::min(nf, b + g)});
    ComputeSpec sp = spec;
    sp.fp64 = true;
    sp.plain_kernel = 1;
    sp.fixed_segments = true;
    std::vector<int64_t> local(n, -1);
    for (size_t A = 0; A < groups.size(); ++A)
        for (size_t B = A; B < groups.size(); ++B) {
            std::vector<uint32_t> idx;
            for (uint32_t q = groups[A].first; q < groups[A].second; ++q) idx.push_back(F[q]);
            const uint32_t na = (uint32_t)idx.size();
            if (B != A)
                for (uint32_t q = groups[B].first; q < groups[B].second; ++q) idx.push_back(F[q]);
            const uint32_t ns = (uint32_t)idx.size();
            for (uint32_t a = 0; a < ns; ++a) local[idx[a]] = a;
            std::vector<uint32_t> sip(1, 0), six;
            std::vector<double> sdata;
            for (uint32_t i = 0; i < n_obs; ++i) {
                for (uint32_t j = ip[i]; j < ip[i + 1]; ++j)
                    if (local[ix[j]] >= 0) { six.push_back((uint32_t)local[ix[j]]); sdata.push_back(vals[j]); }
                sip.push_back((uint32_t)six.size());
            }
            for (uint32_t a = 0; a < ns; ++a) local[idx[a]] = -1;
            Table sub = Table::own(std::vector<std::string>(n_obs), std::vector<std::string>(ns), std::move(sip),
                                   std::move(six), std::move(sdata));
            for (uint32_t a = 0; a < ns; ++a) sub.totals[a] = tot[idx[a]];
            std::unique_ptr<Plan> sp_plan = Plan::create(sub, nodes, sp, err);
            if (!sp_plan || !sp_plan->execute(err)) return false;
            if (!upload(d_ridx, idx, st, err)) return false;
            // (refine_vaw: every pair of the marked samples, the FP64 value being the better one)
            SSU_HIP(launch_refine_scatter((float*)out_ptr(), n, (const double*)sp_plan->device_output(), ns,
                                          d_ridx.as<uint32_t>(), na, B != A, refine_S ? S_ptr() : nullptr,
                                          refine_vaw ? INFINITY : refine_tau, st));
            SSU_HIP(hipStreamSynchronize(st));   // idx and the sub-plan's result are read until here
        }
    refined = count;
    return true;
}

bool Plan::copy_matrix(void* host, std::string* err) const {
    const PlanImpl& P = *p_;
    const size_t bytes = (size_t)P.n * P.n * (P.spec.fp64 ? sizeof(double) : sizeof(float));
    return copy_to_host(host, P.d_out.p, bytes, P.st, err);
}

bool Plan::copy_rows(uint32_t row0, uint32_t nrows, void* host, std::string* err) const {
    const PlanImpl& P = *p_;
    if ((uint64_t)row0 + nrows > P.n) {
        if (err) *err = "ssu_plan_copy_rows: rows " + std::to_string(row0) + ".." + std::to_string((uint64_t)row0 + nrows) +
                        " outside the " + std::to_string(P.n) + "-sample matrix";
        return false;
    }
    const size_t es = P.spec.fp64 ? sizeof(double) : sizeof(float);
    return copy_to_host(host, P.d_out.as<char>() + (size_t)row0 * P.n * es, (size_t)nrows * P.n * es, P.st, err);
}

bool Plan::copy_stripes(uint32_t start, uint32_t stop, double* const* host, std::string* err) const {
    PlanImpl& P = *p_;
    if (stop <= start) return true;
    const size_t rows = stop - start;
    if (!P.d_stripes.ensure(rows * P.n * sizeof(double), err)) return false;
    if (P.spec.fp64)
        SSU_HIP(launch_matrix_to_stripes<double>(P.d_stripes.as<double>(), P.d_out.as<double>(), P.n, start, stop, P.st));
    else
        SSU_HIP(launch_matrix_to_stripes<float>(P.d_stripes.as<double>(), P.d_out.as<float>(), P.n, start, stop, P.st));
    std::vector<double> tmp(rows * P.n);
    if (!copy_to_host(tmp.data(), P.d_stripes.p, tmp.size() * sizeof(double), P.st, err)) return false;
    for (size_t r = 0; r < rows; ++r) std::memcpy(host[r], tmp.data() + r * P.n, P.n * sizeof(double));
    return true;
}

bool Plan::faith(double* host, std::string* err) {
    PlanImpl& P = *p_;
    P.x_data = nullptr;    // the plan's own table
    P.x_totals = nullptr;
    if (!P.d_faith.ensure(std::max<uint32_t>(P.n, 1) * sizeof(double), err)) return false;
    for (size_t c0 = 0; c0 < P.ldn && c0 < P.n; c0 += P.blk) {
        const size_t w = std::min(P.blk, P.ldn - c0);
        if (!P.embed(P.d_E, true, c0, w, err)) return false;
        SSU_HIP(launch_faith(P.d_faith.as<double>() + c0, P.d_E.as<double>(), w, P.d_len_all.as<double>(), P.m,
                             (uint32_t)std::min<size_t>(w, P.n - c0), P.st));
    }
    SSU_HIP(hipStreamSynchronize(P.st));
    if (P.n) SSU_HIP(hipMemcpy(host, P.d_faith.p, P.n * sizeof(double), hipMemcpyDeviceToHost));
    return true;
}

const float* Plan::timings() const { return p_->times; }
void Plan::set_stream(void* stream) {
    (void)hipStreamSynchronize(p_->st);   // work already queued finishes before later work goes elsewhere
    p_->st = stream ? (hipStream_t)stream : p_->own_st;
}
void* Plan::stream() const { return p_->st; }
uint32_t Plan::n_samples() const { return p_->n; }
uint64_t Plan::processed_nodes() const { return p_->knodes.size(); }

uint64_t Plan::kernel_terms() const {
    const PlanImpl& P = *p_;
    const uint64_t K = P.knodes.size();
    if (P.gram) return (uint64_t)P.order_used().size() * GRAM_BM * GRAM_BM * (uint64_t)P.Kp;
    if (P.sparse) {
        // visited nodes per block, times the pairs a block evaluates
        const size_t nbi = P.rect() ? (P.n + 31) / 32 : (P.n + 63) / 64, nbj = (P.n + 63) / 64;
        if (!nbj || !P.d_visits.p) return 0;
        std::vector<uint32_t> v(nbi * nbj);
        if (hipMemcpy(v.data(), P.d_visits.p, v.size() * sizeof(uint32_t), hipMemcpyDeviceToHost) != hipSuccess)
            return 0;
        uint64_t sum = 0;
        for (uint32_t bij : P.order_used()) sum += v[(size_t)(bij >> 16) * nbj + (bij & 0xffffu)];
        return sum * (P.rect() ? 32 * 64 : 64 * 64);
    }
    // dense tiles: 128 x 128 for weighted / plain unweighted, 64 x 64 for generalized and VAW
    const bool small = P.spec.vaw || P.spec.method == Method::generalized;
    const uint64_t edge = small ? 64 : PAIR_TILE;
    const uint64_t t = (P.n + edge - 1) / edge;
    return t * (t + 1) / 2 * edge * edge * (P.vsplit ? (uint64_t)P.n_vdense : K);
}
uint64_t Plan::sparse_terms() const {
    const PlanImpl& P = *p_;
    if (!(((P.sparse || P.vsplit) && P.split_thr) || (P.gram && P.gsplit_slots)) || !P.d_walked.p) return 0;
    std::vector<uint32_t> wk(sparse_rows_launched(P.n, (P.gen || P.vsplit) && !P.gram));
    if (hipMemcpy(wk.data(), P.d_walked.p, wk.size() * sizeof(uint32_t), hipMemcpyDeviceToHost) != hipSuccess) return 0;
    uint64_t sum = 0;
    for (uint32_t x : wk) sum += x;
    return sum;
}
bool Plan::out_fp64() const { return p_->spec.fp64; }
double Plan::split_theta() const {
    const PlanImpl& P = *p_;
    if (P.gram) return P.gsplit_slots ? P.split_theta : 0.0;
    return P.split_thr ? (double)P.split_thr / std::max<uint32_t>(P.n, 1) : 0.0;
}
const char* Plan::main_kernel() const {
    if (p_->gram) return "k_gram_i8r";
    if (p_->sparse)
        return p_->gen ? (p_->spec.fp64 ? "k_pairs_gen64" : "k_pairs_gen32")
                       : (p_->spec.fp64 ? "k_pairs_sparse" : "k_pairs_sparse32");
    return "k_pairs";
}
const void* Plan::device_output() const { return p_->d_out.p; }

}  // namespace ssu
