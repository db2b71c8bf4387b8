// Rarefaction replicates on the device, captured into one HIP graph.
//
// Random numbers: SplitMix64 (Steele, Lea & Flood, "Fast splittable
// pseudorandom number generators", OOPSLA 2014) applied to a counter built
// from (seed, replicate, sample, draw), so every draw is independent of launch
// geometry and replays reproduce the same tables.  Parity with the
// reference's std::mt19937-based sampler is statistical, not bitwise.
#include "ssu_rarefy.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <unordered_set>

#include "ssu_kernels.hpp"

namespace ssu {

namespace {

#define RF_HIP(x)                                                                                    \
    do {                                                                                             \
        hipError_t e_ = (x);                                                                         \
        if (e_ != hipSuccess) {                                                                      \
            if (err) *err = std::string(#x) + ": " + hipGetErrorString(e_);                          \
            return false;                                                                            \
        }                                                                                            \
    } while (0)

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__device__ __forceinline__ uint64_t stream_key(uint64_t seed, uint32_t rep, uint32_t sample) {
    return mix64(seed ^ mix64(((uint64_t)rep << 32) | sample));
}

// first q in [0, len) with cum[q] > u (cum ascending, cum[len-1] > u)
__device__ __forceinline__ uint32_t upper_pos(const double* __restrict__ cum, uint32_t len, double u) {
    uint32_t lo = 0, hi = len - 1;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (cum[mid] > u) hi = mid;
        else lo = mid + 1;
    }
    return lo;
}

constexpr uint32_t RF_THREADS = 256;
constexpr uint32_t RF_LDS = 8192;   // per-sample counters kept in LDS up to this many observations

// With replacement: `depth` draws, observation q of sample s with probability
// count_q / total_s (std::discrete_distribution semantics).  One workgroup per
// (sample, replicate); counts gather in LDS, or in the output for wide samples.
__global__ void __launch_bounds__(RF_THREADS) k_rarefy_with(double* __restrict__ out, size_t nnz,
                                                            const uint32_t* __restrict__ col_ptr,
                                                            const uint32_t* __restrict__ col_pos,
                                                            const double* __restrict__ col_cum, uint32_t depth,
                                                            uint64_t seed) {
    const uint32_t s = blockIdx.x, rep = blockIdx.y;
    const uint32_t b = col_ptr[s], len = col_ptr[s + 1] - b;
    double* o = out + (size_t)rep * nnz;
    __shared__ uint32_t cnt[RF_LDS];
    if (len == 0) return;
    const bool lds = len <= RF_LDS;
    for (uint32_t q = threadIdx.x; q < len; q += RF_THREADS) {
        if (lds) cnt[q] = 0;
        else o[col_pos[b + q]] = 0.0;
    }
    __syncthreads();
    const double* cum = col_cum + b;
    const double total = cum[len - 1];
    const uint64_t key = stream_key(seed, rep, s);
    for (uint32_t j = threadIdx.x; j < depth; j += RF_THREADS) {
        const double u = (double)(mix64(key + j) >> 11) * 0x1.0p-53 * total;
        const uint32_t q = upper_pos(cum, len, u);
        if (lds) atomicAdd(&cnt[q], 1u);
        else atomicAdd(&o[col_pos[b + q]], 1.0);
    }
    __syncthreads();
    if (lds)
        for (uint32_t q = threadIdx.x; q < len; q += RF_THREADS) o[col_pos[b + q]] = (double)cnt[q];
}

// Without replacement: `depth` distinct items out of the sample's total_s
// (integer) items, by Floyd's algorithm (Bentley & Floyd, CACM 1987), one
// thread per (sample, replicate) with an open-addressing set of hsize slots
// in `set`; each chosen item counts for the observation that holds it.
__global__ void __launch_bounds__(64) k_rarefy_without(double* __restrict__ out, size_t nnz,
                                                       const uint32_t* __restrict__ col_ptr,
                                                       const uint32_t* __restrict__ col_pos,
                                                       const double* __restrict__ col_cum, uint32_t n,
                                                       uint32_t reps, uint32_t depth, uint64_t seed,
                                                       uint64_t* __restrict__ set, uint32_t hsize) {
    const uint32_t t = blockIdx.x * 64 + threadIdx.x;
    if (t >= n * reps) return;
    const uint32_t s = t % n, rep = t / n;
    const uint32_t b = col_ptr[s], len = col_ptr[s + 1] - b;
    double* o = out + (size_t)rep * nnz;
    if (len == 0) return;
    for (uint32_t q = 0; q < len; ++q) o[col_pos[b + q]] = 0.0;
    const double* cum = col_cum + b;
    const uint64_t total = (uint64_t)cum[len - 1];
    if (total <= depth) {   // everything is taken
        double prev = 0.0;
        for (uint32_t q = 0; q < len; ++q) { o[col_pos[b + q]] = cum[q] - prev; prev = cum[q]; }
        return;
    }
    uint64_t* h = set + (size_t)t * hsize;
    for (uint32_t i = 0; i < hsize; ++i) h[i] = ~0ull;
    const uint64_t key = stream_key(seed, rep, s);
    const uint32_t mask = hsize - 1;
    uint32_t j_i = 0;
    for (uint64_t j = total - depth; j < total; ++j, ++j_i) {
        uint64_t pick = mix64(key + j_i) % (j + 1);
        // insert pick, or j when pick is already taken
        for (int pass = 0; pass < 2; ++pass) {
            uint32_t slot = (uint32_t)mix64(pick) & mask;
            bool found = false;
            while (h[slot] != ~0ull) {
                if (h[slot] == pick) { found = true; break; }
                slot = (slot + 1) & mask;
            }
            if (!found) { h[slot] = pick; break; }
            pick = j;
        }
    }
    for (uint32_t i = 0; i < hsize; ++i)
        if (h[i] != ~0ull) {
            const uint32_t q = upper_pos(cum, len, (double)h[i]);
            o[col_pos[b + q]] += 1.0;
        }
}

}  // namespace

struct RarefyImpl {
    Table table;
    std::vector<uint32_t> sample_index, obs_index;
    std::vector<ComputeSpec> specs;
    // lanes x methods plans, plans[lane * specs.size() + k]: the replicates go round
    // robin over the lanes, each lane on a stream of its own, so small tables'
    // replicates run side by side (one lane's kernels alone leave most CUs idle)
    std::vector<std::unique_ptr<Plan>> plans;
    uint32_t lanes = 1;
    std::vector<hipStream_t> lane_st;
    std::vector<hipEvent_t> lane_ev;
    hipEvent_t fork_ev = nullptr;
    Plan* plan_of(size_t k, uint32_t rep) const { return plans[(rep % lanes) * specs.size() + k].get(); }
    uint32_t depth = 0, reps = 0, hsize = 0;
    bool with = true;
    uint64_t seed = 0;
    size_t nnz = 0;
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    bool captured = false, used_graph = false;
    float ms = 0.0f;
    void *d_col_ptr = nullptr, *d_col_pos = nullptr, *d_col_cum = nullptr, *d_rdata = nullptr, *d_depth = nullptr,
         *d_set = nullptr;
    std::vector<void*> d_out;   // per method: reps consecutive n x n matrices
    // per (method, replicate): n + 1 words of near-duplicate marks (methods whose plan refines)
    uint32_t* d_flags = nullptr;
    double* d_sums = nullptr;   // per (method, replicate): the n per-sample sums S the marks were tested against
    uint64_t refined = 0;

    uint32_t* flags_of(size_t k, uint32_t r) const {
        return d_flags + (k * reps + r) * ((size_t)table.n_samples + 1);
    }
    double* sums_of(size_t k, uint32_t r) const { return d_sums + (k * reps + r) * (size_t)table.n_samples; }
    // after the replicates have run: FP64 recomputation of the marked pairs
    bool refine(std::string* err) {
        refined = 0;
        bool any = false;
        for (const auto& p : plans) any = any || p->refines();
        if (!any) return true;
        const size_t stride = (size_t)table.n_samples + 1;
        std::vector<uint32_t> counts(specs.size() * reps);
        RF_HIP(hipMemcpy2D(counts.data(), sizeof(uint32_t), d_flags, stride * sizeof(uint32_t), sizeof(uint32_t),
                           counts.size(), hipMemcpyDeviceToHost));
        for (size_t k = 0; k < specs.size(); ++k)
            for (uint32_t r = 0; r < reps; ++r) {
                Plan* pl = plan_of(k, r);
                if (!pl->refines() || !counts[k * reps + r]) continue;
                char* out = (char*)d_out[k] + out_bytes(k) * r;
                if (!pl->refine((const double*)d_rdata + (size_t)r * nnz, (const double*)d_depth, out, flags_of(k, r),
                                sums_of(k, r), err))
                    return false;
                refined += pl->refined_pairs();
            }
        return true;
    }

    ~RarefyImpl() {
        plans.clear();
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        for (void* p : {d_col_ptr, d_col_pos, d_col_cum, d_rdata, d_depth, d_set, (void*)d_flags, (void*)d_sums})
            if (p) (void)hipFree(p);
        for (void* p : d_out)
            if (p) (void)hipFree(p);
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        if (fork_ev) (void)hipEventDestroy(fork_ev);
        for (hipEvent_t e : lane_ev)
            if (e) (void)hipEventDestroy(e);
        for (hipStream_t q : lane_st)
            if (q) (void)hipStreamDestroy(q);
        if (st) (void)hipStreamDestroy(st);
    }

    size_t out_bytes(size_t k) const {
        return (size_t)table.n_samples * table.n_samples * (specs[k].fp64 ? sizeof(double) : sizeof(float));
    }

    bool enqueue(bool quiet, std::string* err) {
        const uint32_t n = table.n_samples;
        if (with) {
            hipLaunchKernelGGL(k_rarefy_with, dim3(n, reps), dim3(RF_THREADS), 0, st, (double*)d_rdata, nnz,
                               (const uint32_t*)d_col_ptr, (const uint32_t*)d_col_pos, (const double*)d_col_cum, depth,
                               seed);
        } else {
            const uint32_t threads = n * reps;
            hipLaunchKernelGGL(k_rarefy_without, dim3((threads + 63) / 64), dim3(64), 0, st, (double*)d_rdata, nnz,
                               (const uint32_t*)d_col_ptr, (const uint32_t*)d_col_pos, (const double*)d_col_cum, n,
                               reps, depth, seed, (uint64_t*)d_set, hsize);
        }
        RF_HIP(hipGetLastError());
        // fork: every lane waits for the drawn tables
        RF_HIP(hipEventRecord(fork_ev, st));
        for (hipStream_t q : lane_st) RF_HIP(hipStreamWaitEvent(q, fork_ev, 0));
        for (uint32_t r = 0; r < reps; ++r)
            for (size_t k = 0; k < specs.size(); ++k) {
                Plan* pl = plan_of(k, r);
                char* out = (char*)d_out[k] + out_bytes(k) * r;
                // non-quiet (the eager pass) refines at once; captured work leaves it to refine()
                const bool rf = pl->refines();
                if (!pl->execute_into((const double*)d_rdata + (size_t)r * nnz, (const double*)d_depth, out, quiet, err,
                                      rf ? flags_of(k, r) : nullptr, rf ? sums_of(k, r) : nullptr))
                    return false;
                if (!quiet) refined += pl->refined_pairs();
            }
        // join: the batch's stream waits for every lane
        for (uint32_t l = 0; l < (uint32_t)lane_st.size(); ++l) {
            RF_HIP(hipEventRecord(lane_ev[l], lane_st[l]));
            RF_HIP(hipStreamWaitEvent(st, lane_ev[l], 0));
        }
        return true;
    }
};

RarefyBatch::RarefyBatch() : p_(new RarefyImpl) {}
RarefyBatch::~RarefyBatch() = default;

std::unique_ptr<RarefyBatch> RarefyBatch::create(const Table& table, const PhyloTree& tree,
                                                 const std::vector<ComputeSpec>& specs, uint32_t depth,
                                                 bool with_replacement, uint64_t seed, uint32_t reps,
                                                 std::string* err) {
    std::string e2;
    if (!err) err = &e2;
    if (!device_available(err)) return nullptr;
    if (depth == 0 || reps == 0) { *err = "depth and replicates must be non-zero"; return nullptr; }
    std::unique_ptr<RarefyBatch> b(new RarefyBatch());
    RarefyImpl& R = *b->p_;
    R.specs = specs;
    R.depth = depth;
    R.reps = reps;
    R.with = with_replacement;
    R.seed = seed;
    for (uint32_t s = 0; s < table.n_samples; ++s)
        if (table.totals[s] >= depth) R.sample_index.push_back(s);
    R.table = table.filtered((double)depth);
    if (R.table.n_samples == 0 || R.table.n_obs == 0) { *err = "no sample reaches the rarefaction depth"; return nullptr; }
    // the kept rows, in order (Table::filtered keeps rows with an entry in a kept sample)
    {
        std::vector<char> keep(table.n_samples, 0);
        for (uint32_t s : R.sample_index) keep[s] = 1;
        for (uint32_t i = 0; i < table.n_obs; ++i)
            for (uint32_t j = table.indptr[i]; j < table.indptr[i + 1]; ++j)
                if (keep[table.indices[j]]) { R.obs_index.push_back(i); break; }
    }
    const Table& t = R.table;
    const uint32_t n = t.n_samples;
    R.nnz = t.nnz();
    const uint32_t base = t.indptr[0];
    // sample-major view: per kept sample, its CSR positions and cumulative counts
    std::vector<uint32_t> col_ptr(n + 1, 0), col_pos(R.nnz);
    std::vector<double> col_cum(R.nnz);
    for (uint64_t j = 0; j < R.nnz; ++j) ++col_ptr[t.indices[base + j] + 1];
    for (uint32_t s = 0; s < n; ++s) col_ptr[s + 1] += col_ptr[s];
    {
        std::vector<uint32_t> fill(col_ptr.begin(), col_ptr.end() - 1);
        for (uint32_t i = 0; i < t.n_obs; ++i)
            for (uint32_t j = t.indptr[i]; j < t.indptr[i + 1]; ++j) col_pos[fill[t.indices[j]]++] = j - base;
        for (uint32_t s = 0; s < n; ++s) {
            double acc = 0.0;
            for (uint32_t q = col_ptr[s]; q < col_ptr[s + 1]; ++q) {
                const double v = t.data[base + col_pos[q]];
                // without replacement the items are whole counts (src/biom_subsampled.cpp:186-190)
                acc += with_replacement ? std::max(v, 0.0) : std::floor(std::max(v, 0.0));
                col_cum[q] = acc;
            }
        }
    }
    if (!with_replacement) {
        R.hsize = 1;
        while (R.hsize < 2 * depth) R.hsize <<= 1;
    }
    // one plan per method over the kept table, all on the batch's stream (no
    // methods: the batch only draws the tables)
    PostorderNodes nodes;
    if (!specs.empty()) {
        std::unordered_set<std::string> keep(t.obs_ids.begin(), t.obs_ids.end());
        PhyloTree sheared = tree.shear(keep).collapse();
        nodes = postorder_nodes(sheared, t.obs_index);
    }
    if (hipStreamCreateWithFlags(&R.st, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&R.e0) != hipSuccess || hipEventCreate(&R.e1) != hipSuccess) {
        *err = "stream/event creation failed";
        return nullptr;
    }
    // lanes: SSU_RAREFY_LANES (default 4), at most one per replicate
    R.lanes = 4;
    if (const char* env = getenv("SSU_RAREFY_LANES")) R.lanes = (uint32_t)std::max(1L, atol(env));
    R.lanes = std::max<uint32_t>(1, std::min(R.lanes, reps));
    if (specs.empty()) R.lanes = 1;
    if (hipEventCreateWithFlags(&R.fork_ev, hipEventDisableTiming) != hipSuccess) {
        *err = "event creation failed";
        return nullptr;
    }
    for (uint32_t l = 0; l < R.lanes; ++l) {
        hipStream_t q = nullptr;
        hipEvent_t e = nullptr;
        if (hipStreamCreateWithFlags(&q, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) {
            if (q) (void)hipStreamDestroy(q);
            *err = "stream/event creation failed";
            return nullptr;
        }
        R.lane_st.push_back(q);
        R.lane_ev.push_back(e);
        for (const auto& sp : specs) {
            if (sp.bypass_tips) { *err = "bypass_tips is not supported with batched rarefaction"; return nullptr; }
            auto plan = Plan::create(t, nodes, sp, err);
            if (!plan) return nullptr;
            plan->set_stream(q);
            R.plans.push_back(std::move(plan));
        }
    }
    auto dmalloc = [&](void** p, size_t bytes) {
        if (hipMalloc(p, std::max<size_t>(bytes, 16)) != hipSuccess) {
            *err = "hipMalloc(" + std::to_string(bytes) + ") failed";
            return false;
        }
        return true;
    };
    // every replicate of a sample sums to depth, or, without replacement, to all
    // of its whole items when those are fewer (fractional counts)
    std::vector<double> depth_totals(n, (double)depth);
    if (!with_replacement)
        for (uint32_t s = 0; s < n; ++s)
            if (col_ptr[s + 1] > col_ptr[s]) depth_totals[s] = std::min((double)depth, col_cum[col_ptr[s + 1] - 1]);
    if (!dmalloc(&R.d_col_ptr, col_ptr.size() * 4) || !dmalloc(&R.d_col_pos, col_pos.size() * 4) ||
        !dmalloc(&R.d_col_cum, col_cum.size() * 8) || !dmalloc(&R.d_rdata, R.nnz * 8 * reps) ||
        !dmalloc(&R.d_depth, n * 8))
        return nullptr;
    if (!with_replacement && !dmalloc(&R.d_set, (size_t)n * reps * R.hsize * 8)) return nullptr;
    if (!specs.empty() && (!dmalloc((void**)&R.d_flags, specs.size() * reps * ((size_t)n + 1) * sizeof(uint32_t)) ||
                           !dmalloc((void**)&R.d_sums, specs.size() * reps * (size_t)n * sizeof(double))))
        return nullptr;
    for (size_t k = 0; k < specs.size(); ++k) {
        void* p = nullptr;
        if (!dmalloc(&p, R.out_bytes(k) * reps)) return nullptr;
        R.d_out.push_back(p);
    }
    if (hipMemcpy(R.d_col_ptr, col_ptr.data(), col_ptr.size() * 4, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(R.d_col_pos, col_pos.data(), col_pos.size() * 4, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(R.d_col_cum, col_cum.data(), col_cum.size() * 8, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(R.d_depth, depth_totals.data(), n * 8, hipMemcpyHostToDevice) != hipSuccess) {
        *err = "upload failed";
        return nullptr;
    }
    return b;
}

bool RarefyBatch::execute(std::string* err) {
    RarefyImpl& R = *p_;
    if (!R.captured) {
        // eager pass: allocates every buffer the plans need, then the same work is captured
        R.refined = 0;
        RF_HIP(hipEventRecord(R.e0, R.st));
        if (!R.enqueue(false, err)) return false;
        RF_HIP(hipEventRecord(R.e1, R.st));
        RF_HIP(hipStreamSynchronize(R.st));
        RF_HIP(hipEventElapsedTime(&R.ms, R.e0, R.e1));
        RF_HIP(hipStreamBeginCapture(R.st, hipStreamCaptureModeThreadLocal));
        const bool ok = R.enqueue(true, err);
        hipGraph_t g = nullptr;
        const hipError_t ec = hipStreamEndCapture(R.st, &g);
        if (!ok) {
            if (g) (void)hipGraphDestroy(g);
            return false;
        }
        RF_HIP(ec);
        R.graph = g;
        RF_HIP(hipGraphInstantiate(&R.exec, R.graph, nullptr, nullptr, 0));
        R.captured = true;
        R.used_graph = false;
        return true;
    }
    RF_HIP(hipEventRecord(R.e0, R.st));
    RF_HIP(hipGraphLaunch(R.exec, R.st));
    RF_HIP(hipEventRecord(R.e1, R.st));
    RF_HIP(hipStreamSynchronize(R.st));
    RF_HIP(hipEventElapsedTime(&R.ms, R.e0, R.e1));
    R.used_graph = true;
    return R.refine(err);
}

struct RarefyGroupImpl {
    std::vector<RarefyBatch*> batches;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    float ms = 0.0f;
    ~RarefyGroupImpl() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        if (st) (void)hipStreamDestroy(st);
    }
};

RarefyGroup::RarefyGroup() : p_(new RarefyGroupImpl) {}
RarefyGroup::~RarefyGroup() = default;

std::unique_ptr<RarefyGroup> RarefyGroup::create(const std::vector<RarefyBatch*>& batches, std::string* err) {
    std::unique_ptr<RarefyGroup> g(new RarefyGroup);
    RarefyGroupImpl& G = *g->p_;
    auto fail = [&](const char* what) {
        *err = std::string(what) + " failed";
        return nullptr;
    };
    if (batches.empty()) { *err = "a rarefaction group needs at least one batch"; return nullptr; }
    for (RarefyBatch* b : batches) {
        if (!b) { *err = "a rarefaction group holds a null batch"; return nullptr; }
        if (!b->p_->captured && !b->execute(err)) return nullptr;   // the eager pass, then the capture
        if (!b->p_->graph) { *err = "batch without a captured graph"; return nullptr; }
    }
    G.batches = batches;
    if (hipStreamCreateWithFlags(&G.st, hipStreamNonBlocking) != hipSuccess) return fail("stream creation");
    if (hipEventCreate(&G.e0) != hipSuccess || hipEventCreate(&G.e1) != hipSuccess) return fail("event creation");
    if (hipGraphCreate(&G.graph, 0) != hipSuccess) return fail("hipGraphCreate");
    for (RarefyBatch* b : batches) {
        hipGraphNode_t node = nullptr;
        if (hipGraphAddChildGraphNode(&node, G.graph, nullptr, 0, b->p_->graph) != hipSuccess)
            return fail("hipGraphAddChildGraphNode");
    }
    if (hipGraphInstantiate(&G.exec, G.graph, nullptr, nullptr, 0) != hipSuccess) return fail("hipGraphInstantiate");
    return g;
}

bool RarefyGroup::execute(std::string* err) {
    RarefyGroupImpl& G = *p_;
    RF_HIP(hipEventRecord(G.e0, G.st));
    RF_HIP(hipGraphLaunch(G.exec, G.st));
    RF_HIP(hipEventRecord(G.e1, G.st));
    RF_HIP(hipStreamSynchronize(G.st));
    RF_HIP(hipEventElapsedTime(&G.ms, G.e0, G.e1));
    for (RarefyBatch* b : G.batches) {
        RarefyImpl& R = *b->p_;
        R.used_graph = true;
        R.ms = 0.0f;   // (timed for the group as a whole)
        if (!R.refine(err)) return false;
    }
    return true;
}

float RarefyGroup::last_ms() const { return p_->ms; }

uint32_t RarefyBatch::n_samples() const { return p_->table.n_samples; }
uint32_t RarefyBatch::n_obs() const { return p_->table.n_obs; }
uint64_t RarefyBatch::nnz() const { return p_->nnz; }
uint32_t RarefyBatch::reps() const { return p_->reps; }
const Table& RarefyBatch::kept() const { return p_->table; }
const std::vector<uint32_t>& RarefyBatch::sample_index() const { return p_->sample_index; }
const std::vector<uint32_t>& RarefyBatch::obs_index() const { return p_->obs_index; }
bool RarefyBatch::out_fp64(size_t k) const { return k < p_->specs.size() && p_->specs[k].fp64; }
float RarefyBatch::last_ms() const { return p_->ms; }
bool RarefyBatch::graph_used() const { return p_->used_graph; }
uint64_t RarefyBatch::refined_pairs() const { return p_->refined; }

bool RarefyBatch::copy_matrix(size_t k, uint32_t rep, void* host, std::string* err) const {
    const RarefyImpl& R = *p_;
    if (k >= R.specs.size() || rep >= R.reps) { if (err) *err = "method or replicate out of range"; return false; }
    return copy_to_host(host, (const char*)R.d_out[k] + R.out_bytes(k) * rep, R.out_bytes(k), R.st, err);
}

bool RarefyBatch::copy_counts(uint32_t rep, double* host, std::string* err) const {
    const RarefyImpl& R = *p_;
    if (rep >= R.reps) { if (err) *err = "replicate out of range"; return false; }
    RF_HIP(hipMemcpy(host, (const double*)R.d_rdata + (size_t)rep * R.nnz, R.nnz * 8, hipMemcpyDeviceToHost));
    return true;
}

uint64_t RarefyBatch::pair_nodes() const {
    const uint64_t n = p_->table.n_samples;
    uint64_t sum = 0;
    for (size_t k = 0; k < p_->specs.size(); ++k) sum += n * (n - 1) / 2 * p_->plans[k]->processed_nodes();
    return sum * p_->reps;
}

}  // namespace ssu
