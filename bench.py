#!/usr/bin/env python3
"""Throughput of the Strided State UniFrac hot path on one MI355X.

Workload (default): random binary tree with 100,000 tips (Exp(0.1) lengths,
FP32-rounded), 16,384 samples at 1% tip density with Zipf(1.1) tip popularity,
lognormal counts -- the "16k samples" configuration.  One step computes the
full 16,384 x 16,384 unweighted matrix (exact int8-MFMA Gram over the nodes
held by at least 6% of the samples, the rarer ones from posting lists) and the full
weighted_normalized matrix (nodes held by at least 25% of the samples through
32x64 pair tiles summing l |u - v| in FP32, merged into FP64, visiting per
tile only the nodes either sample panel holds; rarer nodes through per-node
posting lists walked row by row into exact fixed-point sums of l min(u, v)).

Timed region of `value` (host to host): each step calls
one_off_matrix_inmem_fp32_v3 once per method, from the CSR table and the tree
arrays in host memory to the n x n matrix in host memory (the result buffer
is passed back in, as src/api.cpp:674-677 allows, so its pages are mapped);
it includes tree shear/collapse, uploads, plan building, the device work,
near-duplicate refinement and the copy-back.  `kernel_only` repeats the step
on device-resident plans (inputs in HBM, matrices left in HBM) with HIP-event
timings per phase.

metric: pair-node updates/sec = sum over both methods of P * N / t, with
P = n(n-1)/2 and N the nodes the method processes after shear/collapse.

Prints one JSON line (rank 0, N=1).  Extra fields: the per-call phase
breakdown, per-kernel HIP-event timings, the max relative error of both
matrices against the FP64 oracle on a 48-sample subset, the plain dense
pair-tile kernels (unweighted and weighted) for comparison, a roofline entry
for the dominant kernel and the CPU baseline.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "unifrac-binaries_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

PEAK_FP32_TFLOPS = 157.3    # MI355X published FP32 vector peak (packed FP32)
PEAK_FP64_TFLOPS = 78.6     # MI355X published FP64 vector peak
FLOPS_PER_TERM_SPARSE = 4   # sub, abs, fused multiply-add (the pair-node count of the stripe-kernel model)
PEAK_INT8_TOPS = 5000.0     # MI355X published dense INT8 MFMA peak, TOP/s
PEAK_HBM_GBS = 8000.0       # MI355X published HBM3E bandwidth, GB/s
GRAM_TILE = 128             # Gram output tile edge


def bytes_per_posting(name, n):
    """Bytes a walked posting reads: its sample index (16-bit up to 65,536
    samples, else 32-bit), its value (FP32, or FP64 for "_fp64" names; none for
    the Gram path's walk) and, for VAW, its count in the value's type."""
    vaw = ":vaw" in name
    val = 8 if "_fp64" in name.split(":")[0] else 4
    gram_walk = name.startswith("unweighted") and not vaw
    return (2 if n <= 65536 else 4) + (0 if gram_walk else val) + (val if vaw else 0)


def gram_units(n, kernel_terms):
    """(sample pair i <= j, Gram column) units one k_gram_i8r launch processes:
    kernel_terms counts the padded upper-triangle tiles times the columns."""
    t = (n + GRAM_TILE - 1) // GRAM_TILE
    cols = kernel_terms / (t * (t + 1) // 2 * GRAM_TILE * GRAM_TILE)
    return n * (n + 1) / 2 * cols
FLOPS_PER_PAIR_NODE_W = 4   # sub, abs, fused multiply-add
OPS_PER_PAIR_NODE_GRAM = 8  # 4 int8 digit passes x one multiply-add


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def sub_table(table, idx):
    """CSR restricted to the samples idx (re-indexed), empty observations dropped."""
    obs_ids, sample_ids, indptr, indices, data = table
    remap = -np.ones(len(sample_ids), np.int64)
    remap[idx] = np.arange(len(idx))
    keep = remap[indices] >= 0
    rows = np.repeat(np.arange(len(obs_ids)), np.diff(indptr.astype(np.int64)))
    r, c, v = rows[keep], remap[indices[keep]], data[keep]
    order = np.lexsort((c, r))
    r, c, v = r[order], c[order], v[order]
    used, counts = np.unique(r, return_counts=True)
    ip = np.zeros(used.size + 1, np.uint32)
    np.cumsum(counts, out=ip[1:])
    return ([obs_ids[i] for i in used], [sample_ids[i] for i in idx], ip, c.astype(np.uint32), v)


def max_rel_err(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    off = ~np.eye(a.shape[0], dtype=bool)
    d = np.abs(a - b)[off]
    s = np.abs(b)[off]
    return float(np.max(np.where(s > 0, d / np.where(s > 0, s, 1), d)))


def parse_methods(text):
    """'weighted_normalized,generalized:alpha=0.5,weighted_normalized:vaw' -> {label: (method, kwargs)}"""
    out = {}
    for item in text.split(","):
        parts = item.strip().split(":")
        kw = {}
        for opt in parts[1:]:
            if opt == "vaw":
                kw["variance_adjust"] = True
            elif opt.startswith("alpha="):
                kw["alpha"] = float(opt[6:])
            else:
                raise SystemExit(f"bench.py: unknown method option {opt!r}")
        out[item.strip()] = (parts[0], kw)
    return out


def cpu_baseline(tree, table, methods, budget_s):
    """Reference CPU code (oracle/_ref) on stripe sub-ranges; extrapolated to the full matrix."""
    import oracle as O
    kind = "reference" if O.ref_available() else "port"
    otree = O.Tree(tree[0].astype(np.uint8), tree[1], tree[2])
    otab = O.Table(*table)
    n = otab.n_samples
    S = (n + 1) // 2
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or len(os.sched_getaffinity(0))
    total_time, per = 0.0, {}
    for label, (method, kw) in methods.items():
        base, fp32 = method.replace("_fp64", ""), not method.endswith("_fp64")

        def run(k):
            t0 = time.perf_counter()
            if kind == "reference":
                O.ref_stripes(otree, otab, base, fp32=fp32, start=0, stop=k, **kw)
            else:
                O.stripes(otree, otab, base, start=0, stop=k, threads=threads, **kw)
            return time.perf_counter() - t0
        t1 = run(1)
        if t1 * 4 < budget_s / len(methods):
            k2 = 3
            t2 = run(k2)
            slope = max((t2 - t1) / (k2 - 1), 1e-9)
            intercept = max(t1 - slope, 0.0)
        else:
            slope, intercept = t1, 0.0
        full = intercept + S * slope
        per[label] = {"stripe_s": slope, "setup_s": intercept, "full_matrix_s": full}
        total_time += full
        log(f"cpu baseline {label}: {slope:.3f} s/stripe, setup {intercept:.2f} s, full {full:.0f} s")
    per["_host"] = {"isa": O.REF_ISA if kind == "reference" else "gcc -O3 (port)",
                    "threads_from": "OMP_NUM_THREADS" if os.environ.get("OMP_NUM_THREADS") else "sched_getaffinity",
                    "affinity_cpus": len(os.sched_getaffinity(0)), "os_cpu_count": os.cpu_count()}
    return kind, threads, total_time, per


class OneOff:
    """one_off_matrix_inmem_fp32_v3 / _v3 with the result buffer handed back in
    on every later call (src/api.cpp:674-677 accepts a caller-preallocated
    result), so the timed calls write into mapped pages."""

    def __init__(self, ssu, inp, method, kw):
        import ctypes as C
        self.C, self.ssu, self.inp, self.method, self.kw = C, ssu, inp, method, kw
        self.fp64 = method.endswith("_fp64")
        self.res = C.POINTER(ssu.MatFullFP64 if self.fp64 else ssu.MatFullFP32)()

    def __call__(self):
        C, L, kw = self.C, self.ssu.lib(), self.kw
        fn = L.one_off_matrix_inmem_v3 if self.fp64 else L.one_off_matrix_inmem_fp32_v3
        st = fn(C.byref(self.inp.table), C.byref(self.inp.tree), self.method.encode(),
                C.c_bool(kw.get("variance_adjust", False)), C.c_double(kw.get("alpha", 1.0)), C.c_bool(False),
                C.c_bool(True), C.c_uint(1), C.c_uint(0), C.c_bool(True), None, C.byref(self.res))
        if st != self.ssu.OKAY:
            raise RuntimeError(f"{self.method}: status {st}: {self.ssu.last_error()}")
        return self.ssu.last_call_timings()

    def matrix(self):
        n = self.res.contents.n_samples
        return np.ctypeslib.as_array(self.res.contents.matrix, shape=(n * n,)).reshape(n, n)

    def close(self):
        if self.res:
            L = self.ssu.lib()
            (L.destroy_mat_full_fp64 if self.fp64 else L.destroy_mat_full_fp32)(self.C.byref(self.res))


def copy_probe(nbytes):
    """hipMemcpy device-to-device of nbytes: the runtime's copy kernel reads and writes nbytes each."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(C.byref(b), C.c_size_t(nbytes)) == 0
    hip.hipMemset(a, 1, C.c_size_t(nbytes))
    hip.hipDeviceSynchronize()
    for _ in range(2):
        hip.hipMemcpy(b, a, C.c_size_t(nbytes), 3)  # hipMemcpyDeviceToDevice
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)
    log(f"copy probe: {nbytes} bytes device-to-device, twice")


def bench_rarefy(args):
    """Batched rarefaction: every table's replicates x methods captured in one HIP graph per table; a
    step is one launch of a parent graph holding every table's graph (the tables run side by side)."""
    import ssu
    import synth
    import oracle as O
    c = synth.RAREFY
    t0 = time.perf_counter()
    tree, tables = synth.rarefy_workload(tips=args.tips)
    log(f"workload rarefy: {len(tree[0]) // 2} nodes, tables {[len(t[1]) for t in tables]} "
        f"({time.perf_counter() - t0:.1f} s to generate)")
    specs = parse_methods(args.methods)
    methods = list(specs)
    reps = args.reps or c["reps"]
    batches = []
    for k, t in enumerate(tables):
        inp = ssu.Inputs(*tree, *t)
        b = ssu.Rarefaction(inp, [specs[m][0] for m in methods], c["depth"], reps, with_replacement=True,
                            seed=1000 + k, **(specs[methods[0]][1]))
        batches.append((inp, b, t))
        log(f"table {k}: {len(t[1])} samples -> {b.n} kept, pair-nodes {b.pair_nodes():.4g}")
    for _ in range(max(args.warmup, 1)):        # the first execute runs eagerly and captures the graph
        for _, b, _ in batches:
            b.execute()
    log("warmup done (graphs captured)")
    # each table's graph on its own, one after another (per-table device times)
    dev_ms = [[] for _ in batches]
    t_seq = time.perf_counter()
    for k, (_, b, _) in enumerate(batches):
        b.execute()
        dev_ms[k].append(b.last_ms())
    seq_ms = 1000.0 * (time.perf_counter() - t_seq)
    log(f"tables one after another: {seq_ms:.1f} ms, device {[round(v[0], 1) for v in dev_ms]}")
    # the timed steps: one launch of the parent graph of every table's graph
    group = ssu.RarefactionGroup([b for _, b, _ in batches])
    group.execute()
    group_ms = []
    t_start = time.perf_counter()
    for step in range(args.steps):
        group.execute()
        group_ms.append(group.last_ms())
        log(f"step {step + 1}/{args.steps} done")
    elapsed = time.perf_counter() - t_start
    assert all(b.graph_used() for _, b, _ in batches)
    pair_nodes = sum(b.pair_nodes() for _, b, _ in batches)
    value = pair_nodes / (elapsed / args.steps)
    # accuracy: replicate 0 of the smallest table, 48 samples, against the FP64 oracle on its counts
    _, b0, t0_ = batches[0]
    si, oi, ip, ix = b0.kept()
    rt = ([t0_[0][i] for i in oi], [t0_[1][s] for s in si], ip, ix, b0.counts(0))
    rng = np.random.default_rng(7)
    idx = np.sort(rng.choice(b0.n, size=min(48, b0.n), replace=False))
    sub = sub_table(rt, idx)
    otree = O.Tree(tree[0].astype(np.uint8), tree[1], tree[2])
    errors = {}
    for k, m in enumerate(methods):
        ref = O.matrix(otree, O.Table(*sub), specs[m][0], **specs[m][1])
        errors[m] = max_rel_err(b0.matrix(k, 0)[np.ix_(idx, idx)], ref)
    log(f"max rel err vs FP64 oracle (table 0, replicate 0): {errors}")
    cpu = None
    if not args.no_cpu_baseline:
        # the reference's per-replicate work: one replicate of the 1,536-sample table
        _, b1, t1_ = batches[1]
        si, oi, ip, ix = b1.kept()
        rt1 = ([t1_[0][i] for i in oi], [t1_[1][s] for s in si], ip, ix, b1.counts(0))
        kind, threads, t_full, per = cpu_baseline(tree, rt1, specs, args.cpu_budget)
        cpu = {"value": (b1.pair_nodes() / reps) / t_full, "unit": "pair-node updates/s", "cores": threads,
               "kind": kind,
               "sample": f"one replicate of the {b1.n}-sample table: stripes [0,1) and [0,3) of each matrix "
                         f"(FP32 reference kernels, OpenMP), extrapolated to all {(b1.n + 1) // 2} stripes; "
                         f"host subsampling not timed",
               "per_method": {k: v for k, v in per.items() if k != "_host"}, **per["_host"]}
    out = {
        "metric": "pair-node updates/sec, batched rarefaction (4 tables x 32 replicates, depth 1000)",
        "value": value,
        "unit": "pair-node updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1000.0 * elapsed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int8 MFMA (unweighted, exact) + fp32 terms summed in fp64 and int64 fixed point (weighted)",
        "data": "synthetic",
        "config": {"workload": f"rarefy_20k_4tables: random binary tree {len(tree[0]) // 4 + 1} tips, tables of "
                               f"{', '.join(str(len(t[1])) for t in tables)} samples at 2% density, "
                               f"{reps} replicates at depth {c['depth']} with replacement, {' + '.join(methods)}",
                   "kept_samples": [b.n for _, b, _ in batches], "replicates": reps, "depth": c["depth"],
                   "methods": methods},
        "max_rel_err_vs_fp64": errors,
        "device_ms_per_step": float(np.mean(group_ms)),
        "device_ms_per_table": [float(np.mean(v)) for v in dev_ms],
        "ms_tables_one_after_another": seq_ms,
        "pair_nodes_per_step": pair_nodes,
        "roofline": None,
        "roofline_note": "one HIP graph launch per step: a parent graph with every table's graph (sampler + "
                         "every replicate's matrices) as a child node; device_ms_per_table times each table's "
                         "graph alone; per-kernel durations come from the rocprofv3 kernel trace",
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="unw_100k_16384",
                    help="unw_100k_16384, cpu_1k_256, w_50k_8192, large_200k_65536 or rarefy_20k_4tables")
    ap.add_argument("--reps", type=int, default=None, help="rarefaction replicates (rarefy_20k_4tables)")
    ap.add_argument("--methods", default="unweighted,weighted_normalized",
                    help="comma list of method[:alpha=A][:vaw]; one step runs every listed matrix")
    ap.add_argument("--samples", type=int, default=None)
    ap.add_argument("--tips", type=int, default=None)
    ap.add_argument("--cpu-budget", type=float, default=20.0, help="seconds of reference CPU work")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-plain", action="store_true", help="skip the plain dense pair-tile kernels")
    ap.add_argument("--sequential", action="store_true",
                    help="keep one plan in device memory at a time (for workloads whose plans do not fit together)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="also run a 1 GiB device-to-device copy (reference point for HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); "
              "multi-GPU sharding is a design note, not an implemented path", file=sys.stderr)
        sys.exit(2)

    import ssu
    import synth

    if args.workload == "rarefy_20k_4tables":
        if not ssu.device_available():
            print("bench.py: no HIP device visible", file=sys.stderr)
            sys.exit(1)
        bench_rarefy(args)
        return
    if args.copy_probe:
        copy_probe(1 << 30)
    if not ssu.device_available():
        print("bench.py: no HIP device visible", file=sys.stderr)
        sys.exit(1)
    t0 = time.perf_counter()
    tree, table = synth.workload(args.workload, samples=args.samples, tips=args.tips)
    n = len(table[1])
    log(f"workload {args.workload}: {len(tree[0]) // 2} nodes, {n} samples, nnz {table[4].size} "
        f"({time.perf_counter() - t0:.1f} s to generate)")
    inp = ssu.Inputs(*tree, *table)
    specs = parse_methods(args.methods)
    import oracle as O
    rng = np.random.default_rng(7)
    idx = np.sort(rng.choice(n, size=min(48, n), replace=False))
    sub = sub_table(table, idx)
    otree = O.Tree(tree[0].astype(np.uint8), tree[1], tree[2])
    errors = {}

    times = {k: [] for k in specs}
    plans_info = {}
    thetas = {}
    k_steps = max(1, min(args.steps, 10))
    # kernel-only: device-resident plans, inputs in HBM, matrices left in HBM
    if not args.sequential:
        # every matrix of a step computed back to back, all plans resident in HBM
        plans = {label: ssu.Plan(inp, m, **kw) for label, (m, kw) in specs.items()}
        for name, p in plans.items():
            log(f"plan {name}: main kernel {p.main_kernel()}, pair-nodes {p.pair_nodes():.4g}")
        for _ in range(max(args.warmup, 1)):
            for p in plans.values():
                p.execute()
        t_start = time.perf_counter()
        for step in range(k_steps):
            for name, p in plans.items():
                p.execute()
                times[name].append(p.timings())
        k_elapsed = time.perf_counter() - t_start
        for name, p in plans.items():
            plans_info[name] = (p.main_kernel(), p.pair_nodes(), p.kernel_terms(), p.sparse_terms())
            thetas[name] = p.split_theta()
            p.close()
    else:
        # one plan resident at a time (workloads whose plans do not fit in HBM together)
        k_elapsed = 0.0
        for name, (m, kw) in specs.items():
            p = ssu.Plan(inp, m, **kw)
            log(f"plan {name}: main kernel {p.main_kernel()}, pair-nodes {p.pair_nodes():.4g}")
            for _ in range(max(args.warmup, 1)):
                p.execute()
            t_start = time.perf_counter()
            for step in range(k_steps):
                p.execute()
                times[name].append(p.timings())
            k_elapsed += time.perf_counter() - t_start
            plans_info[name] = (p.main_kernel(), p.pair_nodes(), p.kernel_terms(), p.sparse_terms())
            thetas[name] = p.split_theta()
            p.close()
    log(f"kernel-only: {1000.0 * k_elapsed / k_steps:.1f} ms per step; occupancy splits chosen {thetas}")

    # host to host: one C-API call per method and step, CSR and tree in host memory -> matrix in host memory
    calls = {label: OneOff(ssu, inp, m, kw) for label, (m, kw) in specs.items()}
    phases = {k: [] for k in specs}
    for _ in range(args.warmup):
        for c in calls.values():
            c()
    log("warmup done")
    t_start = time.perf_counter()
    for step in range(args.steps):
        for name, c in calls.items():
            phases[name].append(c())
        log(f"step {step + 1}/{args.steps} done")
    elapsed = time.perf_counter() - t_start
    for name, c in calls.items():
        m = c.matrix()
        ref = O.matrix(otree, O.Table(*sub), specs[name][0].replace("_fp64", ""), **specs[name][1])
        errors[name] = max_rel_err(m[np.ix_(idx, idx)], ref)
        c.close()
    log(f"max rel err vs FP64 oracle: {errors}")
    ms_per_step = 1000.0 * elapsed / args.steps
    pair_nodes = {k: v[1] for k, v in plans_info.items()}
    value = sum(pair_nodes.values()) / (elapsed / args.steps)
    call_phases = {name: {k: float(np.mean([ph[k] for ph in v])) for k in v[0]} for name, v in phases.items()}
    kernel_only = {"value": sum(pair_nodes.values()) / (k_elapsed / k_steps),
                   "ms_per_step": 1000.0 * k_elapsed / k_steps, "steps": k_steps}

    def avg(name, key):
        return float(np.mean([t[key] for t in times[name]]))

    kernels = {}
    for name, (kern, pn, terms, sterms) in plans_info.items():
        kernels[name] = {"kernel": kern, "main_ms": avg(name, "main"), "embed_ms": avg(name, "embed"),
                         "prep_ms": avg(name, "prep"), "combine_ms": avg(name, "combine"),
                         "sparse_ms": avg(name, "sparse"), "device_total_ms": avg(name, "total"),
                         "pair_node_per_s": pn / (avg(name, "total") / 1000.0),
                         "kernel_terms": terms, "sparse_terms": sterms}
        kernels[name]["terms_per_pair_node"] = (terms + sterms) / max(pair_nodes[name], 1)
        kernels[name]["split_theta"] = thetas.get(name, 0.0)
    # dominant kernel: the largest of the main kernels and the posting-list phases (k_pairs_rows)
    cands = [(k["main_ms"], name, k["kernel"]) for name, k in kernels.items()]
    cands += [(k["sparse_ms"], name, "k_pairs_rows" + ("[gram]" if name.startswith("unweighted") else ""))
              for name, k in kernels.items() if k["sparse_terms"]]
    dom_ms, dom, dom_kernel = max(cands)
    dk = kernels[dom]
    if dom_kernel == "k_gram_i8r":
        achieved = OPS_PER_PAIR_NODE_GRAM * gram_units(n, dk["kernel_terms"]) / (dom_ms / 1000.0) / 1e12
        roofline = {"kernel": dom_kernel, "bound": "mfma", "achieved": achieved, "peak": PEAK_INT8_TOPS,
                    "unit": "TOP/s", "frac": achieved / PEAK_INT8_TOPS, "traffic": None}
    elif dom_kernel.startswith("k_pairs_rows"):
        # each walked posting reads its sample index (and value)
        achieved = bytes_per_posting(dom, n) * dk["sparse_terms"] / (dom_ms / 1000.0) / 1e9
        roofline = {"kernel": dom_kernel, "bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS,
                    "unit": "GB/s", "frac": achieved / PEAK_HBM_GBS, "traffic": None}
    elif dom_kernel in ("k_pairs_sparse", "k_pairs_sparse32"):
        # executed work: the (pair, shared node) terms the kernel evaluated, 4 flops each (sub, abs, FMA)
        peak = PEAK_FP64_TFLOPS if dom_kernel == "k_pairs_sparse" else PEAK_FP32_TFLOPS
        achieved = FLOPS_PER_TERM_SPARSE * dk["kernel_terms"] / (dom_ms / 1000.0) / 1e12
        roofline = {"kernel": dom_kernel, "bound": "valu", "achieved": achieved, "peak": peak,
                    "unit": "TFLOP/s", "frac": achieved / peak, "traffic": None}
    else:
        achieved = FLOPS_PER_PAIR_NODE_W * pair_nodes[dom] / (dom_ms / 1000.0) / 1e12
        roofline = {"kernel": dom_kernel, "bound": "valu", "achieved": achieved, "peak": PEAK_FP32_TFLOPS,
                    "unit": "TFLOP/s", "frac": achieved / PEAK_FP32_TFLOPS, "traffic": None}
    for name, k in kernels.items():
        if k["kernel"] == "k_gram_i8r":
            g_ops = OPS_PER_PAIR_NODE_GRAM * gram_units(n, k["kernel_terms"]) / (k["main_ms"] / 1000.0) / 1e12
            k["mfma_int8_tops"] = g_ops
            k["mfma_frac"] = g_ops / PEAK_INT8_TOPS
    prof = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(prof):
        tr = json.load(open(prof)).get(f"{args.workload}:{dom_kernel}")
        if tr and not args.samples and not args.tips:
            roofline["traffic"] = tr
            roofline["traffic_source"] = ("profiles/traffic.json: rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE "
                                          "of this kernel on this workload, scaled by a 1 GiB copy's counts")

    plain = None
    if not args.no_plain:
        plain = {}
        for name in [x for x in specs if x in ("unweighted", "weighted_normalized")]:
            pp = ssu.Plan(inp, name, plain_kernel=1)
            pp.execute()
            pp.execute()
            pt = pp.timings()
            plain[name] = {"kernel": pp.main_kernel(), "main_ms": pt["main"], "device_total_ms": pt["total"],
                           "speedup_main": pt["main"] / kernels[name]["main_ms"],
                           "speedup_total": pt["total"] / kernels[name]["device_total_ms"]}
            pp.close()
            log(f"plain dense kernel {name}: {pt['main']:.1f} ms main "
                f"({plain[name]['speedup_main']:.1f}x slower than {kernels[name]['kernel']})")

    cpu = None
    if not args.no_cpu_baseline:
        kind, threads, t_full, per = cpu_baseline(tree, table, specs, args.cpu_budget)
        cpu = {"value": sum(pair_nodes.values()) / t_full, "unit": "pair-node updates/s", "cores": threads,
               "kind": kind,
               "sample": f"stripes [0,1) and [0,3) of each matrix (the reference kernels in each method's "
                         f"precision, OpenMP), extrapolated linearly to all {(n + 1) // 2} stripes",
               "per_method": {k: v for k, v in per.items() if k != "_host"}, **per["_host"]}

    out = {
        "metric": "pair-node updates/sec, 16k samples (unweighted, weighted); max rel err vs FP64",
        "value": value,
        "unit": "pair-node updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int8 MFMA (unweighted, exact) + fp32 terms summed in fp64 and int64 fixed point (weighted)",
        "data": "synthetic",
        "config": {"workload": f"{args.workload}: random binary tree {len(tree[0]) // 4 + 1} tips, "
                               f"{n} samples, {' + '.join(specs)}, full n x n matrices",
                   "samples": n, "tree_nodes": len(tree[0]) // 2, "nnz": int(table[4].size),
                   "methods": list(specs)},
        "max_rel_err_vs_fp64": errors,
        "call_phases_ms": call_phases,
        "kernel_only": kernel_only,
        "kernels": kernels,
        "plain_dense_kernels": plain,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
