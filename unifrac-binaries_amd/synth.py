"""Seeded synthetic workloads for benchmarks and parity tests.

Tree: random binary topology built by repeatedly joining two uniformly chosen
active nodes; branch lengths Exp(mean 0.1) rounded to FP32 (so that a newick
round trip through a float parse reproduces them exactly); root length 0;
tips named t<i>.

Table: each sample draws its nonzero tips without replacement, uniformly or
with Zipf(s) tip popularity over a random rank permutation (Efraimidis-
Spirakis keys, implemented as Gumbel top-k); counts lognormal or negative
binomial.  Only observed tips become table rows.  Tables with more than 2^32
key draws (65,536 samples x 200,000 tips) draw them from one generator per
128-sample batch, spawned from the seed, so that they generate in parallel.
"""
from __future__ import annotations

import os

import numpy as np


def random_tree(n_tips: int, seed: int, mean_len: float = 0.1):
    """Balanced-parentheses arrays (structure bool, lengths float64, names list)."""
    rng = np.random.default_rng(seed)
    n_nodes = 2 * n_tips - 1
    left = np.full(n_nodes, -1, np.int64)
    right = np.full(n_nodes, -1, np.int64)
    active = list(range(n_tips))
    picks = rng.random((n_tips - 1, 2))
    for step in range(n_tips - 1):
        na = len(active)
        i = int(picks[step, 0] * na)
        a = active[i]
        active[i] = active[-1]
        active.pop()
        na -= 1
        j = int(picks[step, 1] * na)
        b = active[j]
        active[j] = active[-1]
        active.pop()
        node = n_tips + step
        left[node], right[node] = a, b
        active.append(node)
    root = active[0]
    node_len = rng.exponential(mean_len, n_nodes).astype(np.float32).astype(np.float64)
    node_len[root] = 0.0
    # iterative DFS -> parentheses
    structure = np.zeros(2 * n_nodes, dtype=bool)
    lengths = np.zeros(2 * n_nodes)
    open_node = np.full(2 * n_nodes, -1, np.int64)
    pos = 0
    stack = [(root, 0)]
    while stack:
        node, state = stack.pop()
        if state == 0:
            structure[pos] = True
            lengths[pos] = node_len[node]
            open_node[pos] = node
            pos += 1
            stack.append((node, 1))
            if left[node] >= 0:
                stack.append((right[node], 0))
                stack.append((left[node], 0))
        else:
            pos += 1  # closing parenthesis, structure stays False
    names = [""] * (2 * n_nodes)
    tip_pos = np.nonzero((open_node >= 0) & (open_node < n_tips))[0]
    for p in tip_pos:
        names[p] = f"t{open_node[p]}"
    return structure, lengths, names


def caterpillar_tree(n_tips: int, seed: int, mean_len: float = 0.1, shuffle: bool = True):
    """A caterpillar (maximally deep) binary tree: every internal node joins one
    tip and the subtree of the remaining tips, n_tips - 1 levels deep; with
    shuffle the tip sits left or right at random.  Same arrays as random_tree."""
    rng = np.random.default_rng(seed)
    n_nodes = 2 * n_tips - 1
    node_len = rng.exponential(mean_len, n_nodes).astype(np.float32).astype(np.float64)
    tip_left = rng.random(n_tips) < 0.5 if shuffle else np.ones(n_tips, bool)
    structure, lengths, names = [], [], []
    # internal node i (0 = root) has tip i and internal node i + 1 (the last one: tips n-2, n-1)
    closes = []
    ln = 0
    for i in range(n_tips - 1):
        structure.append(True)
        lengths.append(0.0 if i == 0 else node_len[n_tips + i])
        names.append("")
        tip = [(True, node_len[i], f"t{i}"), (False, 0.0, "")]
        if i == n_tips - 2:
            last = [(True, node_len[i + 1], f"t{i + 1}"), (False, 0.0, "")]
            pair = tip + last if tip_left[i] else last + tip
            for st_, le_, nm_ in pair:
                structure.append(st_)
                lengths.append(le_)
                names.append(nm_)
        elif tip_left[i]:
            for st_, le_, nm_ in tip:
                structure.append(st_)
                lengths.append(le_)
                names.append(nm_)
        else:
            closes.append(tip)
            continue
        closes.append(None)
    # close the internal nodes from the deepest up, a right-hand tip before its parent's ')'
    for tip in reversed(closes):
        if tip is not None:
            for st_, le_, nm_ in tip:
                structure.append(st_)
                lengths.append(le_)
                names.append(nm_)
        structure.append(False)
        lengths.append(0.0)
        names.append("")
    return np.array(structure, bool), np.array(lengths), names


def contract_tree(tree, frac: float, seed: int):
    """Delete the parentheses of a random fraction of the internal non-root
    nodes: their children join the parent, which then has three or more."""
    st, le, nm = tree
    st = np.asarray(st, bool)
    n = st.size
    partner = np.zeros(n, np.int64)
    stack = []
    for i in range(n):
        if st[i]:
            stack.append(i)
        else:
            o = stack.pop()
            partner[o], partner[i] = i, o
    internal = [i for i in range(1, n) if st[i] and i + 1 < n and st[i + 1]]
    rng = np.random.default_rng(seed)
    drop = rng.choice(internal, int(frac * len(internal)), replace=False) if internal else []
    keep = np.ones(n, bool)
    for o in drop:
        keep[o] = keep[partner[o]] = False
    return st[keep], np.asarray(le)[keep], [x for x, k in zip(nm, keep) if k]


def bp_to_newick(structure, lengths, names) -> str:
    """Newick text for parentheses arrays, lengths printed with %.9g."""
    out = []
    n = len(structure)
    i = 0
    stack = []
    last_closed = False
    while i < n:
        if structure[i]:
            if last_closed:
                out.append(",")
            if i + 1 < n and not structure[i + 1]:  # tip
                out.append(f"{names[i]}:{lengths[i]:.9g}")
                i += 2
                last_closed = True
                continue
            out.append("(")
            stack.append(i)
            last_closed = False
            i += 1
        else:
            o = stack.pop()
            nm = names[o]
            out.append(")" + nm + (f":{lengths[o]:.9g}" if stack else ""))
            last_closed = True
            i += 1
    return "".join(out) + ";"


# tables needing more key draws than this take them from one generator per
# sample batch (spawned from the seed), in parallel; smaller ones from the
# single seeded generator
PARALLEL_DRAWS = 1 << 32


def random_table(n_tips: int, n_samples: int, density: float, seed: int, zipf_s: float | None = None,
                 counts: str = "lognormal", batch: int = 128):
    """Observation-major CSR: (obs_ids, sample_ids, indptr, indices, data)."""
    rng = np.random.default_rng(seed)
    k = max(1, int(round(density * n_tips)))
    if zipf_s is not None:
        ranks = rng.permutation(n_tips) + 1
        logw = (-zipf_s * np.log(ranks)).astype(np.float32)
    tips_all = np.empty(n_samples * k, np.int64)
    inv_w = None if zipf_s is None else np.exp(-logw)[None, :]

    def top_k(b0, u):
        # the draws come from the one generator in batch order; the keys and the
        # top-k of each batch are independent of the others (numpy releases the GIL)
        b1 = b0 + u.shape[0]
        np.maximum(u, np.float32(1e-30), out=u)
        if inv_w is None:
            keys = u
        else:
            # Efraimidis-Spirakis: top-k of u^(1/w)  <=>  top-k of log(u)/w
            keys = np.log(u)
            keys *= inv_w
        np.negative(keys, out=keys)
        top = np.argpartition(keys, k - 1, axis=1)[:, :k]
        tips_all[b0 * k:b1 * k] = top.reshape(-1)

    from concurrent.futures import ThreadPoolExecutor
    workers = max(1, min(16, len(os.sched_getaffinity(0))))
    starts = list(range(0, n_samples, batch))
    with ThreadPoolExecutor(workers) as pool:
        if n_samples * n_tips <= PARALLEL_DRAWS:
            pending = []
            for b0 in starts:
                b1 = min(n_samples, b0 + batch)
                pending.append(pool.submit(top_k, b0, rng.random((b1 - b0, n_tips), dtype=np.float32)))
                if len(pending) >= 2 * workers:   # bounded memory: a few batches in flight
                    pending.pop(0).result()
            for f in pending:
                f.result()
        else:
            # one generator per batch, spawned from the seed: the draws themselves run in parallel
            subs = np.random.SeedSequence(seed).spawn(len(starts))

            def batch_job(q):
                b0 = starts[q]
                g = np.random.default_rng(subs[q])
                top_k(b0, g.random((min(n_samples, b0 + batch) - b0, n_tips), dtype=np.float32))

            list(pool.map(batch_job, range(len(starts))))
    samples_all = np.repeat(np.arange(n_samples, dtype=np.int64), k)
    if counts == "lognormal":
        vals = np.maximum(1.0, np.ceil(rng.lognormal(3.0, 1.5, tips_all.size)))
    elif counts == "negbin":
        # mean 20, dispersion 0.5 read as var = mu + 0.5 mu^2: r = 2, p = r / (r + mu)
        vals = np.maximum(1.0, rng.negative_binomial(2, 2.0 / 22.0, tips_all.size).astype(np.float64))
    else:
        raise ValueError(counts)
    order = np.lexsort((samples_all, tips_all))
    tips_s, samp_s, vals_s = tips_all[order], samples_all[order], vals[order]
    observed, row_counts = np.unique(tips_s, return_counts=True)
    indptr = np.zeros(observed.size + 1, np.uint32)
    np.cumsum(row_counts, out=indptr[1:])
    obs_ids = [f"t{t}" for t in observed]
    sample_ids = [f"s{i}" for i in range(n_samples)]
    return obs_ids, sample_ids, indptr, samp_s.astype(np.uint32), vals_s


# Workloads (tips, samples, density, zipf, counts, seed)
CONFIGS = {
    "cpu_1k_256": dict(tips=1000, samples=256, density=0.05, zipf=None, counts="lognormal", seed=20261201),
    "unw_100k_16384": dict(tips=100_000, samples=16384, density=0.01, zipf=1.1, counts="lognormal", seed=20261202),
    "w_50k_8192": dict(tips=50_000, samples=8192, density=0.02, zipf=None, counts="negbin", seed=20261203),
    "large_200k_65536": dict(tips=200_000, samples=65536, density=0.005, zipf=1.1, counts="lognormal",
                             seed=20261205),
}


# Batched rarefaction: four tables of uneven sample counts on one 20k-tip tree,
# 2% density, lognormal counts; 32 replicates each at depth 1000 with replacement.
RAREFY = dict(tips=20_000, tables=(512, 1536, 3000, 6000), density=0.02, counts="lognormal", depth=1000, reps=32,
              seed=20261204)


def rarefy_workload(tables=None, tips=None):
    """(tree arrays, [table arrays, ...]) of the batched rarefaction workload."""
    c = dict(RAREFY)
    tips = tips or c["tips"]
    tree = random_tree(tips, c["seed"])
    sizes = tables or c["tables"]
    out = [random_table(tips, n, c["density"], c["seed"] + 1 + k, None, c["counts"]) for k, n in enumerate(sizes)]
    return tree, out


def workload(name: str, samples: int | None = None, tips: int | None = None):
    """(tree arrays, table arrays) for a named workload; sizes may be overridden for smaller runs."""
    c = dict(CONFIGS[name])
    if samples is not None:
        c["samples"] = samples
    if tips is not None:
        c["tips"] = tips
    tree = random_tree(c["tips"], c["seed"])
    table = random_table(c["tips"], c["samples"], c["density"], c["seed"] + 1, c["zipf"], c["counts"])
    return tree, table
