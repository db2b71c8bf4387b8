"""Seeded synthetic workloads for benchmarks and parity tests.

Tree: random binary topology built by repeatedly joining two uniformly chosen
active nodes; branch lengths Exp(mean 0.1) rounded to FP32 (so that a newick
round trip through a float parse reproduces them exactly); root length 0;
tips named t<i>.

Table: each sample draws its nonzero tips without replacement, uniformly or
with Zipf(s) tip popularity over a random rank permutation (Efraimidis-
Spirakis keys, implemented as Gumbel top-k); counts lognormal or negative
binomial.  Only observed tips become table rows.
"""
from __future__ import annotations

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


def random_table(n_tips: int, n_samples: int, density: float, seed: int, zipf_s: float | None = None,
                 counts: str = "lognormal", batch: int = 128):
    """Observation-major CSR: (obs_ids, sample_ids, indptr, indices, data)."""
    rng = np.random.default_rng(seed)
    k = max(1, int(round(density * n_tips)))
    if zipf_s is not None:
        ranks = rng.permutation(n_tips) + 1
        logw = (-zipf_s * np.log(ranks)).astype(np.float32)
    tips_all = np.empty(n_samples * k, np.int64)
    for b0 in range(0, n_samples, batch):
        b1 = min(n_samples, b0 + batch)
        u = rng.random((b1 - b0, n_tips), dtype=np.float32)
        np.maximum(u, np.float32(1e-30), out=u)
        if zipf_s is None:
            keys = u
        else:
            # Efraimidis-Spirakis: top-k of u^(1/w)  <=>  top-k of log(u)/w
            keys = np.log(u) * np.exp(-logw)[None, :]
        top = np.argpartition(-keys, k - 1, axis=1)[:, :k]
        tips_all[b0 * k:b1 * k] = top.reshape(-1)
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
