"""Rarefaction replicates on the device (src/biom_subsampled.cpp:155-321).

The device sampler draws from a counter-based generator, not the reference's
std::mt19937, so parity is pinned four ways:
  * exactly, against a Python model of this library's own draws (CSR order,
    cumulative-count search, Floyd's set): a wrong index or offset fails it;
  * statistically, against the distributions the reference samples from
    (multinomial with replacement, multivariate hypergeometric without):
    replicate means and the pooled variance ratio;
  * the reference's own subsampling tests (src/test_ska.cpp:639-920);
  * every replicate's distance matrix against the oracle run on the
    replicate's copied-back table.
Graph replays are checked against the eager first pass.
"""
import os
import sys

import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu

M64 = (1 << 64) - 1
TOL = {"fp64": 1e-11, "fp32": 1e-6}


def say(*a):
    print("[gpu-case]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def _need_device(ssu):
    if not ssu.device_available():
        pytest.fail("no HIP device visible: the HIP path cannot run (no CPU fallback exists)")


def mix64(z):
    z = (z + 0x9E3779B97F4A7C15) & M64
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
    return z ^ (z >> 31)


def model_counts(indptr, indices, data, n, depth, with_rep, seed, rep):
    """This library's draws for one replicate, restated in Python (kept-table CSR in, counts out)."""
    out = np.zeros(len(data))
    for s in range(n):
        pos = np.nonzero(indices == s)[0]          # ascending CSR positions = ascending rows
        if pos.size == 0:
            continue
        vals = np.maximum(data[pos], 0.0)
        if not with_rep:
            vals = np.floor(vals)
        cum = np.cumsum(vals)
        key = mix64(seed ^ mix64((rep << 32) | s))
        if with_rep:
            total = cum[-1]
            for j in range(depth):
                u = float(mix64((key + j) & M64) >> 11) * 2.0 ** -53 * total
                q = min(int(np.searchsorted(cum, u, side="right")), pos.size - 1)
                out[pos[q]] += 1
        else:
            total = int(cum[-1])
            if total <= depth:
                out[pos] = vals
                continue
            chosen = set()
            for j_i, j in enumerate(range(total - depth, total)):
                pick = mix64((key + j_i) & M64) % (j + 1)
                chosen.add(j if pick in chosen else pick)
            for h in chosen:
                out[pos[int(np.searchsorted(cum, float(h), side="right"))]] += 1
    return out


def filtered(table, depth):
    """Samples with total >= depth and the rows that have an entry in one of them (src/biom_subsampled.cpp:290)."""
    obs_ids, sids, indptr, indices, data = table
    indptr = np.asarray(indptr)
    indices = np.asarray(indices)
    data = np.asarray(data, float)
    n = len(sids)
    totals = np.zeros(n)
    for i in range(len(obs_ids)):
        for j in range(indptr[i], indptr[i + 1]):
            totals[indices[j]] += data[j]
    keep = np.nonzero(totals >= depth)[0]
    newi = -np.ones(n, np.int64)
    newi[keep] = np.arange(keep.size)
    rows, ip, ix, d = [], [0], [], []
    for i in range(len(obs_ids)):
        sel = [j for j in range(indptr[i], indptr[i + 1]) if newi[indices[j]] >= 0]
        if sel:
            rows.append(i)
            ix += [int(newi[indices[j]]) for j in sel]
            d += [data[j] for j in sel]
            ip.append(len(ix))
    return keep, np.array(rows), np.array(ip), np.array(ix), np.array(d)


def star_tree(ssu, names):
    return ssu.newick_to_bp("(" + ",".join(f"{x}:1.0" for x in names) + ");")


def inputs(ssu, tree, table):
    st, le, nm = tree
    return ssu.Inputs(st, le, nm, *table)


def small_case(synth, tips=300, n=60, dens=0.1, seed=7):
    tree = synth.random_tree(tips, seed)
    table = synth.random_table(tips, n, dens, seed + 1, None, "lognormal")
    return tree, table


def depth_dropping_some(table, q=30):
    _, sids, indptr, indices, data = table
    totals = np.bincount(np.asarray(indices), weights=np.asarray(data, float), minlength=len(sids))
    return int(np.percentile(totals, q))


@pytest.mark.parametrize("with_rep", [True, False])
def test_counts_match_model_and_kept_table(ssu, synth, with_rep):
    tree, table = small_case(synth)
    depth = depth_dropping_some(table)
    say("rarefy counts", with_rep, depth)
    r = ssu.Rarefaction(inputs(ssu, tree, table), ["unweighted"], depth, 3, with_replacement=with_rep, seed=12345)
    keep, rows, ip, ix, d = filtered(table, depth)
    assert 0 < keep.size < len(table[1])          # the depth drops some samples
    si, oi, kip, kix = r.kept()
    assert np.array_equal(si, keep) and np.array_equal(oi, rows)
    assert np.array_equal(kip, ip) and np.array_equal(kix, ix)
    r.execute()
    for rep in range(3):
        got = r.counts(rep)
        exp = model_counts(ip, ix, d, keep.size, depth, with_rep, 12345, rep)
        assert np.array_equal(got, exp), rep
        sums = np.bincount(ix, weights=got, minlength=keep.size)
        assert np.all(sums == depth)
        assert np.all(got[d == 0] == 0)
        if not with_rep:
            assert np.all(got <= np.floor(d))
    assert not np.array_equal(r.counts(0), r.counts(1))   # replicates differ


@pytest.mark.parametrize("methods", [["unweighted", "weighted_normalized"],
                                     ["unweighted_fp64", "weighted_normalized_fp64", "weighted_unnormalized_fp64",
                                      "unweighted_unnormalized_fp64"]])
@pytest.mark.parametrize("with_rep", [True, False])
def test_replicate_matrices_match_oracle(ssu, O, synth, methods, with_rep):
    tree, table = small_case(synth, seed=31)
    depth = depth_dropping_some(table, 20)
    say("rarefy matrices", methods, with_rep, depth)
    r = ssu.Rarefaction(inputs(ssu, tree, table), methods, depth, 4, with_replacement=with_rep, seed=99)
    r.execute()
    si, oi, ip, ix = r.kept()
    st, le, nm = tree
    otree = O.Tree(np.asarray(st, np.uint8), le, nm)
    for rep in range(4):
        counts = r.counts(rep)
        t = O.Table([table[0][i] for i in oi], [table[1][s] for s in si], ip, ix, counts)
        for k, m in enumerate(methods):
            base = m[:-5] if m.endswith("_fp64") else m
            ref = O.matrix(otree, t, base)
            err = rel_err(r.matrix(k, rep), ref)
            assert err <= TOL["fp64" if m.endswith("_fp64") else "fp32"], (m, rep, err)


def test_normalized_counts_and_generalized(ssu, O, synth):
    """normalize_sample_counts=False and generalized alpha reach the per-replicate plans."""
    tree, table = small_case(synth, seed=47)
    depth = depth_dropping_some(table, 10)
    methods = ["generalized_fp64", "weighted_normalized_fp64"]
    say("rarefy generalized / unnormalized counts", depth)
    r = ssu.Rarefaction(inputs(ssu, tree, table), methods, depth, 2, seed=5, alpha=0.5, normalize=False)
    r.execute()
    si, oi, ip, ix = r.kept()
    st, le, nm = tree
    otree = O.Tree(np.asarray(st, np.uint8), le, nm)
    for rep in range(2):
        t = O.Table([table[0][i] for i in oi], [table[1][s] for s in si], ip, ix, r.counts(rep))
        assert rel_err(r.matrix(0, rep), O.matrix(otree, t, "generalized", alpha=0.5, normalize=False)) <= 1e-11
        assert rel_err(r.matrix(1, rep), O.matrix(otree, t, "weighted_normalized", normalize=False)) <= 1e-11


def test_graph_replays_equal_eager(ssu, synth):
    tree, table = small_case(synth, tips=400, n=150, seed=61)
    depth = depth_dropping_some(table, 5)
    methods = ["unweighted", "weighted_normalized", "unweighted_fp64"]
    say("rarefy graph replays", depth)
    inp = inputs(ssu, tree, table)
    r = ssu.Rarefaction(inp, methods, depth, 5, seed=2024)
    r.execute()
    assert not r.graph_used()
    eager_c = [r.counts(k) for k in range(5)]
    eager_m = [[r.matrix(k, rep) for rep in range(5)] for k in range(len(methods))]
    other = ssu.Plan(inp, "weighted_normalized")        # eager work of another plan between replays
    for it in range(3):
        other.execute()
        r.execute()
        assert r.graph_used()
        for rep in range(5):
            assert np.array_equal(r.counts(rep), eager_c[rep]), (it, rep)
            for k in range(len(methods)):
                assert np.array_equal(r.matrix(k, rep), eager_m[k][rep]), (it, rep, methods[k])
    other.close()
    assert r.last_ms() > 0 and r.pair_nodes() > 0


@pytest.mark.parametrize("with_rep", [True, False])
def test_distribution(ssu, with_rep):
    """Replicate means against depth*c/T, and the pooled variance against the
    multinomial (with replacement) or hypergeometric (without) variance."""
    rng = np.random.default_rng(3)
    n_obs, n, reps = 12, 6, 1000
    names = [f"O{i}" for i in range(n_obs)]
    counts = rng.integers(1, 40, size=(n_obs, n)).astype(float)
    counts[rng.random((n_obs, n)) < 0.2] = 0
    ip = [0]
    ix, d = [], []
    for i in range(n_obs):
        for s in range(n):
            if counts[i, s] > 0:
                ix.append(s)
                d.append(counts[i, s])
        ip.append(len(ix))
    T = counts.sum(axis=0)
    depth = int(T.min() // 2)
    say("rarefy distribution", with_rep, depth)
    inp = ssu.Inputs(*star_tree(ssu, names), names, [f"S{s}" for s in range(n)], ip, ix, d)
    r = ssu.Rarefaction(inp, ["unweighted"], depth, reps, with_replacement=with_rep, seed=77)
    r.execute()
    draws = np.stack([r.counts(k) for k in range(reps)])            # reps x nnz
    ix = np.asarray(ix)
    d = np.asarray(d)
    p = d / T[ix]
    mean_exp = depth * p
    var_exp = depth * p * (1 - p)
    if not with_rep:
        var_exp = var_exp * (T[ix] - depth) / (T[ix] - 1)
    z = (draws.mean(axis=0) - mean_exp) / np.sqrt(var_exp / reps)
    assert np.max(np.abs(z)) < 5.5, np.max(np.abs(z))
    ratio = draws.var(axis=0, ddof=1).sum() / var_exp.sum()
    assert abs(ratio - 1) < 0.06, ratio


def test_reference_subsample_tests(ssu, ref_tables, ref_trees):
    """src/test_ska.cpp:639-695 and 779-835: test.biom at depth 5 keeps
    Sample1 and Sample4, each column sums to 5; depth 8 leaves nothing."""
    t = ref_tables["test"]
    st, le, nm = ssu.newick_to_bp(ref_trees["test"])
    inp = ssu.Inputs(st, le, nm, t["obs_ids"], t["sample_ids"], t["indptr"], t["indices"], t["data"])
    for with_rep in (True, False):
        say("reference subsample test", with_rep)
        r = ssu.Rarefaction(inp, ["unweighted"], 5, 4, with_replacement=with_rep, seed=1)
        si, oi, ip, ix = r.kept()
        assert [t["sample_ids"][s] for s in si] == ["Sample1", "Sample4"]
        assert set(t["obs_ids"][i] for i in oi) <= {"GG_OTU_2", "GG_OTU_3", "GG_OTU_4"}
        r.execute()
        for rep in range(4):
            assert np.array_equal(np.bincount(ix, weights=r.counts(rep), minlength=2), [5.0, 5.0])
        with pytest.raises(RuntimeError, match=f"status {ssu.TABLE_EMPTY}"):
            ssu.Rarefaction(inp, ["unweighted"], 8, 2, with_replacement=with_rep)


@pytest.mark.parametrize("with_rep,expect", [(True, "more"), (False, "exactly")])
def test_reference_limit_case(ssu, with_rep, expect):
    """src/test_ska.cpp:698-777 and 838-920: one sample of 35 items drawn to 34,
    1000 times: with replacement more than 10 distinct outcomes, without exactly 10."""
    names = [f"OTU{i}" for i in range(10)]
    data = [2., 1., 2., 1., 7., 6., 3., 3., 5., 5.]
    inp = ssu.Inputs(*star_tree(ssu, names), names, ["S1"], list(range(11)), [0] * 10, data)
    say("reference limit case", with_rep)
    r = ssu.Rarefaction(inp, ["unweighted"], 34, 1000, with_replacement=with_rep, seed=11)
    r.execute()
    outcomes = {tuple(r.counts(k)) for k in range(1000)}
    assert all(sum(o) == 34 for o in outcomes)
    if expect == "more":
        assert len(outcomes) > 10
    else:
        assert len(outcomes) == 10
        assert all(np.all(np.array(o) <= np.array(data)) for o in outcomes)


def test_rejects_bad_arguments(ssu, synth):
    tree, table = small_case(synth, n=10)
    inp = inputs(ssu, tree, table)
    with pytest.raises(RuntimeError):
        ssu.Rarefaction(inp, ["unweighted"], 0, 2)
    with pytest.raises(RuntimeError):
        ssu.Rarefaction(inp, ["unweighted"], 10, 0)
    with pytest.raises(RuntimeError, match=f"status {ssu.UNKNOWN_METHOD}"):
        ssu.Rarefaction(inp, ["bogus"], 10, 2)


def test_generalized_fp32_replicates_and_replays(ssu, O, synth):
    """FP32 generalized (shared-node path, occupancy split) on rarefied
    replicates: every replicate against the oracle, graph replays equal eager."""
    tree, table = small_case(synth, seed=53)
    depth = depth_dropping_some(table, 10)
    say("rarefy generalized fp32", depth)
    r = ssu.Rarefaction(inputs(ssu, tree, table), ["generalized"], depth, 3, seed=8, alpha=0.5)
    r.execute()
    si, oi, ip, ix = r.kept()
    st, le, nm = tree
    otree = O.Tree(np.asarray(st, np.uint8), le, nm)
    eager = [r.matrix(0, rep) for rep in range(3)]
    for rep in range(3):
        t = O.Table([table[0][i] for i in oi], [table[1][s] for s in si], ip, ix, r.counts(rep))
        assert rel_err(eager[rep], O.matrix(otree, t, "generalized", alpha=0.5)) <= TOL["fp32"], rep
    r.execute()
    assert r.graph_used()
    for rep in range(3):
        assert np.array_equal(r.matrix(0, rep), eager[rep]), rep


def test_group_launch_equals_each_batch(ssu, O, synth):
    """Three tables of uneven sample counts in one group launch (every batch's
    graph a child of one parent graph): every replicate matrix equal to the
    batch's own eager pass, on two launches in a row; replicate 0 against the
    oracle."""
    tree = synth.random_tree(300, 61)
    tables = [synth.random_table(300, n, 0.1, 62 + n, None, "lognormal") for n in (40, 90, 150)]
    methods = ["unweighted", "weighted_normalized", "weighted_unnormalized_fp64"]
    batches, eager = [], []
    for k, t in enumerate(tables):
        depth = depth_dropping_some(t, 10)
        say("group batch", k, len(t[1]), depth)
        b = ssu.Rarefaction(inputs(ssu, tree, t), methods, depth, 3, seed=20 + k)
        b.execute()   # eager, captures the graph
        eager.append({(m, rep): b.matrix(m, rep) for m in range(len(methods)) for rep in range(3)})
        batches.append(b)
    g = ssu.RarefactionGroup(batches)
    st, le, nm = tree
    otree = O.Tree(np.asarray(st, np.uint8), le, nm)
    for launch in range(2):
        say("group launch", launch)
        g.execute()
        assert g.last_ms() > 0.0
        for k, b in enumerate(batches):
            assert b.graph_used()
            for (m, rep), x in eager[k].items():
                assert np.array_equal(b.matrix(m, rep), x), (k, methods[m], rep)
    for k, (b, t) in enumerate(zip(batches, tables)):
        si, oi, ip, ix = b.kept()
        ot = O.Table([t[0][i] for i in oi], [t[1][s] for s in si], ip, ix, b.counts(0))
        for m, name in enumerate(methods):
            base = name[:-5] if name.endswith("_fp64") else name
            err = rel_err(b.matrix(m, 0), O.matrix(otree, ot, base))
            assert err <= TOL["fp64" if name.endswith("_fp64") else "fp32"], (k, name, err)
    g.close()
    for b in batches:
        b.close()


def test_vaw_shared_node_replicates_and_replays(ssu, O, synth, monkeypatch):
    """VAW on the shared-node path in rarefied replicates (the executions bring
    their own totals: one-sided sums at every sample's total): every replicate
    against the oracle, graph replays equal eager."""
    monkeypatch.setenv("SSU_VAW_SPLIT", "0.5")
    tree, table = small_case(synth, seed=57)
    depth = depth_dropping_some(table, 10)
    methods = ["weighted_normalized_fp64", "unweighted", "weighted_unnormalized"]
    say("rarefy vaw split", depth)
    r = ssu.Rarefaction(inputs(ssu, tree, table), methods, depth, 3, seed=9, variance_adjust=True)
    r.execute()
    si, oi, ip, ix = r.kept()
    st, le, nm = tree
    otree = O.Tree(np.asarray(st, np.uint8), le, nm)
    eager = {(k, rep): r.matrix(k, rep) for k in range(len(methods)) for rep in range(3)}
    for rep in range(3):
        t = O.Table([table[0][i] for i in oi], [table[1][s] for s in si], ip, ix, r.counts(rep))
        for k, m in enumerate(methods):
            base = m[:-5] if m.endswith("_fp64") else m
            err = rel_err(eager[(k, rep)], O.matrix(otree, t, base, variance_adjust=True))
            assert err <= TOL["fp64" if m.endswith("_fp64") else "fp32"], (m, rep, err)
    r.execute()
    assert r.graph_used()
    for (k, rep), m in eager.items():
        assert np.array_equal(r.matrix(k, rep), m), (methods[k], rep)


def test_near_duplicate_refinement_in_replicates(ssu, O, synth):
    """Near-duplicate samples survive the draws (samples 0..39 hold exactly
    `depth` items and are taken whole without replacement; 40..79 repeat them
    with one extra item, of which the draw leaves out one): the FP32 weighted
    and generalized matrices of every replicate are refined to 1e-6 of the
    FP64 oracle, in the eager pass and in graph replays alike."""
    tips, n_base, depth = 300, 40, 2000
    st, le, nm = synth.random_tree(tips, 71)
    rng = np.random.default_rng(72)
    dense = np.zeros((tips, 2 * n_base))
    for s in range(n_base):
        held = rng.choice(tips, 40, replace=False)
        dense[held, s] = rng.multinomial(depth, rng.dirichlet(np.ones(held.size)))
    dense[:, n_base:] = dense[:, :n_base]
    for s in range(n_base, 2 * n_base):
        dense[rng.choice(np.flatnonzero(dense[:, s])), s] += 1.0
    rows = [np.flatnonzero(dense[r]) for r in range(tips)]
    keep = [r for r in range(tips) if len(rows[r])]
    table = ([f"t{r}" for r in keep], [f"s{i}" for i in range(2 * n_base)],
             np.concatenate([[0], np.cumsum([len(rows[r]) for r in keep])]).astype(np.uint32),
             np.concatenate([rows[r] for r in keep]).astype(np.uint32),
             np.concatenate([dense[r, rows[r]] for r in keep]))
    methods = ["unweighted", "weighted_normalized", "generalized"]
    say("rarefy near duplicates")
    r = ssu.Rarefaction(inputs(ssu, (st, le, nm), table), methods, depth, 3, with_replacement=False, seed=4,
                        alpha=0.5)
    r.execute()
    assert not r.graph_used() and r.refined_pairs() > 0
    si, oi, ip, ix = r.kept()
    assert si.size == 2 * n_base
    otree = O.Tree(np.asarray(st, np.uint8), le, nm)
    eager = [[r.matrix(k, rep) for rep in range(3)] for k in range(len(methods))]
    for rep in range(3):
        t = O.Table([table[0][i] for i in oi], [table[1][s] for s in si], ip, ix, r.counts(rep))
        for k, m in enumerate(methods):
            kw = {"alpha": 0.5} if m == "generalized" else {}
            err = rel_err(eager[k][rep], O.matrix(otree, t, m, **kw))
            assert err <= TOL["fp32"], (m, rep, err)
    refined_eager = r.refined_pairs()
    r.execute()
    assert r.graph_used() and r.refined_pairs() == refined_eager
    for k in range(len(methods)):
        for rep in range(3):
            assert np.array_equal(r.matrix(k, rep), eager[k][rep]), (methods[k], rep)
