"""Unweighted occupancy split on the device (through the C ABI of libssu.so).

Nodes held by fewer than theta * n samples of the plan's table leave the int8
Gram; their shared lengths are summed from posting lists in 64-bit fixed
point (a power-of-two scale per row) and added to the Gram's FP64 entries.
SSU_GRAM_SPLIT sets theta (0: off) and SSU_GRAM_SPLIT_MIN_N the smallest
table that is split; it is lowered here so that oracle-sized cases take the
split.

Tolerances: FP64 method names 1e-11 relative against the FP64 oracle, FP32
names 1e-6; unit lengths bit-exact; samples holding the same nodes exactly 0
apart; repeated executions and graph replays bitwise identical.
"""
import sys

import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu

TOL = {"fp64": 1e-11, "fp32": 1e-6}


def say(*a):
    print("[gpu-case]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def _need_device(ssu):
    if not ssu.device_available():
        pytest.fail("no HIP device visible: the HIP path cannot run (no CPU fallback exists)")


@pytest.fixture
def split(monkeypatch):
    def on(theta):
        monkeypatch.setenv("SSU_GRAM_SPLIT", str(theta))
        monkeypatch.setenv("SSU_GRAM_SPLIT_MIN_N", "2")
    return on


def oracle_matrix(O, tree, table, method, **kw):
    st, le, nm = tree
    return O.matrix(O.Tree(np.asarray(st, np.uint8), le, nm), O.Table(*table), method, **kw)


def _case(synth, tips, n, dens, seed, zipf=None):
    return synth.random_tree(tips, seed), synth.random_table(tips, n, dens, seed + 1, zipf)


def _sub_table(table, idx):
    """The columns idx of an observation-major CSR table (observations left empty dropped)."""
    obs_ids, sample_ids, indptr, indices, data = table
    remap = -np.ones(len(sample_ids), np.int64)
    remap[idx] = np.arange(len(idx))
    rows = np.repeat(np.arange(len(obs_ids)), np.diff(np.asarray(indptr, np.int64)))
    keep = remap[np.asarray(indices)] >= 0
    r, c, v = rows[keep], remap[np.asarray(indices)[keep]], np.asarray(data)[keep]
    order = np.lexsort((c, r))
    r, c, v = r[order], c[order], v[order]
    used, counts = np.unique(r, return_counts=True)
    ip = np.zeros(used.size + 1, np.uint32)
    np.cumsum(counts, out=ip[1:])
    return [obs_ids[i] for i in used], [sample_ids[i] for i in idx], ip, c.astype(np.uint32), v


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("theta", ["0", "0.05", "0.3", "1.0"])
def test_split_matches_oracle(ssu, O, synth, split, theta, prec):
    """theta 0: Gram only; 1.0: every node not held by all samples goes to the
    posting lists.  Symmetric, zero diagonal, repeated executions identical."""
    tree, table = _case(synth, 500, 300, 0.05, 73, zipf=1.1)
    split(theta)
    inp = ssu.Inputs(*tree, *table)
    for method, kw in (("unweighted", {}), ("unweighted_unnormalized", {}), ("unweighted", {"bypass_tips": True})):
        say("gram split", theta, prec, method, kw)
        p = ssu.Plan(inp, method + ("_fp64" if prec == "fp64" else ""), **kw)
        p.execute()
        m = p.matrix()
        err = rel_err(m, oracle_matrix(O, tree, table, method, **kw))
        say("max rel err", err)
        assert err <= TOL[prec]
        assert np.array_equal(m, m.T) and np.all(np.diag(m) == 0)
        if theta != "0":
            assert p.sparse_terms() > 0   # the posting lists were walked
        p.execute()
        assert np.array_equal(p.matrix(), m)
        p.close()


def test_split_against_gram_only_across_row_chunks(ssu, O, synth, split):
    """4,500 samples: rows span two 4,096-column chunks of the rows kernel, and
    the diagonal of rows past 4,096 comes from the second chunk."""
    tree, table = _case(synth, 300, 4500, 0.04, 79)
    inp = ssu.Inputs(*tree, *table)
    out = {}
    for theta in ("0", "0.2"):
        split(theta)
        say("gram split row chunks", theta)
        p = ssu.Plan(inp, "unweighted_fp64")
        p.execute()
        out[theta] = p.matrix()
        p.close()
    assert rel_err(out["0.2"], out["0"]) <= 1e-13
    idx = np.sort(np.random.default_rng(5).choice(4500, 48, replace=False))
    idx[-1] = 4499
    ref = oracle_matrix(O, tree, _sub_table(table, idx), "unweighted")
    assert rel_err(out["0.2"][np.ix_(idx, idx)], ref) <= TOL["fp64"]


def test_split_duplicate_samples_exact_zero(ssu, O, synth, split):
    """Samples 20..39 hold the same nodes as 0..19 (other counts): their
    fixed-point shares, the Gram share and the diagonals agree bit for bit."""
    st, le, nm = synth.random_tree(300, 91)
    obs, sids, ip, ix, d = synth.random_table(300, 40, 0.1, 92)
    dense = np.zeros((len(obs), 40))
    for r in range(len(obs)):
        dense[r, ix[ip[r]:ip[r + 1]]] = d[ip[r]:ip[r + 1]]
    dense[:, 20:] = dense[:, :20] * 3.0
    rows = [np.flatnonzero(dense[r]) for r in range(len(obs))]
    keep = [r for r in range(len(obs)) if len(rows[r])]
    table = ([obs[r] for r in keep], sids,
             np.concatenate([[0], np.cumsum([len(rows[r]) for r in keep])]).astype(np.uint32),
             np.concatenate([rows[r] for r in keep]).astype(np.uint32),
             np.concatenate([dense[r, rows[r]] for r in keep]))
    inp = ssu.Inputs(st, le, nm, *table)
    for theta in ("0.3", "1.0"):
        split(theta)
        for name in ("unweighted_fp64", "unweighted", "unweighted_unnormalized_fp64"):
            say("gram split duplicates", theta, name)
            s, m = ssu.matrix_inmem(inp, name)
            assert s == ssu.OKAY
            base = name.replace("_fp64", "")
            assert rel_err(m, oracle_matrix(O, (st, le, nm), table, base)) <= TOL["fp64" if "fp64" in name else "fp32"]
            assert all(m[i, i + 20] == 0.0 for i in range(20))


def test_split_unit_lengths_bit_exact(ssu, O, synth, split):
    """All lengths 1: the fixed-point shares are exact integers, so the split
    keeps the unweighted matrices bit-exact against the oracle."""
    st, le, nm = synth.random_tree(900, 77)
    ones = np.where(st, 1.0, 0.0)
    table = synth.random_table(900, 260, 0.03, 78)
    inp = ssu.Inputs(st, ones, nm, *table)
    split("0.5")
    for method in ("unweighted_unnormalized_fp64", "unweighted_fp64"):
        say("gram split unit", method)
        s, m = ssu.matrix_inmem(inp, method)
        ref = oracle_matrix(O, (st, ones, nm), table, method.replace("_fp64", ""))
        assert s == ssu.OKAY and np.array_equal(m, ref)


def test_split_explicit_zeros_and_empty_sample(ssu, O, synth, split):
    """Stored zeros count toward a node's structural occupancy (the posting
    capacity) but never make a sample present."""
    st, le, nm = synth.random_tree(60, 3)
    obs, sids, ip, ix, d = synth.random_table(60, 10, 0.2, 4)
    d = d.copy()
    d[ix == 3] = 0.0   # sample 3 keeps explicit zeros only
    d[::7] = 0.0       # and some other stored zeros
    table = (obs, sids, ip, ix, d)
    split("1.0")
    for method in ("unweighted_fp64", "unweighted_unnormalized_fp64"):
        say("gram split zeros", method)
        s, m = ssu.matrix_inmem(ssu.Inputs(st, le, nm, *table), method)
        ref = oracle_matrix(O, (st, le, nm), table, method.replace("_fp64", ""))
        assert s == ssu.OKAY and rel_err(m, ref) <= 1e-11


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_split_blocked_embedding(ssu, O, synth, split, monkeypatch, prec):
    """Occupancy counted and sample bits gathered over 128-sample embedding blocks."""
    tree, table = _case(synth, 400, 300, 0.05, 131, zipf=1.1)
    split("0.3")
    monkeypatch.setenv("SSU_EMBED_BLOCK_SAMPLES", "128")
    say("gram split blocked", prec)
    s, m = ssu.matrix_inmem(ssu.Inputs(*tree, *table), "unweighted" + ("_fp64" if prec == "fp64" else ""))
    assert s == ssu.OKAY, ssu.last_error()
    assert rel_err(m, oracle_matrix(O, tree, table, "unweighted")) <= TOL[prec]


def test_split_rarefaction_replicates_and_replays(ssu, O, synth, split):
    """Plans split on the kept table serve every rarefied replicate (presence
    only shrinks); graph replays equal the eager first execution."""
    tree, table = _case(synth, 300, 60, 0.1, 31)
    split("0.3")
    totals = np.bincount(table[3], weights=table[4], minlength=60)
    depth = int(np.sort(totals)[20])
    methods = ["unweighted", "unweighted_fp64"]
    say("gram split rarefaction", depth)
    r = ssu.Rarefaction(ssu.Inputs(*tree, *table), methods, depth, 4, seed=99)
    r.execute()
    si, oi, ip, ix = r.kept()
    st, le, nm = tree
    otree = O.Tree(np.asarray(st, np.uint8), le, nm)
    eager = [[r.matrix(k, rep) for rep in range(4)] for k in range(2)]
    for rep in range(4):
        t = O.Table([table[0][i] for i in oi], [table[1][s] for s in si], ip, ix, r.counts(rep))
        ref = O.matrix(otree, t, "unweighted")
        assert rel_err(eager[0][rep], ref) <= TOL["fp32"]
        assert rel_err(eager[1][rep], ref) <= TOL["fp64"]
    for it in range(2):
        r.execute()
        assert r.graph_used()
        for rep in range(4):
            for k in range(2):
                assert np.array_equal(r.matrix(k, rep), eager[k][rep]), (it, rep, k)


@pytest.mark.parametrize("theta", ["0", "0.3"])
def test_presence_bits_match_fp64_embedding(ssu, O, synth, split, monkeypatch, theta):
    """The Gram path on presence bits (non-negative tables) and on the FP64
    embedding (SSU_GRAM_BITS=0) build the same operands: equal bit for bit."""
    tree, table = _case(synth, 500, 300, 0.05, 83, zipf=1.1)
    split(theta)
    inp = ssu.Inputs(*tree, *table)
    out = {}
    for bits in ("1", "0"):
        monkeypatch.setenv("SSU_GRAM_BITS", bits)
        for name in ("unweighted_fp64", "unweighted", "unweighted_unnormalized"):
            say("presence bits", bits, theta, name)
            s, m = ssu.matrix_inmem(inp, name)
            assert s == ssu.OKAY, ssu.last_error()
            out[bits, name] = m
    for name in ("unweighted_fp64", "unweighted", "unweighted_unnormalized"):
        assert np.array_equal(out["1", name], out["0", name]), name
    assert rel_err(out["1", "unweighted_fp64"], oracle_matrix(O, tree, table, "unweighted")) <= TOL["fp64"]


def test_negative_values_take_fp64_embedding(ssu, O, synth, split):
    """A table with negative values: presence is E > 0 of the summed embedding
    (a node can cancel to zero), so the Gram path keeps the FP64 embedding."""
    st, le, nm = synth.random_tree(200, 17)
    obs, sids, ip, ix, d = synth.random_table(200, 90, 0.1, 18)
    d = d.copy()
    d[::5] *= -1.0
    table = (obs, sids, ip, ix, d)
    split("0.3")
    for name in ("unweighted_fp64", "unweighted"):
        say("negative values", name)
        s, m = ssu.matrix_inmem(ssu.Inputs(st, le, nm, *table), name)
        assert s == ssu.OKAY, ssu.last_error()
        ref = oracle_matrix(O, (st, le, nm), table, "unweighted")
        assert rel_err(m, ref) <= TOL["fp64" if "fp64" in name else "fp32"]
