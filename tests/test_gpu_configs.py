"""The BASELINE configurations other than the 16k-sample bench workload, at
their real sizes, on the device (C ABI of libssu.so):

* cpu_1k_256 (1,000 tips, 256 samples, 5% density, lognormal counts): the
  full 256 x 256 matrices against the FP64 oracle -- FP64 names within 1e-11,
  FP32 names within 1e-6 -- and Faith PD bit-exact;
* w_50k_8192 (50,000 tips, 8,192 samples, 2% density, negative-binomial
  counts): weighted_normalized, generalized (alpha 0.5), VAW weighted and VAW
  generalized (alpha 0.5) through their default paths (occupancy splits):
  symmetric, zero diagonal, within [0, 1]; the first 64 rows against the plain
  dense pair tiles (2e-6 relative); a 48-sample subset within 1e-6 of the FP64
  oracle;
* rarefy_20k_4tables (20,000 tips, tables of 512 / 1,536 / 3,000 / 6,000
  samples, 32 replicates at depth 1,000 with replacement, unweighted +
  weighted_normalized): graph replays equal the eager pass, replicate 0 of
  every table within 1e-6 of the FP64 oracle on a 48-sample subset of its
  drawn counts;
* large_200k_65536 (200,000 tips, 65,536 samples, 0.5% density): unweighted
  and weighted_normalized, one plan resident at a time: the first 64 rows
  symmetric with the columns, zero diagonal, within [0, 1], against the
  no-split paths, and a 48-sample subset within 1e-6 of the FP64 oracle.
"""
import sys

import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-case]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def _need_device(ssu):
    if not ssu.device_available():
        pytest.fail("no HIP device visible: the HIP path cannot run (no CPU fallback exists)")


def sub_table(table, idx):
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


def oracle_on(O, tree, table, idx, method, **kw):
    st, le, nm = tree
    return O.matrix(O.Tree(np.asarray(st, np.uint8), le, nm), O.Table(*sub_table(table, idx)), method, **kw)


def subset(n, seed=7, k=46):
    return np.unique(np.concatenate([np.random.default_rng(seed).choice(n, k, replace=False), [0, n - 1]]))


def test_cfg1_cpu_1k_256_full_matrix(ssu, O, synth):
    tree, table = synth.workload("cpu_1k_256")
    n = len(table[1])
    assert n == 256
    inp = ssu.Inputs(*tree, *table)
    st, le, nm = tree
    otree, otab = O.Tree(np.asarray(st, np.uint8), le, nm), O.Table(*table)
    for name in ("unweighted_fp64", "weighted_normalized_fp64", "unweighted", "weighted_normalized"):
        say("cfg1", name)
        s, m = ssu.matrix_inmem(inp, name)
        assert s == ssu.OKAY, ssu.last_error()
        ref = O.matrix(otree, otab, name.replace("_fp64", ""))
        err = rel_err(m, ref)
        say("max rel err", err)
        assert err <= (1e-11 if name.endswith("_fp64") else 1e-6)
    say("cfg1 faith")
    s, vals, _ = ssu.faith_pd_inmem(inp)
    assert s == ssu.OKAY and np.array_equal(vals, O.faith_pd(otree, otab))


@pytest.fixture(scope="module")
def cfg3(synth):
    return synth.workload("w_50k_8192")


@pytest.mark.parametrize("name,kw", [("weighted_normalized", {}), ("generalized", {"alpha": 0.5}),
                                     ("weighted_normalized", {"variance_adjust": True}),
                                     ("generalized", {"alpha": 0.5, "variance_adjust": True})],
                         ids=["weighted_normalized", "generalized-0.5", "vaw-weighted_normalized",
                              "vaw-generalized-0.5"])
def test_cfg3_w_50k_8192(ssu, O, cfg3, name, kw):
    tree, table = cfg3
    n = len(table[1])
    assert n == 8192
    inp = ssu.Inputs(*tree, *table)
    say("cfg3", name, kw)
    p = ssu.Plan(inp, name, **kw)
    p.execute()
    m = p.matrix()
    say("main kernel", p.main_kernel(), "timings", p.timings())
    p_theta = p.split_theta()
    assert p_theta > 0.0   # every one of these takes an occupancy split at this size
    p.close()
    assert np.all(np.isfinite(m))
    assert np.array_equal(m, m.T) and np.all(np.diag(m) == 0)
    assert m.min() >= 0.0 and m.max() <= 1.0 + 1e-6
    say("cfg3", name, "split theta", p_theta, "plain dense tiles, rows 0..63")
    q = ssu.Plan(inp, name, plain_kernel=1, **kw)
    q.execute()
    band = q.rows(0, 64)
    q.close()
    assert rel_err(m[:64], band) <= 2e-6
    idx = subset(n)
    ref = oracle_on(O, tree, table, idx, name, **kw)
    err = rel_err(m[np.ix_(idx, idx)], ref)
    say("oracle subset max rel err", err)
    assert err <= 1e-6


def test_cfg4_rarefy_20k_4tables(ssu, O, synth):
    c = synth.RAREFY
    tree, tables = synth.rarefy_workload()
    methods = ["unweighted", "weighted_normalized"]
    st, le, nm = tree
    otree = O.Tree(np.asarray(st, np.uint8), le, nm)
    for k, t in enumerate(tables):
        say("cfg4 table", k, len(t[1]), "samples")
        r = ssu.Rarefaction(ssu.Inputs(*tree, *t), methods, c["depth"], c["reps"], with_replacement=True,
                            seed=1000 + k)
        r.execute()
        assert not r.graph_used()
        reps_checked = (0, c["reps"] - 1)
        eager = {(mi, rep): r.matrix(mi, rep) for mi in range(len(methods)) for rep in reps_checked}
        counts0 = r.counts(0)
        r.execute()
        assert r.graph_used()
        for key, m in eager.items():
            assert np.array_equal(r.matrix(*key), m), key
        si, oi, ip, ix = r.kept()
        assert si.size == r.n and r.n > 0
        rt = ([t[0][i] for i in oi], [t[1][s] for s in si], ip, ix, counts0)
        idx = subset(r.n, seed=k + 1)
        for mi, m in enumerate(methods):
            ref = oracle_on(O, tree, rt, idx, m)
            err = rel_err(eager[(mi, 0)][np.ix_(idx, idx)], ref)
            say("table", k, m, "replicate 0 oracle subset max rel err", err)
            assert err <= 1e-6
        r.close()


@pytest.fixture(scope="module")
def cfg5(synth):
    return synth.workload("large_200k_65536")


@pytest.mark.parametrize("method", ["unweighted", "weighted_normalized"])
def test_cfg5_large_200k_65536(ssu, O, cfg5, monkeypatch, method):
    tree, table = cfg5
    n = len(table[1])
    assert n == 65536
    inp = ssu.Inputs(*tree, *table)
    idx = subset(n)
    say("cfg5", method, "default paths")
    p = ssu.Plan(inp, method)
    p.execute()
    assert p.sparse_terms() > 0
    say("timings", p.timings())
    band = p.rows(0, 64)
    sub = np.stack([p.rows(int(i), 1)[0] for i in idx])[:, idx]
    cols = np.stack([p.rows(int(i), 1)[0][:64] for i in range(64, 128)])   # rows 64..127, columns 0..63
    p.close()
    assert np.all(np.isfinite(band))
    assert np.all(band[np.arange(64), np.arange(64)] == 0)
    assert band.min() >= 0.0 and band.max() <= 1.0 + 1e-6
    assert np.array_equal(cols, band[:, 64:128].T)   # symmetry across the 64 x 64 block off the diagonal
    say("cfg5", method, "no split, rows 0..63")
    monkeypatch.setenv("SSU_GRAM_SPLIT" if method == "unweighted" else "SSU_SPARSE_SPLIT", "0")
    q = ssu.Plan(inp, method)
    q.execute()
    assert q.sparse_terms() == 0
    band0 = q.rows(0, 64)
    q.close()
    assert rel_err(band, band0) <= 2e-6
    ref = oracle_on(O, tree, table, idx, method)
    err = rel_err(sub, ref)
    say("oracle subset max rel err", err)
    assert err <= 1e-6
