"""Parity of the HIP path (through the C ABI of libssu.so) with the oracle,
the reference's golden vectors and the reference-written vectors.

Tolerances (stated per test):
  * FP64 method names: relative 1e-11 against the FP64 oracle
    (the unweighted Gram is exact up to one FP64 rounding per K-chunk)
  * FP32 method names: relative 1e-6 per matrix entry against the FP64 oracle
  * Faith PD and unit-length unweighted counts: bit-exact
"""
import ctypes as C
import os
import sys

import numpy as np
import pytest

from conftest import GOLDEN, rel_err, stripes_of

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-case]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def _need_device(ssu):
    if not ssu.device_available():
        pytest.fail("no HIP device visible: the HIP path cannot run (no CPU fallback exists)")


def inputs_from(ssu, tree, table):
    st, le, nm = tree
    return ssu.Inputs(st, le, nm, *table)


def oracle_matrix(O, tree, table, method, **kw):
    st, le, nm = tree
    return O.matrix(O.Tree(np.asarray(st, np.uint8), le, nm), O.Table(*table), method, **kw)


@pytest.fixture(scope="module")
def test_fixture(ssu, ref_tables, ref_trees):
    t = ref_tables["test"]
    st, le, nm = ssu.newick_to_bp(ref_trees["test"])
    return ssu.Inputs(st, le, nm, t["obs_ids"], t["sample_ids"], t["indptr"], t["indices"], t["data"])


GOLDEN_KEYS = ["unweighted", "unweighted_bypass", "unweighted_unnormalized", "weighted_unnormalized",
               "weighted_normalized", "generalized_a1", "generalized_a0", "generalized_a05",
               "vaw_weighted_normalized"]


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("key", GOLDEN_KEYS)
def test_goldens_through_c_api(ssu, goldens, test_fixture, key, prec):
    g = goldens[key]
    method = g["method"] + ("_fp64" if prec == "fp64" else "")
    say(key, prec)
    st, m = ssu.matrix_inmem(test_fixture, method, variance_adjust=g["vaw"], alpha=g["alpha"],
                             bypass_tips=g["bypass_tips"])
    assert st == ssu.OKAY, ssu.last_error()
    # src/test_helper.hpp:81-93: absolute 1e-6
    assert np.max(np.abs(stripes_of(m) - np.array(g["stripes"]))) < 1e-6, g["cite"]


def test_faith_goldens_through_c_api(ssu, goldens, ref_tables, ref_trees):
    t = ref_tables["test"]
    for tree_key, gold in [("test", "faith_test_tre"), ("test_faith", "faith_test_faith_tre"),
                           ("test_faith_shear", "faith_test_faith_shear_tre")]:
        say("faith", tree_key)
        st, le, nm = ssu.newick_to_bp(ref_trees[tree_key])
        inp = ssu.Inputs(st, le, nm, t["obs_ids"], t["sample_ids"], t["indptr"], t["indices"], t["data"])
        s, vals, ids = ssu.faith_pd_inmem(inp)
        assert s == ssu.OKAY
        assert np.array_equal(vals, np.array(goldens[gold]["values"], float))


def test_crawford_through_c_api(ssu, O, ref_tables, ref_trees, ref_expected):
    t = ref_tables["crawford"]
    st, le, nm = ssu.newick_to_bp(ref_trees["crawford"])
    inp = ssu.Inputs(st, le, nm, t["obs_ids"], t["sample_ids"], t["indptr"], t["indices"], t["data"])
    for key, vaw in [("crawford_unweighted_fp64", False), ("crawford_unweighted_fp64_vaw", True)]:
        say(key)
        s, m = ssu.matrix_inmem(inp, "unweighted_fp64", variance_adjust=vaw)
        assert s == ssu.OKAY
        exp = ref_expected[key]
        idx = [t["sample_ids"].index(x) for x in exp["ids"]]
        assert rel_err(m[np.ix_(idx, idx)], np.array(exp["matrix"])) < 1e-12
    say("crawford faith")
    s, vals, ids = ssu.faith_pd_inmem(inp)
    exp = dict(ref_expected["crawford_faith"])
    for sid, v in zip(ids, vals):
        assert "%.16g" % v == exp[sid]   # ci/crawford_test.sh:40-44 (md5 of all digits)


def _case(synth, tips, n, dens, seed, zipf=None, counts="lognormal"):
    tree = synth.random_tree(tips, seed)
    table = synth.random_table(tips, n, dens, seed + 1, zipf, counts)
    return tree, table


PARITY = [
    # method, kwargs
    ("unweighted", {}),
    ("unweighted_unnormalized", {}),
    ("unweighted", {"bypass_tips": True}),
    ("weighted_normalized", {}),
    ("weighted_normalized", {"normalize": False}),
    ("weighted_unnormalized", {}),
    ("weighted_normalized", {"bypass_tips": True}),
    ("generalized", {"alpha": 0.5}),
    ("generalized", {"alpha": 0.0}),
    ("generalized", {"alpha": 1.0}),
    ("unweighted", {"variance_adjust": True}),
    ("unweighted_unnormalized", {"variance_adjust": True}),
    ("weighted_normalized", {"variance_adjust": True}),
    ("weighted_unnormalized", {"variance_adjust": True}),
    ("generalized", {"alpha": 0.5, "variance_adjust": True}),
]
TOL = {"fp64": 1e-11, "fp32": 1e-6}


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("method,kw", PARITY)
def test_parity_vs_oracle(ssu, O, synth, method, kw, prec):
    tree, table = _case(synth, 700, 201, 0.04, 101, zipf=1.1)
    say("parity", method, kw, prec)
    s, m = ssu.matrix_inmem(inputs_from(ssu, tree, table), method + ("_fp64" if prec == "fp64" else ""), **kw)
    assert s == ssu.OKAY, ssu.last_error()
    ref = oracle_matrix(O, tree, table, method, **kw)
    err = rel_err(m, ref)
    say("max rel err", err)
    assert err <= TOL[prec]


@pytest.mark.parametrize("n", [1, 2, 3, 127, 128, 129, 300])
def test_sample_count_edges(ssu, O, synth, n):
    tree, table = _case(synth, 150, n, 0.1, 500 + n)
    for method in ("unweighted_fp64", "weighted_normalized_fp64", "weighted_unnormalized"):
        say("edge n", n, method)
        s, m = ssu.matrix_inmem(inputs_from(ssu, tree, table), method)
        assert s == ssu.OKAY, ssu.last_error()
        base = method.replace("_fp64", "")
        ref = oracle_matrix(O, tree, table, base)
        assert rel_err(m, ref) <= TOL["fp64" if method.endswith("fp64") else "fp32"]


def test_unit_lengths_bit_exact(ssu, O, synth):
    """All lengths 1: unweighted counts are integers, the Gram is exact, ratios correctly rounded."""
    st, le, nm = synth.random_tree(900, 77)
    ones = np.where(st, 1.0, 0.0)
    table = synth.random_table(900, 260, 0.03, 78)
    inp = ssu.Inputs(st, ones, nm, *table)
    for method in ("unweighted_unnormalized_fp64", "unweighted_fp64"):
        say("unit", method)
        s, m = ssu.matrix_inmem(inp, method)
        ref = oracle_matrix(O, (st, ones, nm), table, method.replace("_fp64", ""))
        assert s == ssu.OKAY and np.array_equal(m, ref)
    s, m32 = ssu.matrix_inmem(inp, "unweighted_unnormalized")
    ref_u = oracle_matrix(O, (st, ones, nm), table, "unweighted_unnormalized").astype(np.float32)
    assert s == ssu.OKAY and np.array_equal(m32, ref_u)


def test_faith_bit_exact_vs_oracle(ssu, O, synth):
    tree, table = _case(synth, 3000, 400, 0.02, 901, zipf=1.1)
    say("faith random")
    s, vals, ids = ssu.faith_pd_inmem(inputs_from(ssu, tree, table))
    st, le, nm = tree
    ref = O.faith_pd(O.Tree(st.astype(np.uint8), le, nm), O.Table(*table))
    assert s == ssu.OKAY and np.array_equal(vals, ref)


def test_empty_sample_and_zero_lengths(ssu, O, synth):
    st, le, nm = synth.random_tree(60, 3)
    obs, sids, ip, ix, d = synth.random_table(60, 10, 0.2, 4)
    d = d.copy()
    d[ix == 3] = 0.0  # sample 3 keeps explicit zeros only
    table = (obs, sids, ip, ix, d)
    for method in ("unweighted_fp64", "weighted_normalized_fp64", "weighted_unnormalized_fp64"):
        say("empty sample", method)
        s, m = ssu.matrix_inmem(ssu.Inputs(st, le, nm, *table), method)
        ref = oracle_matrix(O, (st, le, nm), table, method.replace("_fp64", ""))
        assert s == ssu.OKAY and rel_err(m, ref) <= 1e-11
    for method, tol in (("generalized_fp64", 1e-11), ("generalized", 1e-6)):   # the shared-node path
        for theta in ("0", "0.5"):
            say("empty sample", method, theta)
            os.environ["SSU_SPARSE_SPLIT"] = theta
            try:
                s, m = ssu.matrix_inmem(ssu.Inputs(st, le, nm, *table), method, alpha=0.5)
            finally:
                del os.environ["SSU_SPARSE_SPLIT"]
            ref = oracle_matrix(O, (st, le, nm), table, "generalized", alpha=0.5)
            assert s == ssu.OKAY and rel_err(m, ref) <= tol
            assert np.array_equal(np.isnan(m), np.isnan(ref))
    for method in ("unweighted", "weighted_normalized", "weighted_unnormalized", "generalized"):   # VAW, both paths
        for prec, tol in (("_fp64", 1e-11), ("", 1e-6)):
            for theta in ("0", "0.5"):
                say("empty sample vaw", method + prec, theta)
                os.environ["SSU_VAW_SPLIT"] = theta
                try:
                    s, m = ssu.matrix_inmem(ssu.Inputs(st, le, nm, *table), method + prec, variance_adjust=True,
                                            alpha=0.5)
                finally:
                    del os.environ["SSU_VAW_SPLIT"]
                ref = oracle_matrix(O, (st, le, nm), table, method, variance_adjust=True, alpha=0.5)
                assert s == ssu.OKAY and rel_err(m, ref) <= tol
                assert np.array_equal(np.isnan(m), np.isnan(ref))
    zero = np.zeros_like(le)
    s, m = ssu.matrix_inmem(ssu.Inputs(st, zero, nm, *table), "unweighted_fp64")
    off = ~np.eye(10, dtype=bool)
    assert s == ssu.OKAY and np.all(np.isnan(m[off]))   # 0/0, as the reference computes it


def test_reference_vectors_on_gpu(ssu):
    """tests/golden/ref_vectors.npz holds outputs of the reference implementation."""
    z = np.load(os.path.join(GOLDEN, "ref_vectors.npz"), allow_pickle=False)
    cases = sorted(k[:-len("__out")] for k in z.files if k.endswith("__out"))
    for c in cases:
        prec = c.split("__")[-1]
        meta = z[c + "__meta"]
        method = str(z[c + "__method"])
        inp = ssu.Inputs(z[c + "__structure"].astype(bool), z[c + "__lengths"], list(z[c + "__names"]),
                         list(z[c + "__obs_ids"]), list(z[c + "__sample_ids"]), z[c + "__indptr"],
                         z[c + "__indices"], z[c + "__data"])
        say("ref vectors", c)
        s, m = ssu.matrix_inmem(inp, method + "_fp64", variance_adjust=bool(meta[0]), alpha=float(meta[1]),
                                bypass_tips=bool(meta[2]), normalize=bool(meta[3]))
        assert s == ssu.OKAY
        got = stripes_of(m, z[c + "__out"].shape[0])
        if prec == "fp64":
            assert rel_err(got, z[c + "__out"]) < 1e-11, c
        else:
            # the reference's own FP32 path vs our FP64: bounded by the reference's FP32 error
            assert rel_err(got, z[c + "__out"]) < 1e-4, c


def test_plain_fp32_unweighted_matches_gram(ssu, O, synth):
    tree, table = _case(synth, 500, 300, 0.05, 61, zipf=1.1)
    inp = inputs_from(ssu, tree, table)
    say("plain vs gram")
    p_gram = ssu.Plan(inp, "unweighted", plain_kernel=0)
    p_plain = ssu.Plan(inp, "unweighted", plain_kernel=1)
    p_gram.execute()
    p_plain.execute()
    assert p_gram.main_kernel() == "k_gram_i8r" and p_plain.main_kernel() == "k_pairs"
    ref = oracle_matrix(O, tree, table, "unweighted")
    assert rel_err(p_gram.matrix(), ref) <= 1e-6
    assert rel_err(p_plain.matrix(), ref) <= 1e-5
    p_gram.close()
    p_plain.close()


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("method", ["weighted_normalized", "weighted_unnormalized"])
def test_weighted_shared_node_kernel_matches_dense(ssu, O, synth, method, prec):
    """Shared-node (min) formulation vs the dense pair-tile kernel vs the FP64 oracle;
    n = 300 leaves a ragged last 64-sample panel."""
    tree, table = _case(synth, 500, 300, 0.05, 71, zipf=1.1)
    inp = inputs_from(ssu, tree, table)
    name = method + ("_fp64" if prec == "fp64" else "")
    say("shared-node vs dense", name)
    ps = ssu.Plan(inp, name, plain_kernel=0)
    pd = ssu.Plan(inp, name, plain_kernel=1)
    ps.execute()
    pd.execute()
    assert ps.main_kernel() == ("k_pairs_sparse" if prec == "fp64" else "k_pairs_sparse32") and pd.main_kernel() == "k_pairs"
    ref = oracle_matrix(O, tree, table, method)
    ms, md = ps.matrix(), pd.matrix()
    assert rel_err(ms, ref) <= TOL[prec]
    assert rel_err(md, ref) <= TOL[prec]
    assert np.array_equal(ms, ms.T) and np.all(np.diag(ms) == 0)
    ps.close()
    pd.close()


@pytest.mark.parametrize("prec", ["fp32", "fp64"])
@pytest.mark.parametrize("theta", ["0", "0.05", "0.3", "1.0"])
def test_weighted_occupancy_split(ssu, O, synth, monkeypatch, theta, prec):
    """Shared-node path with nodes held by fewer than theta * n samples summed
    from posting lists (fixed-point row sums; FP32 or FP64 values) and added in
    the tiled kernel's epilogue; theta 0 turns the split off, 1.0 sends almost
    every node to the posting lists.  Repeated executions are bitwise identical."""
    tree, table = _case(synth, 500, 300, 0.05, 73, zipf=1.1)
    monkeypatch.setenv("SSU_SPARSE_SPLIT", theta)
    inp = inputs_from(ssu, tree, table)
    for method in ("weighted_normalized", "weighted_unnormalized"):
        say("occupancy split", theta, method, prec)
        p = ssu.Plan(inp, method + ("_fp64" if prec == "fp64" else ""))
        p.execute()
        m = p.matrix()
        assert rel_err(m, oracle_matrix(O, tree, table, method)) <= TOL[prec]
        if theta != "0":
            assert p.sparse_terms() > 0
        assert np.array_equal(m, m.T) and np.all(np.diag(m) == 0)
        p.execute()
        assert np.array_equal(p.matrix(), m)
        p.close()


def _sub_table(table, idx):
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


@pytest.mark.parametrize("prec", ["fp32", "fp64"])
def test_weighted_split_across_row_chunks(ssu, O, synth, monkeypatch, prec):
    """4,500 samples: rows span two 4,096-column chunks of the posting-list
    kernel.  Split against no split on the device, and a 48-sample subset
    against the FP64 oracle."""
    tree, table = _case(synth, 300, 4500, 0.04, 79)
    inp = inputs_from(ssu, tree, table)
    name = "weighted_normalized" + ("_fp64" if prec == "fp64" else "")
    out = {}
    for theta in ("0", "0.2"):
        monkeypatch.setenv("SSU_SPARSE_SPLIT", theta)
        say("row chunks", theta, prec)
        p = ssu.Plan(inp, name)
        p.execute()
        out[theta] = p.matrix()
        p.close()
    assert rel_err(out["0.2"], out["0"]) <= (2e-6 if prec == "fp32" else 1e-12)
    idx = np.sort(np.random.default_rng(5).choice(4500, 48, replace=False))
    idx[-1] = 4499
    ref = oracle_matrix(O, tree, _sub_table(table, idx), "weighted_normalized")
    assert rel_err(out["0.2"][np.ix_(idx, idx)], ref) <= TOL[prec]


@pytest.mark.parametrize("method,kw", [("unweighted", {}), ("weighted_normalized", {}),
                                       ("weighted_normalized_fp64", {}), ("generalized", {"alpha": 0.5})])
def test_wide_sample_index(ssu, O, synth, method, kw):
    """65,600 samples: posting lists carry 32-bit sample indices (16-bit up to
    65,536) and the row walk spans 17 column chunks.  The occupancy split must
    be in use; rows of 48 samples on both sides of index 65,536, copied as row
    bands, against the FP64 oracle on those samples."""
    n = 65600
    tree, table = _case(synth, 300, n, 0.02, 83)
    inp = inputs_from(ssu, tree, table)
    idx = np.unique(np.concatenate([np.random.default_rng(8).choice(n, 44, replace=False), [0, 65535, 65536, n - 1]]))
    say("wide sample index", method, n)
    p = ssu.Plan(inp, method, **kw)
    p.execute()
    assert p.sparse_terms() > 0
    sub = np.stack([p.rows(int(i), 1)[0, idx] for i in idx])
    band = p.rows(65530, 12)
    p.close()
    assert band.shape == (12, n)
    assert np.array_equal(band[6, idx], sub[list(idx).index(65536)])   # the band's row 65,536
    ref = oracle_matrix(O, tree, _sub_table(table, idx), method.replace("_fp64", ""), **kw)
    assert rel_err(sub, ref) <= TOL["fp64" if method.endswith("fp64") else "fp32"]
    assert np.array_equal(sub, sub.T) and np.all(np.diag(sub) == 0)


def test_sparse_terms_count_launched_rows_only(ssu, synth, monkeypatch):
    """sparse_terms() sums the walked postings of the workgroups the row walk
    launched: a generalized plan (half-width column chunks, twice the
    workgroups) leaves its counts in memory a later weighted plan may get, and
    the weighted count must not include them."""
    tree, table = _case(synth, 300, 4500, 0.04, 79)
    inp = inputs_from(ssu, tree, table)
    monkeypatch.setenv("SSU_SPARSE_SPLIT", "0.2")
    counts = []
    for method, kw in (("weighted_normalized", {}), ("generalized", {"alpha": 0.5}), ("weighted_normalized", {})):
        say("sparse terms", method)
        p = ssu.Plan(inp, method, **kw)
        p.execute()
        counts.append(p.sparse_terms())
        p.close()
    assert counts[0] > 0 and counts[2] == counts[0]
    assert counts[1] == counts[0]   # the same sparse nodes at the same theta: the same postings


def test_weighted_duplicate_samples(ssu, O, synth):
    """Identical samples: d = 0 exactly in FP64 arithmetic; the shared-node
    kernel gets S_i + S_j - 2M, clamped at 0, within the absolute bound."""
    st, le, nm = synth.random_tree(300, 91)
    obs, sids, ip, ix, d = synth.random_table(300, 40, 0.1, 92)
    dense = np.zeros((len(obs), 40))
    for r in range(len(obs)):
        dense[r, ix[ip[r]:ip[r + 1]]] = d[ip[r]:ip[r + 1]]
    dense[:, 20:] = dense[:, :20] * 3.0      # same proportions, other totals
    rows = [np.flatnonzero(dense[r]) for r in range(len(obs))]
    keep = [r for r in range(len(obs)) if len(rows[r])]
    obs2 = [obs[r] for r in keep]
    ip2 = np.concatenate([[0], np.cumsum([len(rows[r]) for r in keep])]).astype(np.uint32)
    ix2 = np.concatenate([rows[r] for r in keep]).astype(np.uint32)
    d2 = np.concatenate([dense[r, rows[r]] for r in keep])
    table = (obs2, sids, ip2, ix2, d2)
    inp = ssu.Inputs(st, le, nm, *table)
    for name, tol in (("weighted_normalized_fp64", 1e-11), ("weighted_normalized", 1e-6),
                      ("weighted_unnormalized", 1e-6)):
        say("duplicates", name)
        s, m = ssu.matrix_inmem(inp, name)
        assert s == ssu.OKAY
        ref = oracle_matrix(O, (st, le, nm), table, name.replace("_fp64", ""))
        assert rel_err(m, ref) <= tol
        dup = np.array([m[i, i + 20] for i in range(20)])
        assert np.all(dup >= 0) and np.max(dup) <= tol


def test_weighted_negative_counts_use_dense_kernel(ssu, O, synth):
    """min(u, v) only vanishes for u, v >= 0: negative counts select the dense kernel."""
    tree, table = _case(synth, 200, 50, 0.1, 93)
    obs, sids, ip, ix, d = table
    d = d.copy()
    d[::7] *= -1.0
    table = (obs, sids, ip, ix, d)
    inp = inputs_from(ssu, tree, table)
    say("negative counts")
    p = ssu.Plan(inp, "weighted_unnormalized_fp64")
    p.execute()
    assert p.main_kernel() == "k_pairs"
    ref = oracle_matrix(O, tree, table, "weighted_unnormalized")
    assert rel_err(p.matrix(), ref) <= 1e-11
    p.close()


# 200 tips x 50 samples: the totals pass runs on one thread; 4,000 x 512 at 60%
# (1,228,800 entries, above 2^20): on several
CHECK_SIZES = [(200, 50, 0.1), (4000, 512, 0.6)]


@pytest.mark.parametrize("tips,n,dens", CHECK_SIZES)
def test_sample_index_out_of_range_is_an_error(ssu, synth, tips, n, dens):
    """A CSR sample index at or past n_samples fails the call with an error and no device work."""
    tree, table = _case(synth, tips, n, dens, 95)
    obs, sids, ip, ix, d = table
    assert (len(d) >= 1 << 20) == (tips > 1000)
    ix = ix.copy()
    ix[len(ix) // 2] = n
    inp = inputs_from(ssu, tree, (obs, sids, ip, ix, d))
    say("index out of range", tips, n)
    s, m = ssu.matrix_inmem(inp, "unweighted")
    assert s == ssu.OUTPUT_ERROR and m is None
    assert "sample index out of range" in ssu.last_error()


@pytest.mark.parametrize("tips,n,dens", CHECK_SIZES)
@pytest.mark.parametrize("bad", [-1.0, float("nan"), float("inf")])
def test_values_outside_shared_node_domain_use_dense_kernel(ssu, synth, tips, n, dens, bad):
    """One value that is negative or not finite anywhere in the table keeps the weighted
    metrics off the shared-node kernels; the same table with that value set to 1 takes them."""
    tree, table = _case(synth, tips, n, dens, 97)
    obs, sids, ip, ix, d = table
    d = d.copy()
    d[len(d) // 3] = bad
    inp = inputs_from(ssu, tree, (obs, sids, ip, ix, d))
    say("values outside domain", tips, n, bad)
    p = ssu.Plan(inp, "weighted_unnormalized")
    assert p.main_kernel() == "k_pairs"
    p.close()
    d[len(d) // 3] = 1.0
    inp = inputs_from(ssu, tree, (obs, sids, ip, ix, d))
    p = ssu.Plan(inp, "weighted_unnormalized")
    assert p.main_kernel() != "k_pairs"
    p.close()


def test_file_entry_points(ssu, O, ref_trees, ref_tables, goldens, tmp_path):
    L = ssu.lib()
    tre = tmp_path / "test.tre"
    tre.write_text(ref_trees["test"])
    biom = os.path.join(GOLDEN, "test.biom").encode()
    say("one_off file")
    res = C.POINTER(ssu.Mat)()
    st = L.one_off(biom, str(tre).encode(), b"unweighted", False, C.c_double(1.0), False, 1, C.byref(res))
    assert st == ssu.OKAY, ssu.last_error()
    cf = np.ctypeslib.as_array(res.contents.condensed_form, shape=(res.contents.cf_size,)).copy()
    L.destroy_mat(C.byref(res))
    assert np.max(np.abs(cf - np.array(goldens["condensed_unweighted"]["values"]))) < 1e-5
    say("faith file")
    rv = C.POINTER(ssu.RVec)()
    assert L.faith_pd_one_off(biom, str(tre).encode(), C.byref(rv)) == ssu.OKAY
    vals = np.ctypeslib.as_array(rv.contents.values, shape=(rv.contents.n_samples,)).copy()
    L.destroy_results_vec(C.byref(rv))
    assert np.array_equal(vals, np.array(goldens["faith_test_tre"]["values"], float))
    say("partial file")
    pm = C.POINTER(ssu.PartialMat)()
    assert L.partial_v3(biom, str(tre).encode(), b"unweighted_fp64", False, C.c_double(1.0), False, True, 1, 1, 3,
                        C.byref(pm)) == ssu.OKAY
    g = np.array(goldens["unweighted"]["stripes"])
    for s in range(1, 3):
        row = np.ctypeslib.as_array(pm.contents.stripes[s - 1], shape=(6,))
        assert np.max(np.abs(row - g[s])) < 1e-6
    L.destroy_partial_mat(C.byref(pm))
    say("dense pair")
    d = goldens["dense_pair_unweighted"]
    op = C.c_void_p()
    L.load_bptree_opaque(ref_trees["test"].encode(), C.byref(op))
    ids = (C.c_char_p * 5)(*[x.encode() for x in d["obs_ids"]])
    s1 = (C.c_double * 5)(*d["sample1"])
    s3 = (C.c_double * 5)(*d["sample3"])
    out = C.c_double()
    assert L.one_dense_pair_v2t(5, ids, s1, s3, op, b"unweighted", False, C.c_double(1.0), False,
                                C.byref(out)) == ssu.OKAY
    assert abs(out.value - d["exp13"]) < 1e-5
    L.destroy_bptree_opaque(C.byref(op))


def test_partial_files_merge_to_full_matrix(ssu, ref_trees, tmp_path):
    """partial_v3 on the GPU for two stripe ranges, written with write_partial,
    merged with merge_partial_to_mmap_matrix: equal to one_off_matrix_v3."""
    L = ssu.lib()
    tre = tmp_path / "crawford.tre"
    tre.write_text(ref_trees["crawford"])
    biom = os.path.join(GOLDEN, "crawford.biom").encode()
    say("partial merge crawford")
    full = C.POINTER(ssu.MatFullFP64)()
    assert L.one_off_matrix_v3(biom, str(tre).encode(), b"weighted_normalized_fp64", False, C.c_double(1.0), False,
                               True, 1, 0, True, None, C.byref(full)) == ssu.OKAY, ssu.last_error()
    n = full.contents.n_samples
    ref = np.ctypeslib.as_array(full.contents.matrix, shape=(n * n,)).copy().reshape(n, n)
    L.destroy_mat_full_fp64(C.byref(full))
    total = (n + 1) // 2
    mats = []
    for a, b in ((0, 2), (2, total)):
        pm = C.POINTER(ssu.PartialMat)()
        assert L.partial_v3(biom, str(tre).encode(), b"weighted_normalized_fp64", False, C.c_double(1.0), False, True,
                            1, a, b, C.byref(pm)) == ssu.OKAY
        path = str(tmp_path / f"part{a}.partial").encode()
        assert L.write_partial(path, pm) == ssu.WRITE_OKAY
        L.destroy_partial_mat(C.byref(pm))
        dyn = C.POINTER(ssu.PartialDynMat)()
        assert L.read_partial_header(path, C.byref(dyn)) == ssu.READ_OKAY
        mats.append(dyn)
    arr = (C.POINTER(ssu.PartialDynMat) * 2)(*mats)
    res = C.POINTER(ssu.MatFullFP64)()
    assert L.merge_partial_to_mmap_matrix(arr, 2, str(tmp_path).encode(), C.byref(res)) == ssu.MERGE_OKAY
    got = np.ctypeslib.as_array(res.contents.matrix, shape=(n * n,)).copy().reshape(n, n)
    L.destroy_mat_full_fp64(C.byref(res))
    for d in mats:
        L.destroy_partial_dyn_mat(C.byref(d))
    assert np.array_equal(got, ref)


BLOCKED = [("unweighted", {}, "fp32"), ("unweighted", {}, "fp64"), ("unweighted_unnormalized", {}, "fp32"),
           ("weighted_normalized", {}, "fp32"), ("weighted_unnormalized", {}, "fp32"),
           ("generalized", {"alpha": 0.5}, "fp32"), ("weighted_normalized", {"variance_adjust": True}, "fp32"),
           ("unweighted", {"variance_adjust": True}, "fp32"), ("weighted_normalized", {"normalize": False}, "fp32")]


@pytest.mark.parametrize("method,kw,prec", BLOCKED)
def test_sample_blocked_embedding(ssu, O, synth, monkeypatch, method, kw, prec):
    """The embedding built in 128-sample column blocks (how samples beyond the
    device-memory budget are handled) gives the same matrices."""
    tree, table = _case(synth, 400, 300, 0.05, 131, zipf=1.1)
    monkeypatch.setenv("SSU_EMBED_BLOCK_SAMPLES", "128")
    say("blocked", method, kw, prec)
    s, m = ssu.matrix_inmem(inputs_from(ssu, tree, table), method + ("_fp64" if prec == "fp64" else ""), **kw)
    assert s == ssu.OKAY, ssu.last_error()
    ref = oracle_matrix(O, tree, table, method, **kw)
    assert rel_err(m, ref) <= TOL[prec]


def test_sample_blocked_faith(ssu, O, synth, monkeypatch):
    tree, table = _case(synth, 400, 300, 0.05, 133, zipf=1.1)
    monkeypatch.setenv("SSU_EMBED_BLOCK_SAMPLES", "128")
    say("blocked faith")
    s, vals, _ = ssu.faith_pd_inmem(inputs_from(ssu, tree, table))
    st, le, nm = tree
    assert s == ssu.OKAY and np.array_equal(vals, O.faith_pd(O.Tree(st.astype(np.uint8), le, nm), O.Table(*table)))


def near_duplicate_table(synth, n_base=40, tips=400, seed=95):
    """Samples n_base..2 n_base - 1 repeat 0..n_base - 1 with one count raised
    by 1: distances of 1e-4..1e-2 between each sample and its copy."""
    st, le, nm = synth.random_tree(tips, seed)
    n = 2 * n_base
    obs, sids, ip, ix, d = synth.random_table(tips, n, 0.08, seed + 1)
    dense = np.zeros((len(obs), n))
    for r in range(len(obs)):
        dense[r, ix[ip[r]:ip[r + 1]]] = d[ip[r]:ip[r + 1]]
    dense[:, n_base:] = dense[:, :n_base]
    rng = np.random.default_rng(seed + 2)
    for s in range(n_base, n):
        held = np.flatnonzero(dense[:, s])
        dense[rng.choice(held), s] += 1.0
    rows = [np.flatnonzero(dense[r]) for r in range(len(obs))]
    keep = [r for r in range(len(obs)) if len(rows[r])]
    table = ([obs[r] for r in keep], [f"s{i}" for i in range(n)],
             np.concatenate([[0], np.cumsum([len(rows[r]) for r in keep])]).astype(np.uint32),
             np.concatenate([rows[r] for r in keep]).astype(np.uint32),
             np.concatenate([dense[r, rows[r]] for r in keep]))
    return (st, le, nm), table


@pytest.mark.parametrize("normalize_name", ["weighted_unnormalized", "weighted_normalized"])
def test_vaw_weighted_near_duplicate_samples(ssu, O, synth, normalize_name):
    """VAW weighted on near-duplicate samples (distances 1e-4..1e-2 of their VAW denominator):
    FP32 entries within 1e-6 of the FP64 oracle.  The unnormalised VAW distance has no
    per-sample scale for the refinement's test; the pair kernel marks the pairs below 0.1
    of their own VAW denominator, and every pair of the marked samples is recomputed in FP64."""
    tree, table = near_duplicate_table(synth)
    say("vaw near duplicates", normalize_name)
    inp = inputs_from(ssu, tree, table)
    s, m = ssu.matrix_inmem(inp, normalize_name, variance_adjust=True)
    assert s == ssu.OKAY, ssu.last_error()
    ref = oracle_matrix(O, tree, table, normalize_name, variance_adjust=True)
    err = rel_err(m, ref)
    say("max rel err", err)
    assert err <= TOL["fp32"]
    p = ssu.Plan(inp, normalize_name, variance_adjust=True)
    p.execute()
    assert p.refined_pairs() > 0
    p.close()


@pytest.mark.parametrize("name", ["weighted_normalized", "weighted_unnormalized", "weighted_normalized_fp64",
                                  "weighted_unnormalized_fp64"])
def test_weighted_near_duplicate_samples(ssu, O, synth, name):
    """Near-duplicate samples (distances 1e-4..1e-2): every entry within the
    stated tolerance of the FP64 oracle -- 1e-6 relative for FP32 names, whose
    small distances are recomputed in FP64 (the FP32 rounding of the
    proportions alone is worth up to 2^-24 / d relative), 1e-11 for FP64 names.
    The reference's own FP32 kernels on the same input (oracle/_ref) are
    reported for comparison."""
    tree, table = near_duplicate_table(synth)
    say("near duplicates", name)
    s, m = ssu.matrix_inmem(inputs_from(ssu, tree, table), name)
    assert s == ssu.OKAY, ssu.last_error()
    base = name.replace("_fp64", "")
    ref = oracle_matrix(O, tree, table, base)
    near = np.array([ref[i, i + 40] for i in range(40)])
    assert near.max() < 0.05
    err = rel_err(m, ref)
    say("max rel err", err, "near-duplicate distances", near.min(), near.max())
    if name.endswith("fp64"):
        assert err <= TOL["fp64"]
        return
    assert err <= TOL["fp32"]
    assert O.ref_available(), "the reference build (oracle/_ref) is part of the tree"
    st, le, nm = tree
    T, Tb = O.Tree(np.asarray(st, np.uint8), le, nm), O.Table(*table)
    ref32 = O.stripes_to_matrix(O.ref_stripes(T, Tb, base, fp32=True), 80)
    say("reference FP32 kernels on the same input", rel_err(ref32, ref))


@pytest.mark.parametrize("name,kw", [("weighted_normalized", {}), ("weighted_unnormalized", {}),
                                     ("generalized", {"alpha": 0.5}),
                                     ("generalized", {"alpha": 1.5}),            # dense kernel (alpha > 1)
                                     ("weighted_normalized", {"variance_adjust": True}),
                                     ("generalized", {"alpha": 0.5, "variance_adjust": True})])
def test_near_duplicate_refinement_through_plan(ssu, O, synth, name, kw):
    """The refinement step itself: the plan reports the recomputed pairs,
    every entry is within 1e-6 of the FP64 oracle, and with the step switched
    off (SSU_REFINE=0) the same pairs are off by more -- the FP32 rounding the
    step removes."""
    tree, table = near_duplicate_table(synth)
    say("refinement via plan", name, kw)
    inp = inputs_from(ssu, tree, table)
    ref = oracle_matrix(O, tree, table, name, **kw)
    p = ssu.Plan(inp, name, **kw)
    p.execute()
    m = p.matrix()
    refined = p.refined_pairs()
    p.close()
    err = rel_err(m, ref)
    say("refined pairs", refined, "max rel err", err)
    assert refined >= 40
    assert err <= TOL["fp32"]
    os.environ["SSU_REFINE"] = "0"
    try:
        p = ssu.Plan(inp, name, **kw)
        p.execute()
        m0 = p.matrix()
        assert p.refined_pairs() == 0
        p.close()
    finally:
        del os.environ["SSU_REFINE"]
    err0 = rel_err(m0, ref)
    say("without refinement", err0)
    assert err0 > err


@pytest.mark.parametrize("alpha", [0.0, 0.5, 1.0])
@pytest.mark.parametrize("theta", ["0", "0.1", "0.3"])
def test_generalized_shared_node_path(ssu, O, synth, monkeypatch, theta, alpha):
    """Generalized (FP32, 0 <= alpha <= 1) on the shared-node path: numerator
    and denominator summed directly over the nodes either sample panel holds,
    the rare nodes' shares from the posting lists (theta 0: no split).  Against
    the FP64 oracle and the dense kernel (SSU_GEN_SPARSE=0)."""
    tree, table = _case(synth, 500, 300, 0.05, 75, zipf=1.1)
    monkeypatch.setenv("SSU_SPARSE_SPLIT", theta)
    inp = inputs_from(ssu, tree, table)
    say("generalized shared-node", theta, alpha)
    p = ssu.Plan(inp, "generalized", alpha=alpha)
    assert p.main_kernel() == "k_pairs_gen32"
    p.execute()
    m = p.matrix()
    p.close()
    ref = oracle_matrix(O, tree, table, "generalized", alpha=alpha)
    assert rel_err(m, ref) <= TOL["fp32"]
    assert np.array_equal(m, m.T) and np.all(np.diag(m) == 0)
    monkeypatch.setenv("SSU_GEN_SPARSE", "0")
    s, dense = ssu.matrix_inmem(inp, "generalized", alpha=alpha)
    assert s == ssu.OKAY
    assert rel_err(m, dense) <= 2e-6


@pytest.mark.parametrize("alpha", [0.0, 0.37, 1.0])
@pytest.mark.parametrize("theta", ["0", "0.1", "0.5"])
def test_generalized_fp64_shared_node_path(ssu, O, synth, monkeypatch, theta, alpha):
    """generalized_fp64 (0 <= alpha <= 1) on the shared-node path: FP64
    embedding rows on the 32 x 64 blocks, FP64 terms, the rare nodes' shares
    from the posting lists with FP64 values (theta 0: no split).  Against the
    FP64 oracle and the dense FP64 kernel (SSU_GEN_SPARSE=0), both at 1e-11."""
    tree, table = _case(synth, 500, 300, 0.05, 76, zipf=1.1)
    monkeypatch.setenv("SSU_SPARSE_SPLIT", theta)
    inp = inputs_from(ssu, tree, table)
    say("generalized_fp64 shared-node", theta, alpha)
    p = ssu.Plan(inp, "generalized_fp64", alpha=alpha)
    assert p.main_kernel() == "k_pairs_gen64"
    p.execute()
    m = p.matrix()
    if theta != "0":
        assert p.sparse_terms() > 0
    p.close()
    ref = oracle_matrix(O, tree, table, "generalized", alpha=alpha)
    assert rel_err(m, ref) <= TOL["fp64"]
    assert np.array_equal(m, m.T) and np.all(np.diag(m) == 0)
    monkeypatch.setenv("SSU_GEN_SPARSE", "0")
    s, dense = ssu.matrix_inmem(inp, "generalized_fp64", alpha=alpha, fp32=False)
    assert s == ssu.OKAY
    assert rel_err(m, dense) <= TOL["fp64"]


VAW_SPLIT_METHODS = [("unweighted", {}), ("unweighted_unnormalized", {}), ("weighted_normalized", {}),
                     ("weighted_unnormalized", {}), ("generalized", {"alpha": 0.5})]


@pytest.mark.parametrize("theta", ["0.1", "0.5", "1"])
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("method,kw", VAW_SPLIT_METHODS)
def test_vaw_shared_node_path(ssu, O, synth, monkeypatch, method, kw, prec, theta):
    """VAW on the shared-node path: nodes held by at least theta * n samples
    on the dense pair kernel, the others from the posting lists as one-sided
    sums at the table's distinct totals plus the walk's corrections.  Against
    the oracle at the precision's tolerance and against the dense kernel
    (SSU_VAW_SPLIT=0)."""
    tree, table = _case(synth, 500, 300, 0.05, 81, zipf=1.1)
    _vaw_split_case(ssu, O, monkeypatch, tree, table, method, kw, prec, theta)


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("alpha", [0.0, 1.0])
def test_vaw_generalized_alpha_ends(ssu, O, synth, monkeypatch, alpha, prec):
    """Generalized VAW at alpha 0 and 1 on the shared-node path; alpha 1.5
    keeps the dense kernel (no split)."""
    tree, table = _case(synth, 400, 260, 0.05, 85)
    _vaw_split_case(ssu, O, monkeypatch, tree, table, "generalized", {"alpha": alpha}, prec, "0.5")
    monkeypatch.setenv("SSU_VAW_SPLIT", "0.5")
    p = ssu.Plan(inputs_from(ssu, tree, table), "generalized", variance_adjust=True, alpha=1.5)
    assert p.split_theta() == 0.0
    p.close()


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("method,kw", [("weighted_normalized", {"bypass_tips": True}),
                                       ("weighted_normalized", {"normalize": False}),
                                       ("unweighted", {"bypass_tips": True}),
                                       ("generalized", {"alpha": 0.3, "normalize": False})])
def test_vaw_shared_node_options(ssu, O, synth, monkeypatch, method, kw, prec):
    """VAW on the shared-node path with bypass_tips (tips leave the node set)
    and normalize_sample_counts=False (raw values in place of proportions)."""
    tree, table = _case(synth, 400, 240, 0.05, 87, zipf=1.1)
    _vaw_split_case(ssu, O, monkeypatch, tree, table, method, kw, prec, "0.5")


@pytest.mark.parametrize("n", [2, 3, 65])
def test_vaw_shared_node_small_tables(ssu, O, synth, monkeypatch, n):
    """VAW on the shared-node path at 2, 3 and 65 samples (every node rare at
    theta 1), both precisions."""
    tree, table = _case(synth, 120, n, 0.2, 900 + n)
    for prec in ("fp64", "fp32"):
        for method in ("weighted_normalized", "unweighted_unnormalized"):
            _vaw_split_case(ssu, O, monkeypatch, tree, table, method, {}, prec, "1", walk=n > 3)


@pytest.mark.parametrize("alpha", [-0.5, 1.5, 2.5])
def test_generalized_fp32_alpha_outside_unit_interval(ssu, O, synth, alpha):
    """alpha outside [0, 1] runs the dense kernel, whose FP32 powers keep the exponent of
    u + v exact: exp2(alpha log2 s) alone is off by |alpha log2 s| ulp, up to 2.4e-6 for s
    near 1e-8, and with alpha < 0 the terms of small s weigh the most.  A sparse table (1% of
    3,000 tips per sample, Zipf) gives many such terms; 1e-6 per entry."""
    tree, table = _case(synth, 3000, 300, 0.01, 41, zipf=1.1)
    say("generalized fp32 alpha", alpha)
    s, m = ssu.matrix_inmem(inputs_from(ssu, tree, table), "generalized", alpha=alpha)
    assert s == ssu.OKAY, ssu.last_error()
    err = rel_err(m, oracle_matrix(O, tree, table, "generalized", alpha=alpha))
    say("max rel err", err)
    assert err <= 1e-6


@pytest.mark.parametrize("prec", ["fp32", "fp64"])
@pytest.mark.parametrize("method,kw", [("weighted_normalized", {"variance_adjust": True}),
                                       ("unweighted", {"variance_adjust": True}),
                                       ("generalized", {"alpha": 0.5, "variance_adjust": True})])
def test_vaw_single_tip_samples_on_shared_node_path(ssu, O, synth, monkeypatch, method, kw, prec):
    """1,030 samples of one tip each (40 tips, integer counts): two samples on the same tip
    hold every node with mi = m, so vaw = 0 on all of them and the reference returns 0 / 0.
    The shared-node path forms that pair's shares as one-sided sums less the same terms;
    what the cancellation leaves must come out as 0, so the entry is NaN, as in the oracle."""
    tree = synth.random_tree(40, 47)
    rng = np.random.default_rng(48)
    n = 1030
    tips = [x for x in tree[2] if x]
    own = rng.integers(0, len(tips), n)
    obs = tips
    ip, ix, d = [0], [], []
    for t in range(len(tips)):
        who = np.nonzero(own == t)[0]
        ix.extend(who.tolist()); d.extend(rng.integers(1, 200, len(who)).astype(float).tolist()); ip.append(len(ix))
    table = (obs, [f"s{i}" for i in range(n)], np.array(ip), np.array(ix), np.array(d))
    say("single-tip samples", method, kw, prec)
    monkeypatch.setenv("SSU_VAW_SPLIT", "0.5")   # every tip rare: the pairs on one tip meet in the walk
    p = ssu.Plan(inputs_from(ssu, tree, table), method + ("_fp64" if prec == "fp64" else ""), **kw)
    assert abs(p.split_theta() - 0.5) < 0.01
    p.execute()
    assert p.sparse_terms() > 0
    m = p.matrix()
    p.close()
    ref = oracle_matrix(O, tree, table, method, **kw)
    same = own[:, None] == own[None, :]
    np.fill_diagonal(same, False)
    assert np.all(np.isnan(ref[same])) and np.array_equal(np.isnan(m), np.isnan(ref))
    assert rel_err(m, ref) <= TOL[prec]


VAW_FP32 = [("weighted_normalized", {"variance_adjust": True}), ("weighted_unnormalized", {"variance_adjust": True}),
            ("unweighted", {"variance_adjust": True}), ("generalized", {"alpha": 0.5, "variance_adjust": True})]


@pytest.mark.parametrize("counts", ["fractional", "large integers"])
@pytest.mark.parametrize("method,kw", VAW_FP32)
def test_vaw_fp32_counts_not_exact_in_fp32(ssu, O, synth, counts, method, kw):
    """FP32 VAW divides by sqrt(mi (m - mi)) over node counts and sample totals; with
    counts FP32 cannot hold (fractional, or two totals summing past 2^24) m - mi cancels
    near the root, so such tables compute in FP64 and round to FP32: 1e-6 per entry
    (2.6e-5 before, on the fractional table).  Integer counts keep the FP32 kernels."""
    tree, table = _case(synth, 565, 65, 0.01, 5, zipf=1.1)
    obs, sids, ip, ix, d = table
    rng = np.random.default_rng(7)
    if counts == "fractional":
        d = np.asarray(d, dtype=np.float64) * rng.random(len(d)) + 0.01
    else:
        d = np.asarray(d, dtype=np.float64) * 123457.0   # integers, totals past 2^23
    table = (obs, sids, ip, ix, d)
    inp = inputs_from(ssu, tree, table)
    say("vaw fp32 counts", counts, method, kw)
    s, m = ssu.matrix_inmem(inp, method, **kw)
    assert s == ssu.OKAY, ssu.last_error()
    assert m.dtype == np.float32
    ref = oracle_matrix(O, tree, table, method, **kw)
    assert rel_err(m, ref) <= 1e-6
    p = ssu.Plan(inp, method, **kw)
    assert ssu.lib().ssu_plan_result_bytes(p.h) == 8
    p.close()
    p = ssu.Plan(inputs_from(ssu, tree, (obs, sids, ip, ix, np.rint(np.asarray(table[4]) / 123457.0 + 1.0))),
                 method, **kw)
    assert ssu.lib().ssu_plan_result_bytes(p.h) == 4   # integer counts, small totals: FP32 kernels
    p.close()


def _vaw_split_case(ssu, O, monkeypatch, tree, table, method, kw, prec, theta, walk=True):
    """walk: the samples share some rare node (the posting walk has terms)"""
    monkeypatch.setenv("SSU_VAW_SPLIT", theta)
    inp = inputs_from(ssu, tree, table)
    name = method + ("_fp64" if prec == "fp64" else "")
    say("vaw shared-node", name, kw, theta)
    p = ssu.Plan(inp, name, variance_adjust=True, **kw)
    assert p.main_kernel() == "k_pairs"
    assert abs(p.split_theta() - float(theta)) < 0.01
    p.execute()
    m = p.matrix()
    assert p.sparse_terms() > 0 or not walk
    p.close()
    ref = oracle_matrix(O, tree, table, method, variance_adjust=True, **kw)
    err = rel_err(m, ref)
    say("max rel err", err)
    assert err <= TOL[prec]
    assert np.array_equal(m, m.T) and np.all(np.diag(m) == 0)
    monkeypatch.setenv("SSU_VAW_SPLIT", "0")
    s, dense = ssu.matrix_inmem(inp, name, variance_adjust=True, **kw)
    assert s == ssu.OKAY
    assert rel_err(m, dense) <= 2 * TOL[prec]


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_vaw_shared_node_stripes_and_equal_totals(ssu, O, synth, monkeypatch, prec):
    """VAW on the shared-node path: stripe ranges give the full matrix's
    stripes; a table whose samples all have one total (one distinct total for
    the one-sided sums) and one of near-duplicate samples (where the one-sided
    sums and the walk's corrections cancel almost entirely) against the oracle."""
    monkeypatch.setenv("SSU_VAW_SPLIT", "0.5")
    tree, table = _case(synth, 400, 200, 0.05, 83)
    inp = inputs_from(ssu, tree, table)
    name = "weighted_normalized" + ("_fp64" if prec == "fp64" else "")
    p = ssu.Plan(inp, name, variance_adjust=True)
    p.execute()
    full = p.stripes(0, 100)
    for a, b in ((0, 13), (13, 60), (60, 100)):
        say("vaw stripes", prec, a, b)
        p.execute_stripes(a, b)
        assert np.array_equal(p.stripes(a, b), full[a:b])
    p.close()
    # every sample scaled to total 1000 (integer counts): one distinct total
    obs, sids, ip, ix, d = table
    tot = np.bincount(ix, weights=d, minlength=len(sids))
    d2 = np.maximum(1.0, np.round(d * 1000.0 / tot[ix]))
    tot2 = np.bincount(ix, weights=d2, minlength=len(sids))
    for smp in range(len(sids)):   # the rounding's remainder onto each sample's largest count
        at = np.nonzero(ix == smp)[0]
        d2[at[np.argmax(d2[at])]] += 1000.0 - tot2[smp]
    assert np.all(d2 > 0) and np.all(np.bincount(ix, weights=d2, minlength=len(sids)) == 1000.0)
    table2 = (obs, sids, ip, ix, d2)
    tree3, table3 = near_duplicate_table(synth)
    for case, (tr, tb) in (("one total", (tree, table2)), ("near duplicates", (tree3, table3))):
        for method in ("weighted_normalized", "unweighted"):
            name = method + ("_fp64" if prec == "fp64" else "")
            say("vaw", case, name)
            s, m = ssu.matrix_inmem(inputs_from(ssu, tr, tb), name, variance_adjust=True)
            assert s == ssu.OKAY
            ref = oracle_matrix(O, tr, tb, method, variance_adjust=True)
            err = rel_err(m, ref)
            say("max rel err", err)
            assert err <= TOL[prec]


def test_generalized_near_duplicates(ssu, O, synth):
    """Near-duplicate samples, generalized alpha 0.5: every entry within 1e-6
    of the FP64 oracle (the reference's own FP32 kernels on the same input are
    reported for comparison)."""
    tree, table = near_duplicate_table(synth)
    say("generalized near duplicates")
    s, m = ssu.matrix_inmem(inputs_from(ssu, tree, table), "generalized", alpha=0.5)
    assert s == ssu.OKAY, ssu.last_error()
    ref = oracle_matrix(O, tree, table, "generalized", alpha=0.5)
    err = rel_err(m, ref)
    st, le, nm = tree
    T, Tb = O.Tree(np.asarray(st, np.uint8), le, nm), O.Table(*table)
    ref32 = O.stripes_to_matrix(O.ref_stripes(T, Tb, "generalized", fp32=True, alpha=0.5), 80)
    say("max rel err", err, "reference FP32 kernels", rel_err(ref32, ref))
    assert err <= TOL["fp32"]


def test_near_duplicate_refinement_with_split_and_blocks(ssu, O, synth, monkeypatch):
    """Refinement on a table large enough for the occupancy split (posting
    lists) and with the embedding in 128-sample blocks: 600 samples whose last
    300 repeat the first 300 with one count raised."""
    tree, table = near_duplicate_table(synth, n_base=300, tips=600, seed=41)
    monkeypatch.setenv("SSU_EMBED_BLOCK_SAMPLES", "128")
    say("refinement, split + blocks")
    inp = inputs_from(ssu, tree, table)
    for name, kw in (("weighted_normalized", {}), ("generalized", {"alpha": 0.5})):
        ref = oracle_matrix(O, tree, table, name, **kw)
        p = ssu.Plan(inp, name, **kw)
        p.execute()
        m, refined = p.matrix(), p.refined_pairs()
        p.close()
        err = rel_err(m, ref)
        say(name, "refined pairs", refined, "max rel err", err)
        assert refined >= 300 and err <= TOL["fp32"]