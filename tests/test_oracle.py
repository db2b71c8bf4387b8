"""Pinning the oracle (oracle/ssu_oracle.c) before it is trusted as the checker.

* golden vectors held by the reference's own tests (src/test_su.cpp, test/capi_test.c)
* the reference's CI / validation data (ci/test.ssu.exp, ci/test.faith.exp,
  src/testdata/test.absquant*.dm)
* a pure-Python brute force from the definitions (tests/brute.py)
* closed forms: unit lengths give integer counts, alpha=1 generalized equals
  weighted_normalized, symmetry, zero diagonal, Faith PD = sum of present lengths
* the reference implementation itself (oracle/_ref, built from the reference
  sources in the development container) on seeded random inputs, and the
  vectors it wrote (tests/golden/ref_vectors.npz)
"""
import os

import numpy as np
import pytest

import brute
from conftest import GOLDEN, rel_err, stripes_of


def table_from(d, O):
    return O.Table(d["obs_ids"], d["sample_ids"], d["indptr"], d["indices"], d["data"])


@pytest.fixture(scope="module")
def test_inputs(O, ref_tables, ref_trees):
    return O.newick_to_tree(ref_trees["test"]), table_from(ref_tables["test"], O)


GOLDEN_KEYS = ["unweighted", "unweighted_bypass", "unweighted_unnormalized", "weighted_unnormalized",
               "weighted_normalized", "generalized_a1", "generalized_a0", "generalized_a05",
               "vaw_weighted_normalized"]


@pytest.mark.parametrize("key", GOLDEN_KEYS)
def test_oracle_test_su_goldens(O, goldens, test_inputs, key):
    g = goldens[key]
    tree, table = test_inputs
    got = O.stripes(tree, table, g["method"], variance_adjust=g["vaw"], alpha=g["alpha"], bypass_tips=g["bypass_tips"])
    # src/test_helper.hpp:81-93 compares with absolute tolerance 1e-6
    assert np.max(np.abs(got - np.array(g["stripes"]))) < 1e-6, g["cite"]


def test_oracle_condensed_golden(O, goldens, test_inputs):
    tree, table = test_inputs
    cf = O.stripes_to_condensed(O.stripes(tree, table, "unweighted"), 6)
    assert np.max(np.abs(cf - np.array(goldens["condensed_unweighted"]["values"]))) < 1e-6


@pytest.mark.parametrize("tree_key,gold", [("test", "faith_test_tre"), ("test_faith", "faith_test_faith_tre"),
                                           ("test_faith_shear", "faith_test_faith_shear_tre")])
def test_oracle_faith_goldens(O, goldens, ref_tables, ref_trees, tree_key, gold):
    got = O.faith_pd(O.newick_to_tree(ref_trees[tree_key]), table_from(ref_tables["test"], O))
    assert np.array_equal(got, np.array(goldens[gold]["values"], dtype=float))


def test_oracle_crawford_unweighted_fp64(O, ref_tables, ref_trees, ref_expected):
    """ci/crawford_test.sh:12-20 checks 4 decimals; the oracle matches the file to FP64 rounding."""
    tab = table_from(ref_tables["crawford"], O)
    m = O.matrix(O.newick_to_tree(ref_trees["crawford"]), tab, "unweighted")
    exp = ref_expected["crawford_unweighted_fp64"]
    idx = [tab.sample_ids.index(s) for s in exp["ids"]]
    assert rel_err(m[np.ix_(idx, idx)], np.array(exp["matrix"])) < 1e-12


def test_oracle_crawford_vaw(O, ref_tables, ref_trees, ref_expected):
    """ci/test.ssu_vaw.exp (the reference CI never compares it, ci/crawford_test.sh:22-30)."""
    tab = table_from(ref_tables["crawford"], O)
    m = O.matrix(O.newick_to_tree(ref_trees["crawford"]), tab, "unweighted", variance_adjust=True)
    exp = ref_expected["crawford_unweighted_fp64_vaw"]
    idx = [tab.sample_ids.index(s) for s in exp["ids"]]
    assert rel_err(m[np.ix_(idx, idx)], np.array(exp["matrix"])) < 1e-12


def test_oracle_crawford_faith_exact(O, ref_tables, ref_trees, ref_expected):
    """ci/crawford_test.sh:40-44 compares by md5, i.e. all 16 printed digits."""
    tab = table_from(ref_tables["crawford"], O)
    got = O.faith_pd(O.newick_to_tree(ref_trees["crawford"]), tab)
    exp = dict(ref_expected["crawford_faith"])
    for sid, v in zip(tab.sample_ids, got):
        assert "%.16g" % v == exp[sid]


@pytest.mark.parametrize("normalize,key", [(False, "absquant_wn_raw"), (True, "absquant_wn_normalized")])
def test_oracle_absquant(O, ref_tables, ref_trees, ref_expected, normalize, key):
    """src/testdata/validation_tests.sh:135-148 (tolerance 1e-5)."""
    tab = table_from(ref_tables["absquant"], O)
    m = O.matrix(O.newick_to_tree(ref_trees["absquant"]), tab, "weighted_normalized", normalize=normalize)
    exp = ref_expected[key]
    idx = [tab.sample_ids.index(s) for s in exp["ids"]]
    assert np.max(np.abs(m[np.ix_(idx, idx)] - np.array(exp["matrix"]))) < 1e-5


def _tiny(seed, synth, n_tips=14, n_samples=6, density=0.35):
    st, le, nm = synth.random_tree(n_tips, seed)
    tab = synth.random_table(n_tips, n_samples, density, seed + 7)
    return st, le, nm, tab


def _brute_table(tab):
    obs_ids, sample_ids, indptr, indices, data = tab
    t = {s: {} for s in sample_ids}
    for i, o in enumerate(obs_ids):
        for j in range(indptr[i], indptr[i + 1]):
            t[sample_ids[indices[j]]][o] = float(data[j])
    return t


BRUTE_CASES = [("unweighted", {}), ("unweighted_unnormalized", {}), ("weighted_normalized", {}),
               ("weighted_unnormalized", {}), ("generalized", {"alpha": 0.5}), ("generalized", {"alpha": 0.0}),
               ("unweighted", {"vaw": True}), ("weighted_normalized", {"vaw": True}),
               ("weighted_unnormalized", {"vaw": True}), ("generalized", {"alpha": 0.5, "vaw": True}),
               ("unweighted_unnormalized", {"vaw": True}), ("unweighted", {"bypass_tips": True}),
               ("weighted_normalized", {"bypass_tips": True}), ("weighted_normalized", {"normalize": False})]


@pytest.mark.parametrize("seed", [1, 2, 3])
@pytest.mark.parametrize("method,kw", BRUTE_CASES)
def test_oracle_vs_bruteforce(O, synth, seed, method, kw):
    st, le, nm, tab = _tiny(seed, synth)
    got = O.matrix(O.Tree(st.astype(np.uint8), le, nm), O.Table(*tab), method,
                   variance_adjust=kw.get("vaw", False), alpha=kw.get("alpha", 1.0),
                   bypass_tips=kw.get("bypass_tips", False), normalize=kw.get("normalize", True))
    exp = brute.distances(st, le, nm, _brute_table(tab), method, alpha=kw.get("alpha", 1.0),
                          vaw=kw.get("vaw", False), bypass_tips=kw.get("bypass_tips", False),
                          normalize=kw.get("normalize", True))
    for (a, b), v in exp.items():
        assert got[a, b] == pytest.approx(v, rel=1e-12, abs=1e-14, nan_ok=True), (a, b)


@pytest.mark.parametrize("seed", [4, 5])
def test_oracle_faith_vs_bruteforce(O, synth, seed):
    st, le, nm, tab = _tiny(seed, synth, n_tips=20, n_samples=5)
    got = O.faith_pd(O.Tree(st.astype(np.uint8), le, nm), O.Table(*tab))
    exp = brute.faith_pd(st, le, nm, _brute_table(tab))
    np.testing.assert_allclose(got, exp, rtol=1e-13)


def test_oracle_unit_lengths_integer_counts(O, synth):
    st, le, nm, tab = _tiny(9, synth, n_tips=40, n_samples=10, density=0.2)
    ones = np.where(st, 1.0, 0.0)
    tree = O.Tree(st.astype(np.uint8), ones, nm)
    m = O.matrix(tree, O.Table(*tab), "unweighted_unnormalized")
    assert np.array_equal(m, np.round(m))
    # Faith PD with unit lengths = number of present non-root nodes: an integer
    f = O.faith_pd(tree, O.Table(*tab))
    assert np.array_equal(f, np.round(f))


def test_oracle_closed_forms(O, synth):
    st, le, nm, tab = _tiny(11, synth, n_tips=60, n_samples=12, density=0.15)
    tree, table = O.Tree(st.astype(np.uint8), le, nm), O.Table(*tab)
    wn = O.matrix(tree, table, "weighted_normalized")
    g1 = O.matrix(tree, table, "generalized", alpha=1.0)
    np.testing.assert_allclose(g1, wn, rtol=1e-13, atol=1e-15)
    for m in (wn, O.matrix(tree, table, "unweighted")):
        assert np.array_equal(m, m.T) and np.all(np.diag(m) == 0)
        assert np.all((m >= 0) & (m <= 1))
    # XOR = T_i + T_j - 2 G_ij with T = Faith PD: unweighted_unnormalized from Faith sums
    uu = O.matrix(tree, table, "unweighted_unnormalized")
    f = O.faith_pd(tree, table)
    u = O.matrix(tree, table, "unweighted")
    # d = XOR / OR and OR = (T_i + T_j + XOR) / 2
    orr = (f[:, None] + f[None, :] + uu) / 2
    off = ~np.eye(len(f), dtype=bool)
    np.testing.assert_allclose(u[off], (uu / orr)[off], rtol=1e-12)


# ---------------------------------------------------------------- vs the reference itself

def test_oracle_vs_committed_reference_vectors(O):
    """tests/golden/ref_vectors.npz was written by the reference implementation (oracle/_ref)."""
    path = os.path.join(GOLDEN, "ref_vectors.npz")
    z = np.load(path, allow_pickle=False)
    cases = [k[:-len("__out")] for k in z.files if k.endswith("__out") and "__fp64" in k]
    assert cases
    for c in cases:
        tree = O.Tree(z[c + "__structure"], z[c + "__lengths"], list(z[c + "__names"]))
        table = O.Table(list(z[c + "__obs_ids"]), list(z[c + "__sample_ids"]), z[c + "__indptr"],
                        z[c + "__indices"], z[c + "__data"])
        meta = z[c + "__meta"]
        method = str(z[c + "__method"])
        got = O.stripes(tree, table, method, variance_adjust=bool(meta[0]), alpha=float(meta[1]),
                        bypass_tips=bool(meta[2]), normalize=bool(meta[3]))
        assert rel_err(got, z[c + "__out"]) < 1e-11, c


@pytest.mark.skipif(not os.path.exists(os.path.join(os.path.dirname(GOLDEN), "..", "oracle", "_ref",
                                                    "libssu_ref.so")), reason="reference build not present")
@pytest.mark.parametrize("method,kw", BRUTE_CASES)
def test_oracle_vs_reference_live(O, synth, method, kw):
    st, le, nm = synth.random_tree(300, 31)
    tab = O.Table(*synth.random_table(300, 37, 0.08, 32))
    tree = O.Tree(st.astype(np.uint8), le, nm)
    a = O.stripes(tree, tab, method, variance_adjust=kw.get("vaw", False), alpha=kw.get("alpha", 1.0),
                  bypass_tips=kw.get("bypass_tips", False), normalize=kw.get("normalize", True))
    b = O.ref_stripes(tree, tab, method, variance_adjust=kw.get("vaw", False), alpha=kw.get("alpha", 1.0),
                      bypass_tips=kw.get("bypass_tips", False), normalize=kw.get("normalize", True))
    assert rel_err(a, b) < 1e-11
    if method == "unweighted":
        np.testing.assert_array_equal(O.faith_pd(tree, tab), O.ref_faith_pd(tree, tab))


# --------------------------------------- distance-matrix statistics (oracle)

def _ska():
    import json
    from conftest import GOLDEN
    return json.load(open(os.path.join(GOLDEN, "ska_vectors.json")))


def test_oracle_centering_matches_reference_vectors(O):
    """src/test_ska.cpp:7-60 (6 x 6) and the centred crawford matrix of :62-140."""
    v = _ska()
    c = v["center_mat"]
    m = np.array(c["matrix"]).reshape(6, 6)
    assert np.max(np.abs(O.center_matrix(m) - np.array(c["expected"]).reshape(6, 6))) < 1e-6
    p = v["pcoa"]
    m = np.array(p["matrix"]).reshape(9, 9)
    assert np.max(np.abs(O.center_matrix(m) - np.array(p["centered"]).reshape(9, 9))) < 1e-5


def test_oracle_pcoa_matches_reference_vectors(O):
    """src/test_ska.cpp:142-245: eigenvalues, coordinates (up to sign) and
    proportion explained of the 9-sample crawford matrix, 5 dimensions."""
    p = _ska()["pcoa"]
    m = np.array(p["matrix"]).reshape(9, 9)
    w, coords, prop = O.pcoa_exact(m, 5)
    assert np.max(np.abs(w - np.array(p["eigenvalues"]))) < 1e-5
    assert np.max(np.abs(np.abs(coords) - np.abs(np.array(p["samples"]).reshape(9, 5)))) < 1e-5
    assert np.max(np.abs(prop - np.array(p["proportion"]))) < 1e-5
    # the coordinates are the unit eigenvectors scaled by sqrt(eigenvalue)
    ev = np.array(p["eigenvectors"]).reshape(9, 5)
    assert np.max(np.abs(np.abs(ev * np.sqrt(w)) - np.abs(coords))) < 1e-5


@pytest.mark.parametrize("key", ["ties", "noties", "unequal"])
def test_oracle_permanova_f_matches_reference_vectors(O, key):
    """src/test_ska.cpp test_permanova_*: pseudo-F of skbio's examples, both
    groupings and both the FP64 (asymmetric in the noties case: only the upper
    triangle counts) and the FP32 matrices."""
    v = _ska()[f"permanova_{key}"]
    n = len(v["grouping"])
    for mk in ("matrix_fp64", "matrix_fp32"):
        m = np.array(v[mk]).reshape(n, n)
        for gk in ("grouping", "grouping_swapped"):
            assert abs(O.permanova_f(m, v[gk]) - v["stat"]) < 1e-5


def test_oracle_permanova_f_closed_form(O):
    """Two groups of two with within-group distance a and between-group b:
    s_T = (2a^2 + 4b^2)/4, s_W = a^2, F = (s_T - s_W) / (s_W / 2)."""
    a, b = 1.5, 4.0
    m = np.array([[0, a, b, b], [a, 0, b, b], [b, b, 0, a], [b, b, a, 0]], float)
    s_t = (2 * a * a + 4 * b * b) / 4
    s_w = a * a
    assert abs(O.permanova_f(m, [0, 0, 1, 1]) - (s_t - s_w) / (s_w / 2)) < 1e-12
