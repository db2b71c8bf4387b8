"""CPU-only checks of libssu.so: the exported C ABI, host-side tree handling,
and the API's argument validation (status codes that return before any
device work)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "ssu.h")


def declared_symbols():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"SSU_EXTERN\s+[\w\s\*]+?\b(\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol(ssu):
    L = ssu.lib()
    names = declared_symbols()
    assert len(names) >= 55
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    # the Python binding covers the same list
    assert set(names) == set(ssu.EXPORTS)


LAYOUT_PROBE = r"""
#include <stdio.h>
#include <stddef.h>
#include HEADER
#define P(T, f) printf(#T "." #f " %zu\n", offsetof(T, f));
#define S(T) printf(#T " %zu\n", sizeof(T));
int main(void) {
  S(mat_t) P(mat_t, cf_size) P(mat_t, is_upper_triangle) P(mat_t, condensed_form) P(mat_t, sample_ids)
  S(mat_full_fp64_t) P(mat_full_fp64_t, flags) P(mat_full_fp64_t, matrix) P(mat_full_fp64_t, sample_ids)
  S(mat_full_fp32_t) P(mat_full_fp32_t, matrix)
  S(r_vec) P(r_vec, values) P(r_vec, sample_ids)
  S(partial_mat_t) P(partial_mat_t, stripes) P(partial_mat_t, stripe_start) P(partial_mat_t, stripe_total)
  P(partial_mat_t, is_upper_triangle)
  S(partial_dyn_mat_t) P(partial_dyn_mat_t, offsets) P(partial_dyn_mat_t, is_upper_triangle) P(partial_dyn_mat_t, filename)
  S(support_biom_t) P(support_biom_t, indices) P(support_biom_t, data) P(support_biom_t, n_obs) P(support_biom_t, nnz)
  S(support_bptree_t) P(support_bptree_t, names) P(support_bptree_t, n_parens)
  printf("enum %d %d %d %d\n", (int)invalid_method, (int)grouping_missing, (int)write_error, (int)stripes_overlap);
  return 0;
}
"""


def _layout(tmp_path, header, extra_inc):
    import subprocess
    src = tmp_path / f"probe_{abs(hash(header))}.c"
    exe = tmp_path / f"probe_{abs(hash(header))}"
    src.write_text(LAYOUT_PROBE.replace("HEADER", f'"{header}"'))
    subprocess.run(["gcc", "-std=c11", "-I", extra_inc, "-o", str(exe), str(src)], check=True)
    return subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout


def test_struct_layouts_match_ctypes_binding(ssu, tmp_path):
    out = dict(line.rsplit(" ", 1) for line in _layout(tmp_path, HEADER, os.path.dirname(HEADER)).splitlines()
               if not line.startswith("enum"))
    assert int(out["support_biom_t"]) == C.sizeof(ssu.SupportBiom)
    assert int(out["support_bptree_t"]) == C.sizeof(ssu.SupportBPTree)
    assert int(out["mat_full_fp64_t"]) == C.sizeof(ssu.MatFullFP64)
    assert int(out["partial_mat_t.is_upper_triangle"]) == ssu.PartialMat.is_upper_triangle.offset


@pytest.mark.skipif(not os.path.exists("/root/reference/src/api.hpp"), reason="reference header not present")
def test_struct_layouts_match_reference_header(tmp_path):
    """Same offsets, sizes and enum values as the reference's src/api.hpp compiled as C."""
    mine = _layout(tmp_path, HEADER, os.path.dirname(HEADER))
    theirs = _layout(tmp_path, "/root/reference/src/api.hpp", "/root/reference/src")
    assert mine == theirs


NEWICKS = [
    "(GG_OTU_1:1,(GG_OTU_2:1,GG_OTU_3:1):1,(GG_OTU_5:1,GG_OTU_4:1):1);",
    "((a:0.1,b:0.2)x:0.3,(c:0.5)y:0.4,d:1e-3)root;",                       # single descendant, internal names
    "('quoted, name':1,'with (parens)':2,(e:3,f:4):5);",                   # quoting
    "((a:1,b:2):3,\n(c:4,d:5):6);",                                        # embedded newline
    "(((a:1)b:2)c:3,d:4);",                                                # chain of single descendants
    "(a,b,(c,d));",                                                        # no lengths
    "(a:0.123456789012,b:3.4028235e38,c:1e-30);",                          # float rounding of lengths
]
# std::stof throws on a length that underflows float (src/tree.cpp:446-449), which aborts the
# reference; this parser keeps strtof's value, so that case is compared with the oracle only.
NEWICK_UNDERFLOW = "(a:1,b:1e-40);"


@pytest.mark.parametrize("nw", NEWICKS + [NEWICK_UNDERFLOW])
def test_newick_parse_matches_oracle(ssu, O, nw):
    s, le, nm = ssu.newick_to_bp(nw)
    t = O.newick_to_tree(nw)
    assert np.array_equal(s, t.structure.astype(bool))
    assert np.array_equal(le, t.lengths) or np.allclose(le, t.lengths, rtol=0, atol=0, equal_nan=True)
    assert nm == t.names


@pytest.mark.skipif(not os.path.exists(os.path.join(ROOT, "oracle", "_ref", "libssu_ref.so")),
                    reason="reference build not present")
@pytest.mark.parametrize("nw", NEWICKS)
def test_newick_parse_matches_reference(ssu, O, nw):
    s, le, nm = ssu.newick_to_bp(nw)
    r = O.ref_newick_to_tree(nw)
    assert np.array_equal(s, r.structure.astype(bool))
    assert np.array_equal(le, r.lengths)
    assert nm == r.names


@pytest.mark.parametrize("seed,tips,keep", [(1, 50, 0.3), (2, 200, 0.9), (3, 500, 0.05), (4, 2000, 0.5)])
def test_postorder_matches_oracle(ssu, O, synth, seed, tips, keep):
    st, le, nm = synth.random_tree(tips, seed)
    rng = np.random.default_rng(seed)
    tipnames = [n for n in nm if n]
    obs = list(rng.choice(tipnames, max(1, int(keep * len(tipnames))), replace=False))
    mine = ssu.postorder(st, le, nm, obs)
    tab = O.Table(obs, ["s0"], [0] * (len(obs) + 1), [], [])
    theirs = O.postorder(O.Tree(st.astype(np.uint8), le, nm), tab)
    for a, b in zip(mine, theirs):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("shape", ["multifurcating", "caterpillar"])
@pytest.mark.parametrize("seed,keep", [(7, 0.05), (8, 0.4), (9, 1.0)])
def test_postorder_shapes_match_oracle(ssu, O, synth, shape, seed, keep):
    """Shear + collapse + postorder over multifurcating trees (nodes of three
    or more children; shearing leaves chains of one-child nodes to collapse)
    and a caterpillar tree, against the oracle's restatement."""
    if shape == "multifurcating":
        st, le, nm = synth.contract_tree(synth.random_tree(1500, seed), 0.4, seed + 1)
    else:
        st, le, nm = synth.caterpillar_tree(3000, seed)
    rng = np.random.default_rng(seed)
    tipnames = [n for n in nm if n]
    obs = list(rng.choice(tipnames, max(1, int(keep * len(tipnames))), replace=False)) + ["not-a-tip"]
    mine = ssu.postorder(st, le, nm, obs)
    tab = O.Table(obs, ["s0"], [0] * (len(obs) + 1), [], [])
    theirs = O.postorder(O.Tree(st.astype(np.uint8), le, nm), tab)
    for a, b in zip(mine, theirs):
        assert np.array_equal(a, b)


def test_postorder_large_threaded_matches_oracle(ssu, O, synth):
    """Trees of 65,536 or more parentheses and tables of 65,536 or more ids
    take the threaded name index and leaf lookups; repeated ids keep their
    last row."""
    st, le, nm = synth.random_tree(70000, 11)
    rng = np.random.default_rng(11)
    tipnames = [n for n in nm if n]
    obs = list(rng.choice(tipnames, 66000, replace=False)) + ["not-a-tip"]
    tab = O.Table(obs, ["s0"], [0] * (len(obs) + 1), [], [])
    theirs = O.postorder(O.Tree(st.astype(np.uint8), le, nm), tab)
    for a, b in zip(ssu.postorder(st, le, nm, obs), theirs):
        assert np.array_equal(a, b)
    # an id given twice maps to its last row, as the reference's id index does
    # (src/biom_inmem.cpp:432-439, map[id] = row in row order)
    rep = obs + obs[:50]
    last = {o: i for i, o in enumerate(rep)}
    mine = ssu.postorder(st, le, nm, rep)
    for a, b in zip(mine[:3], theirs[:3]):
        assert np.array_equal(a, b)
    want = np.array([last[obs[r]] if r >= 0 else -1 for r in theirs[3]])
    assert np.array_equal(mine[3], want) and (want >= len(obs)).sum() > 0


def test_postorder_concurrent_calls(ssu, synth):
    """Eight threads walking different trees at once (each call takes its own set of the
    passes' reused scratch arrays) give what each gives alone."""
    from concurrent.futures import ThreadPoolExecutor
    cases = []
    for seed in range(8):
        st, le, nm = synth.random_tree(3000 + 4000 * seed, 40 + seed)
        tips = [x for x in nm if x]
        rng = np.random.default_rng(seed)
        obs = list(rng.choice(tips, len(tips) // (seed + 2), replace=False))
        cases.append((st, le, nm, obs))
    alone = [ssu.postorder(*c) for c in cases]
    with ThreadPoolExecutor(8) as ex:
        for rep in range(3):
            together = list(ex.map(lambda c: ssu.postorder(*c), cases))
            for a, b in zip(alone, together):
                assert all(np.array_equal(x, y) for x, y in zip(a, b))


@pytest.mark.skipif(not os.path.exists(os.path.join(ROOT, "oracle", "_ref", "libssu_ref.so")),
                    reason="reference build not present")
@pytest.mark.parametrize("seed,tips", [(5, 300), (6, 3000)])
def test_postorder_matches_reference(ssu, O, synth, seed, tips):
    st, le, nm = synth.random_tree(tips, seed)
    tab = synth.random_table(tips, 20, 0.05, seed + 1)
    mine = ssu.postorder(st, le, nm, tab[0])
    theirs = O.ref_postorder(O.Tree(st.astype(np.uint8), le, nm), O.Table(*tab))
    for a, b in zip(mine, theirs):
        assert np.array_equal(a, b)


def _fixture_inputs(ssu, ref_tables, ref_trees, key="test", tree_key=None):
    t = ref_tables[key]
    s, le, nm = ssu.newick_to_bp(ref_trees[tree_key or key])
    return ssu.Inputs(s, le, nm, t["obs_ids"], t["sample_ids"], t["indptr"], t["indices"], t["data"])


def test_api_status_codes_without_device(ssu, ref_tables, ref_trees):
    """src/api.cpp:775-856: validation happens before any compute."""
    inp = _fixture_inputs(ssu, ref_tables, ref_trees)
    assert ssu.matrix_inmem(inp, "unweighted_fp64", fp32=True)[0] == ssu.INVALID_METHOD
    assert ssu.matrix_inmem(inp, "unweighted", fp32=False)[0] == ssu.INVALID_METHOD
    assert ssu.matrix_inmem(inp, "not_a_method", fp32=False)[0] == ssu.UNKNOWN_METHOD
    # a table observation missing from the tree
    t = ref_tables["test"]
    s, le, nm = ssu.newick_to_bp("(GG_OTU_1:1,GG_OTU_2:1);")
    bad = ssu.Inputs(s, le, nm, t["obs_ids"], t["sample_ids"], t["indptr"], t["indices"], t["data"])
    assert ssu.matrix_inmem(bad, "unweighted_fp64")[0] == ssu.TABLE_AND_TREE_DO_NOT_OVERLAP
    empty = ssu.Inputs(s, le, nm, [], [], [0], [], [])
    assert ssu.matrix_inmem(empty, "unweighted_fp64")[0] == ssu.TABLE_EMPTY


def test_file_api_missing_inputs(ssu, tmp_path):
    L = ssu.lib()
    res = C.c_void_p()
    missing = str(tmp_path / "nope.biom").encode()
    tre = tmp_path / "t.tre"
    tre.write_text("(a:1,b:1);")
    assert L.one_off(missing, str(tre).encode(), b"unweighted", False, C.c_double(1.0), False, 1,
                     C.byref(res)) == ssu.TABLE_MISSING
    assert L.faith_pd_one_off(missing, str(tre).encode(), C.byref(res)) == ssu.TABLE_MISSING
    op = C.c_void_p()
    assert L.read_bptree_opaque(str(tmp_path / "none.tre").encode(), C.byref(op)) == ssu.OPEN_ERROR
    assert L.read_bptree_opaque(str(tre).encode(), C.byref(op)) == ssu.READ_OKAY
    assert L.get_bptree_opaque_els(op) == 2
    L.destroy_bptree_opaque(C.byref(op))
    assert not op.value


def test_text_writers_match_reference_format(ssu, tmp_path):
    """write_mat_from_matrix (src/api.cpp:1640-1775) uses std::to_chars fixed; write_vec %.16g lines."""
    L = ssu.lib()
    n = 3
    vals = (C.c_double * 9)(0.0, 0.2, 0.5714285714285714, 0.2, 0.0, 1e-7, 0.5714285714285714, 1e-7, 0.0)
    ids = (C.c_char_p * 3)(b"A", b"B", b"C")
    m = ssu.MatFullFP64(n, 0, C.cast(vals, C.POINTER(C.c_double)), ids)
    out = tmp_path / "m.txt"
    assert L.write_mat_from_matrix(str(out).encode(), C.byref(m)) == ssu.WRITE_OKAY
    lines = out.read_text().splitlines()
    assert lines[0] == "\tA\tB\tC"
    assert lines[1] == "A\t0\t0.2\t0.5714285714285714"
    assert lines[2] == "B\t0.2\t0\t0.0000001"


def test_split_model_choices(ssu):
    """The occupancy-split model: rare nodes go to posting lists, common ones
    stay on the tiles; the choice never models slower than no split, and a
    heavier tile term (generalized) moves the threshold up."""
    n = 16384
    assert ssu.choose_split(0, n, np.full(1000, n))[0] == 0              # every node common: no split
    thr, ms, ms0 = ssu.choose_split(0, n, np.full(1000, 3))
    assert thr > 3 and ms < ms0                                           # every node rare: all to postings
    rng = np.random.default_rng(1)
    occ = np.minimum(n, (rng.pareto(1.2, 50000) * 40).astype(np.int64) + 1)
    got = {k: ssu.choose_split(k, n, occ) for k in range(4)}
    for k, (thr, ms, ms0) in got.items():
        assert 0 < thr <= n and ms <= ms0
    assert got[2][0] > got[0][0]                                          # generalized splits higher
    assert got[3][0] < got[0][0]                                          # the Gram column is cheap
    # VAW (5 FP32, 6 FP64): one-sided terms priced at n totals; the FP64 dense term
    # costs more, which moves the threshold up; unknown kinds choose no split
    n8 = 8192
    occ8 = np.minimum(n8, (rng.pareto(1.2, 50000) * 20).astype(np.int64) + 1)
    v32, v64 = ssu.choose_split(5, n8, occ8), ssu.choose_split(6, n8, occ8)
    assert 0 < v32[0] <= n8 and v32[1] < v32[2] and 0 < v64[0] <= n8 and v64[1] < v64[2]
    assert v64[0] >= v32[0]
    assert ssu.choose_split(7, n8, occ8)[0] == 0
