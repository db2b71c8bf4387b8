"""Stripe ranges (src/api.cpp:447-491 set_tasks, :527-567 partial_v3): a
range computes only the pairs of its stripes -- the Gram tiles, pair blocks
and posting-list columns holding one -- and its stripes are bit-identical to
the same stripes of the full matrix, for every kernel path (exact Gram with
the occupancy split, FP32 and FP64 shared-node kernels with posting lists,
generalized on the shared-node path, the dense pair tiles), with near-duplicate
refinement in play, and through the partial_v3 / write_partial /
merge_partial_to_mmap_matrix file path."""
import ctypes as C
import os
import sys

import numpy as np
import pytest

from conftest import GOLDEN, stripes_of

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-case]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def _need_device(ssu):
    if not ssu.device_available():
        pytest.fail("no HIP device visible: the HIP path cannot run (no CPU fallback exists)")


def ranges(total, k=8, seed=3):
    """k consecutive, uneven stripe ranges covering [0, total)"""
    rng = np.random.default_rng(seed)
    cuts = np.sort(rng.choice(np.arange(1, total), size=k - 1, replace=False))
    b = [0, *cuts.tolist(), total]
    return list(zip(b[:-1], b[1:]))


PATHS = [("unweighted", {}), ("unweighted_fp64", {}), ("unweighted_unnormalized", {}),
         ("weighted_normalized", {}), ("weighted_unnormalized", {}), ("weighted_normalized_fp64", {}),
         ("generalized", {"alpha": 0.5}), ("generalized", {"alpha": 1.5}), ("generalized_fp64", {"alpha": 0.5}),
         ("weighted_normalized", {"variance_adjust": True})]


@pytest.fixture(scope="module")
def case(ssu, synth):
    # 1,100 samples: the Gram path splits (from 1,024 samples), every weighted path too
    tree = synth.random_tree(700, 211)
    table = synth.random_table(700, 1100, 0.05, 212, zipf_s=1.1)
    return ssu.Inputs(*tree, *table)


@pytest.mark.parametrize("name,kw", PATHS, ids=[f"{m}{'-' + '-'.join(f'{k}{v}' for k, v in kw.items()) if kw else ''}"
                                                  for m, kw in PATHS])
def test_stripe_ranges_equal_full_matrix(ssu, case, name, kw):
    say("stripe ranges", name, kw)
    p = ssu.Plan(case, name, **kw)
    p.execute()
    n = case.n_samples
    total = (n + 1) // 2
    full = stripes_of(p.matrix().astype(np.float64))
    for a, b in ranges(total):
        p.execute_stripes(a, b)
        got = p.stripes(a, b)
        assert np.array_equal(got, full[a:b]), (name, a, b, float(np.nanmax(np.abs(got - full[a:b]))))
    p.execute_stripes(0, 0)   # every pair again
    assert np.array_equal(stripes_of(p.matrix().astype(np.float64)), full)
    p.close()


def test_stripe_range_work_scales(ssu, case):
    """A range computes only its share: the Gram tiles, pair blocks and posting
    terms of [0, S/8) are well below those of the full matrix."""
    n = case.n_samples
    total = (n + 1) // 2
    for name in ("unweighted", "weighted_normalized"):
        p = ssu.Plan(case, name)
        p.execute()
        kt, st = p.kernel_terms(), p.sparse_terms()
        p.execute_stripes(0, total // 8)
        kt8, st8 = p.kernel_terms(), p.sparse_terms()
        say(name, "kernel terms", kt, "->", kt8, "posting terms", st, "->", st8)
        assert kt8 < 0.45 * kt
        assert st8 < 0.45 * st
        p.close()


def test_stripe_ranges_with_refinement(ssu, O, synth):
    """Near-duplicate samples (FP64 refinement of small distances): every range's
    stripes still equal the full matrix's, bit for bit -- a pair is refined to
    the same value whichever other samples a range marks."""
    sys.path.insert(0, os.path.dirname(__file__))
    from test_gpu_parity import near_duplicate_table
    tree, table = near_duplicate_table(synth, n_base=200, tips=500, seed=17)
    inp = ssu.Inputs(*tree, *table)
    n = len(table[1])
    total = (n + 1) // 2
    for name, kw in (("weighted_normalized", {}), ("weighted_unnormalized", {}), ("generalized", {"alpha": 0.5})):
        say("stripe ranges with refinement", name)
        p = ssu.Plan(inp, name, **kw)
        p.execute()
        assert p.refined_pairs() >= 200
        full = stripes_of(p.matrix().astype(np.float64))
        for a, b in ranges(total, 5, seed=11):
            p.execute_stripes(a, b)
            assert np.array_equal(p.stripes(a, b), full[a:b]), (name, a, b)
        p.close()


def test_partial_files_merge_bit_equal_to_one_off(ssu, synth, tmp_path):
    """Eight partial_v3 calls over files, written with write_partial and merged
    with merge_partial_to_mmap_matrix_fp32: bit-equal to one_off_matrix_fp32_v3."""
    sys.path.insert(0, os.path.dirname(__file__))
    from test_multi_files import write_biom
    L = ssu.lib()
    tree = synth.random_tree(400, 77)
    table = synth.random_table(400, 1060, 0.06, 78)
    tre = tmp_path / "t.tre"
    tre.write_text(synth.bp_to_newick(*tree))
    biom = tmp_path / "t.biom"
    write_biom(str(biom), *table)   # skips where no h5py interpreter writes the file
    full = C.POINTER(ssu.MatFullFP32)()
    assert L.one_off_matrix_fp32_v3(str(biom).encode(), str(tre).encode(), b"weighted_normalized", False,
                                    C.c_double(1.0), False, True, 1, 0, True, None, C.byref(full)) == ssu.OKAY, \
        ssu.last_error()
    n = full.contents.n_samples
    ref = np.ctypeslib.as_array(full.contents.matrix, shape=(n * n,)).copy().reshape(n, n)
    L.destroy_mat_full_fp32(C.byref(full))
    mats = []
    for a, b in ranges((n + 1) // 2):
        say("partial", a, b)
        pm = C.POINTER(ssu.PartialMat)()
        assert L.partial_v3(str(biom).encode(), str(tre).encode(), b"weighted_normalized", False, C.c_double(1.0),
                            False, True, 1, a, b, C.byref(pm)) == ssu.OKAY, ssu.last_error()
        path = str(tmp_path / f"p{a}.partial").encode()
        assert L.write_partial(path, pm) == ssu.WRITE_OKAY
        L.destroy_partial_mat(C.byref(pm))
        dyn = C.POINTER(ssu.PartialDynMat)()
        assert L.read_partial_header(path, C.byref(dyn)) == ssu.READ_OKAY
        mats.append(dyn)
    arr = (C.POINTER(ssu.PartialDynMat) * len(mats))(*mats)
    res = C.POINTER(ssu.MatFullFP32)()
    assert L.merge_partial_to_mmap_matrix_fp32(arr, len(mats), str(tmp_path).encode(), C.byref(res)) == \
        ssu.MERGE_OKAY
    got = np.ctypeslib.as_array(res.contents.matrix, shape=(n * n,)).copy().reshape(n, n)
    L.destroy_mat_full_fp32(C.byref(res))
    for dm in mats:
        L.destroy_partial_dyn_mat(C.byref(dm))
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("name,kw", [("weighted_normalized", {}), ("weighted_normalized_fp64", {}),
                                     ("generalized", {"alpha": 0.5}), ("generalized_fp64", {"alpha": 0.5}),
                                     ("weighted_unnormalized", {}),
                                     ("unweighted", {}), ("unweighted_fp64", {}), ("unweighted_unnormalized", {})])
def test_one_off_banded_copy_equals_plan(ssu, synth, name, kw):
    """one_off_matrix_inmem* copies each finished band of rows while the rows
    below compute (weighted: the tiled kernel's bands alternating between two
    streams; unweighted: Gram tiles, posting walk and combine per band); with
    near-duplicate samples refined afterwards (the copy is redone): the matrix
    equals a resident plan's, bit for bit."""
    sys.path.insert(0, os.path.dirname(__file__))
    from test_gpu_parity import near_duplicate_table
    tree, table = near_duplicate_table(synth, n_base=2100, tips=900, seed=23)
    inp = ssu.Inputs(*tree, *table)
    say("banded one-off", name, kw)
    s, m = ssu.matrix_inmem(inp, name, **kw)
    assert s == ssu.OKAY, ssu.last_error()
    p = ssu.Plan(inp, name, **kw)
    p.execute()
    ref = p.matrix()
    refined = p.refined_pairs()
    p.close()
    say("refined pairs", refined)
    assert np.array_equal(m, ref)


@pytest.mark.parametrize("name,kw", [("weighted_normalized", {}), ("generalized", {"alpha": 0.5})])
def test_one_off_few_refined_samples_equals_plan(ssu, O, synth, name, kw):
    """A few near-duplicate samples among many: after the banded copies only
    the refined samples' square of entries is copied again; the matrix equals
    a resident plan's, bit for bit, and the refined pairs stay within 1e-6 of
    the FP64 oracle."""
    st, le, nm = synth.random_tree(900, 31)
    n, k = 4200, 40
    obs, sids, ip, ix, d = synth.random_table(900, n, 0.08, 32)
    dense = np.zeros((len(obs), n))
    for r in range(len(obs)):
        dense[r, ix[ip[r]:ip[r + 1]]] = d[ip[r]:ip[r + 1]]
    dense[:, n - k:] = dense[:, :k]   # the last k samples repeat the first k, one count raised
    rng = np.random.default_rng(33)
    for s in range(n - k, n):
        dense[rng.choice(np.flatnonzero(dense[:, s])), s] += 1.0
    rows = [np.flatnonzero(dense[r]) for r in range(len(obs))]
    keep = [r for r in range(len(obs)) if len(rows[r])]
    table = ([obs[r] for r in keep], [f"s{i}" for i in range(n)],
             np.concatenate([[0], np.cumsum([len(rows[r]) for r in keep])]).astype(np.uint32),
             np.concatenate([rows[r] for r in keep]).astype(np.uint32),
             np.concatenate([dense[r, rows[r]] for r in keep]))
    inp = ssu.Inputs(st, le, nm, *table)
    say("few refined", name, kw)
    s, m = ssu.matrix_inmem(inp, name, **kw)
    assert s == ssu.OKAY, ssu.last_error()
    p = ssu.Plan(inp, name, **kw)
    p.execute()
    ref = p.matrix()
    refined = p.refined_pairs()
    p.close()
    say("refined pairs", refined)
    assert refined >= k and np.array_equal(m, ref)
    # the near-duplicate pairs against the FP64 oracle, on the sub-table of the 2k samples
    sel = list(range(k)) + list(range(n - k, n))
    sub_rows = [np.array([j for j, c in enumerate(sel) if dense[r, c] > 0], dtype=np.uint32) for r in keep]
    sub = ([obs[r] for r in keep], [f"s{i}" for i in sel],
           np.concatenate([[0], np.cumsum([len(x) for x in sub_rows])]).astype(np.uint32),
           np.concatenate(sub_rows).astype(np.uint32),
           np.concatenate([dense[r, np.array(sel)[sub_rows[i]]] for i, r in enumerate(keep)]))
    want = O.matrix(O.Tree(st.astype(np.uint8), le, nm), O.Table(*sub), name, **kw)
    got = m[np.ix_(sel, sel)]
    off = ~np.eye(2 * k, dtype=bool)
    assert np.max(np.abs(got - want)[off] / np.maximum(np.abs(want[off]), 1e-300)) <= 1e-6
