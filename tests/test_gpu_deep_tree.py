"""Deep trees: a 20,000-tip caterpillar tree (19,999 levels) against the oracle.

The reference walks the tree in postorder sequentially, so depth does not
matter to it (src/unifrac_cmp.cpp:103-136, src/unifrac_internal.cpp:230-258).
Here the embedding runs by rounds of heavy paths (at most log2(tips) + 1
launches) instead of one launch per height level; both schedules give the same
FP64 sums, so the matrices are bit-identical between them, Faith PD is
bit-exact against the oracle, and every method is within its tolerance of the
FP64 oracle (1e-11 FP64 names, 1e-6 FP32 names).
"""
import os
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


@pytest.fixture(scope="module")
def deep(ssu, synth):
    tree = synth.caterpillar_tree(20000, 5)
    table = synth.random_table(20000, 300, 0.05, 6)
    return tree, table, ssu.Inputs(*tree, *table)


@pytest.fixture(scope="module")
def oracle_tree(O, deep):
    st, le, nm = deep[0]
    return O.Tree(np.asarray(st, np.uint8), le, nm), O.Table(*deep[1])


METHODS = [("unweighted", {}), ("unweighted_fp64", {}), ("weighted_normalized", {}), ("weighted_normalized_fp64", {}),
           ("weighted_unnormalized", {}), ("generalized", {"alpha": 0.5}), ("unweighted_unnormalized_fp64", {}),
           ("weighted_normalized", {"variance_adjust": True})]


@pytest.mark.parametrize("name,kw", METHODS, ids=[m + ("-vaw" if kw.get("variance_adjust") else "") for m, kw in METHODS])
def test_caterpillar_against_oracle(ssu, O, deep, oracle_tree, monkeypatch, name, kw):
    _, _, inp = deep
    otree, otab = oracle_tree
    say("caterpillar", name, kw)
    p = ssu.Plan(inp, name, **kw)
    p.execute()
    m = p.matrix()
    say("timings (heavy paths)", p.timings())
    p.close()
    ref = O.matrix(otree, otab, name.replace("_fp64", ""), **kw)
    err = rel_err(m, ref)
    say("max rel err", err)
    assert err <= (1e-11 if name.endswith("_fp64") else 1e-6)
    monkeypatch.setenv("SSU_EMBED_PATHS", "0")   # one launch per height level
    q = ssu.Plan(inp, name, **kw)
    q.execute()
    say("timings (height levels)", q.timings())
    assert np.array_equal(q.matrix(), m)
    q.close()


def test_caterpillar_faith_bit_exact(ssu, O, deep, oracle_tree):
    _, _, inp = deep
    s, vals, _ = ssu.faith_pd_inmem(inp)
    assert s == ssu.OKAY
    assert np.array_equal(vals, O.faith_pd(*oracle_tree))


def test_paths_match_levels_on_random_trees(ssu, synth, monkeypatch):
    """A bushy random tree, and the same tree with a third of its internal
    nodes contracted (nodes of three or more children, summed left to right):
    heavy paths and height levels give bit-identical matrices."""
    tree = synth.random_tree(3000, 9)
    table = synth.random_table(3000, 1200, 0.03, 10, zipf_s=1.1)
    inp = ssu.Inputs(*tree, *table)
    minp = ssu.Inputs(*synth.contract_tree(tree, 0.33, 12), *table)
    for case, x in (("random", inp), ("multifurcating", minp)):
        for name in ("unweighted", "weighted_normalized", "weighted_normalized_fp64", "generalized"):
            monkeypatch.setenv("SSU_EMBED_PATHS", "1")
            p = ssu.Plan(x, name)
            p.execute()
            a = p.matrix()
            p.close()
            monkeypatch.setenv("SSU_EMBED_PATHS", "0")
            q = ssu.Plan(x, name)
            q.execute()
            b = q.matrix()
            q.close()
            say(case, name)
            assert np.array_equal(a, b)
