"""The benchmark workload itself, at full size, on the device (C ABI of libssu.so).

16,384 samples on a 100,000-tip tree (1% tip density, Zipf(1.1) tip
popularity, lognormal counts): the unweighted and weighted_normalized
matrices through the default paths (occupancy splits on), checked through
properties that hold at any size and against independent paths:

* symmetric, zero diagonal, finite, within [0, 1] (both metrics are normalized);
* the first 64 rows equal those of the no-split paths (the int8 Gram over
  every node; the tiled kernel over every node) within 2e-6 relative;
* a 48-sample subset (first and last sample included) within 1e-6 relative of
  the FP64 oracle.
"""
import sys

import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu

SPLIT_ENV = {"unweighted": "SSU_GRAM_SPLIT", "weighted_normalized": "SSU_SPARSE_SPLIT"}


def say(*a):
    print("[gpu-case]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def _need_device(ssu):
    if not ssu.device_available():
        pytest.fail("no HIP device visible: the HIP path cannot run (no CPU fallback exists)")


@pytest.fixture(scope="module")
def workload(synth):
    return synth.workload("unw_100k_16384")


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


@pytest.mark.parametrize("method", ["unweighted", "weighted_normalized"])
def test_full_size_workload(ssu, O, workload, monkeypatch, method):
    tree, table = workload
    n = len(table[1])
    assert n == 16384
    inp = ssu.Inputs(*tree, *table)
    say("full size", method, "default split")
    p = ssu.Plan(inp, method)
    p.execute()
    assert p.sparse_terms() > 0   # the split is in use at this size
    m = p.matrix()
    p.close()
    assert np.all(np.isfinite(m))
    assert np.array_equal(m, m.T) and np.all(np.diag(m) == 0)
    assert m.min() >= 0.0 and m.max() <= 1.0 + 1e-6

    say("full size", method, "no split, rows 0..63")
    monkeypatch.setenv(SPLIT_ENV[method], "0")
    q = ssu.Plan(inp, method)
    q.execute()
    assert q.sparse_terms() == 0
    band = q.rows(0, 64)
    q.close()
    assert rel_err(m[:64], band) <= 2e-6

    idx = np.unique(np.concatenate([np.random.default_rng(7).choice(n, 46, replace=False), [0, n - 1]]))
    st, le, nm = tree
    say("full size", method, "oracle on", len(idx), "samples")
    ref = O.matrix(O.Tree(np.asarray(st, np.uint8), le, nm), O.Table(*_sub_table(table, idx)), method)
    assert rel_err(m[np.ix_(idx, idx)], ref) <= 1e-6
