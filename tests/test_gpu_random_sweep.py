"""Seeded random cases against the oracle, through the C ABI.

Each case draws:
* a tree shape: random binary, multifurcating (internal edges contracted), or caterpillar;
* branch lengths, with some set to 0 in half the cases;
* a sample count off and on the 32- and 64-sample panel boundaries, some above the 1,024
  samples from which the occupancy splits apply;
* a density from sparse to dense, with or without Zipf tip popularity;
* lognormal or negative-binomial counts, made fractional in some cases;
* a method with its options (every row of the parity list) and a precision.

Each case also computes Faith PD, which must be bit-exact (over the samples whose total is at
least 1, which the reference keeps). Tolerances are those of the parity
tests: FP64 1e-11, FP32 1e-6 relative per matrix entry against the FP64 oracle.
"""
import sys

import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu

METHODS = [
    ("unweighted", {}), ("unweighted_unnormalized", {}), ("unweighted", {"bypass_tips": True}),
    ("weighted_normalized", {}), ("weighted_normalized", {"normalize": False}), ("weighted_unnormalized", {}),
    ("weighted_normalized", {"bypass_tips": True}), ("generalized", {"alpha": 0.5}),
    ("generalized", {"alpha": 0.0}), ("generalized", {"alpha": 1.0}), ("generalized", {"alpha": 1.5}),
    ("unweighted", {"variance_adjust": True}), ("unweighted_unnormalized", {"variance_adjust": True}),
    ("weighted_normalized", {"variance_adjust": True}), ("weighted_unnormalized", {"variance_adjust": True}),
    ("generalized", {"alpha": 0.5, "variance_adjust": True}),
]
TOL = {"fp64": 1e-11, "fp32": 1e-6}
CASES = 96


def say(*a):
    print("[gpu-case]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def _need_device(ssu):
    if not ssu.device_available():
        pytest.fail("no HIP device visible: the HIP path cannot run (no CPU fallback exists)")


def draw(synth, seed):
    rng = np.random.default_rng(9000 + seed)
    n = int(rng.choice([2, 3, 31, 33, 63, 65, 100, 127, 129, 200, 257, 333, 1030, 1100]))
    # from 1,024 samples the occupancy splits apply; smaller trees keep the oracle quick there
    tips = int(rng.integers(20, 400 if n > 1000 else 1500))
    shape = str(rng.choice(["binary", "multifurcating", "caterpillar"], p=[0.5, 0.3, 0.2]))
    if shape == "binary":
        st, le, nm = synth.random_tree(tips, 100 + seed)
    elif shape == "multifurcating":
        st, le, nm = synth.contract_tree(synth.random_tree(tips, 100 + seed), 0.4, 200 + seed)
    else:
        st, le, nm = synth.caterpillar_tree(tips, 100 + seed)
    le = np.array(le, dtype=np.float64)
    zero_len = bool(rng.random() < 0.5)
    if zero_len:
        le[rng.random(le.size) < 0.1] = 0.0
    dens = float(rng.choice([0.01, 0.05, 0.2, 0.6]))
    zipf = 1.1 if rng.random() < 0.5 else None
    counts = str(rng.choice(["lognormal", "negbin"]))
    obs, sids, ip, ix, d = synth.random_table(tips, n, dens, 300 + seed, zipf, counts)
    frac = bool(rng.random() < 0.3)
    if frac:
        d = np.asarray(d, dtype=np.float64) * rng.random(len(d)) + 0.01
    method, kw = METHODS[int(rng.integers(len(METHODS)))]
    prec = "fp64" if rng.random() < 0.4 else "fp32"
    desc = dict(tips=tips, shape=shape, zero_len=zero_len, n=n, dens=dens, zipf=zipf, counts=counts,
                frac=frac, method=method, kw=kw, prec=prec)
    return (st, le, nm), (obs, sids, ip, ix, d), method, kw, prec, desc


@pytest.mark.parametrize("seed", range(CASES))
def test_random_case_against_oracle(ssu, O, synth, seed):
    tree, table, method, kw, prec, desc = draw(synth, seed)
    say("random case", seed, desc)
    st, le, nm = tree
    inp = ssu.Inputs(st, le, nm, *table)
    s, m = ssu.matrix_inmem(inp, method + ("_fp64" if prec == "fp64" else ""), **kw)
    assert s == ssu.OKAY, ssu.last_error()
    ot, otab = O.Tree(np.asarray(st, np.uint8), le, nm), O.Table(*table)
    ref = O.matrix(ot, otab, method, **kw)
    err = rel_err(m, ref)
    say("max rel err", err)
    assert err <= TOL[prec]
    # Faith PD drops samples whose total is below 1 first, as the reference does
    # (src/api.cpp:577-578), then the observations left without entries
    s, vals, ids = ssu.faith_pd_inmem(inp)
    assert s == ssu.OKAY, ssu.last_error()
    obs, sids, ip, ix, d = table
    ix, d = np.asarray(ix), np.asarray(d, dtype=np.float64)
    tot = np.bincount(ix, weights=d, minlength=len(sids))
    keep = np.nonzero(tot >= 1.0)[0]
    if len(keep) == len(sids):
        assert np.array_equal(vals, O.faith_pd(ot, otab))
    else:
        new_ix = -np.ones(len(sids), dtype=np.int64)
        new_ix[keep] = np.arange(len(keep))
        fobs, fip, fix, fd = [], [0], [], []
        for r in range(len(obs)):
            q = np.arange(ip[r], ip[r + 1])
            q = q[new_ix[ix[q]] >= 0]
            if len(q):
                fobs.append(obs[r]); fix.extend(new_ix[ix[q]]); fd.extend(d[q]); fip.append(len(fix))
        ftab = O.Table(fobs, [sids[k] for k in keep], np.array(fip), np.array(fix), np.array(fd))
        assert ids == [sids[k] for k in keep]
        assert np.array_equal(vals, O.faith_pd(ot, ftab))
