"""The ssu and faithpd command-line front ends (src/su.cpp, src/faithpd.cpp).

CPU tests: partial-report, check-partial and merge-partial (partials written
through the C ABI), option errors, HDF5 matrix output read back with h5py
(when the conda h5py is present).  GPU tests: one-off, partial + merge-partial
and faithpd on the reference's own test.biom / test.tre and crawford inputs,
compared with the reference's golden values.
"""
import ctypes as C
import json
import os
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN

HERE = os.path.dirname(os.path.abspath(__file__))
BIN = os.path.join(HERE, "..", "unifrac-binaries_amd", "bin")
SSU = os.path.join(BIN, "ssu")
FAITHPD = os.path.join(BIN, "faithpd")
H5PY = "/opt/conda/bin/python3.9"


@pytest.fixture(scope="module", autouse=True)
def _built(ssu):
    if not (os.path.exists(SSU) and os.path.exists(FAITHPD)):
        pytest.fail("CLI binaries missing: build with __graft_entry__.build()")


def run(*args, cwd=None):
    return subprocess.run([*args], capture_output=True, text=True, cwd=cwd, timeout=120)


def read_txt_matrix(path):
    lines = open(path).read().splitlines()
    ids = lines[0].split("\t")[1:]
    rows = [ln.split("\t") for ln in lines[1:]]
    assert [r[0] for r in rows] == ids
    return ids, np.array([[float(x) for x in r[1:]] for r in rows])


def h5_read(path):
    if not os.path.exists(H5PY):
        pytest.skip("no h5py interpreter to read the HDF5 output back")
    code = ("import h5py, json, sys\n"
            "f = h5py.File(sys.argv[1], 'r')\n"
            "out = {k: None for k in f}\n"
            "out['format'] = f['format'][0].decode(); out['version'] = f['version'][0].decode()\n"
            "out['order'] = [x.decode() if isinstance(x, bytes) else x for x in f['order'][:]]\n"
            "out['matrix'] = f['matrix'][:].tolist() if 'matrix' in f else None\n"
            "for k in ('pcoa_eigvals', 'pcoa_samples', 'pcoa_proportion_explained'):\n"
            "    if k in f: out[k] = f[k][:].tolist()\n"
            "out['dtype'] = str(f['matrix'].dtype) if 'matrix' in f else None\n"
            "print(json.dumps(out))\n")
    r = subprocess.run([H5PY, "-c", code, path], capture_output=True, text=True, timeout=120)
    if r.returncode != 0:
        pytest.skip(f"h5py unavailable: {r.stderr[-200:]}")
    return json.loads(r.stdout)


def write_partials(ssu, tmp_path, ids, full, splits):
    """Partial files of the stripes of a symmetric matrix, through write_partial."""
    n = len(ids)
    total = (n + 1) // 2
    stripes = [np.array([full[k, (k + s + 1) % n] for k in range(n)]) for s in range(total)]
    enc = [i.encode() for i in ids]
    idp = (C.c_char_p * n)(*enc)
    for a, b in zip(splits[:-1], splits[1:]):
        arr = stripes[a:b]
        ptrs = (C.POINTER(C.c_double) * len(arr))(*[x.ctypes.data_as(C.POINTER(C.c_double)) for x in arr])
        pm = ssu.PartialMat(n, idp, ptrs, a, b, total, True)
        assert ssu.lib().write_partial(str(tmp_path / f"m.start{a}.partial").encode(), C.byref(pm)) == ssu.WRITE_OKAY


def sym(n, seed):
    a = np.random.default_rng(seed).random((n, n))
    m = a + a.T
    np.fill_diagonal(m, 0)
    return m


# ----------------------------------------------------------------- CPU tests

def test_usage_and_option_errors():
    r = run(SSU)
    assert r.returncode == 0 and "usage: ssu" in r.stdout
    r = run(SSU, "-i", "x", "-o", "y", "-m", "unweighted", "-t", "z", "--format", "bogus")
    assert r.returncode != 0 and "Invalid format" in r.stderr
    r = run(SSU, "--mode", "partial", "-i", "x", "-t", "y", "-o", "z", "-m", "unweighted", "--start", "3", "--stop", "1")
    assert r.returncode != 0 and "stop stripe must be > start stripe" in r.stderr
    r = run(SSU, "--mode", "merge-partial", "-o", "z")
    assert r.returncode != 0 and "pattern missing" in r.stderr
    r = run(SSU, "-i", "x", "-o", "y", "-m", "unweighted", "-t", "z", "--normalize-sample-counts", "maybe")
    assert r.returncode != 0 and "normalize-sample-counts" in r.stderr
    r = run(FAITHPD)
    assert r.returncode == 0 and "usage: faithpd" in r.stdout
    r = run(FAITHPD, "-i", "x", "-t", "y")
    assert r.returncode != 0 and "output filename missing" in r.stderr


@pytest.mark.parametrize("parts,expect", [(3, [(0, 2), (2, 4), (4, 5)]), (2, [(0, 3), (3, 5)]),
                                          (5, [(0, 1), (1, 2), (2, 3), (3, 4), (4, 5)])])
def test_partial_report(parts, expect):
    """src/su.cpp:122-170 on crawford (9 samples, 5 stripes)."""
    biom = os.path.join(GOLDEN, "crawford.biom")
    r = run(SSU, "--mode", "partial-report", "-i", biom, "--n-partials", str(parts), "--report-bare")
    assert r.returncode == 0, r.stderr
    got = [tuple(int(x) for x in ln.split("\t")) for ln in r.stdout.splitlines()]
    assert got == expect
    r = run(SSU, "--mode", "partial-report", "-i", biom, "--n-partials", str(parts))
    assert "Total samples: 9" in r.stdout and "Total stripes: 5" in r.stdout


def test_check_and_merge_partial_ascii(ssu, tmp_path):
    ids = [f"s{i}" for i in range(13)]
    full = sym(13, 1)
    write_partials(ssu, tmp_path, ids, full, [0, 3, 5, 7])
    pat = str(tmp_path / "m.start*.partial")
    r = run(SSU, "--mode", "check-partial", "--partial-pattern", pat)
    assert r.returncode == 0 and "Partials are ready to be merged" in r.stdout and "n_samples = 13" in r.stdout
    out = str(tmp_path / "merged.txt")
    r = run(SSU, "--mode", "merge-partial", "--partial-pattern", pat, "-o", out)
    assert r.returncode == 0, r.stderr
    got_ids, m = read_txt_matrix(out)
    assert got_ids == ids
    assert np.max(np.abs(m - full)) <= 1e-15 * 4          # %.16g text round trip
    os.remove(tmp_path / "m.start3.partial")
    r = run(SSU, "--mode", "check-partial", "--partial-pattern", pat)
    assert r.returncode == 1 and "Insufficient number of stripes" in r.stderr
    r = run(SSU, "--mode", "merge-partial", "--partial-pattern", pat, "-o", out)
    assert r.returncode != 0 and "Unable to complete merge; err 1" in r.stderr


@pytest.mark.parametrize("fmt,dtype", [("hdf5_fp64", "float64"), ("hdf5_fp32", "float32")])
def test_merge_partial_hdf5(ssu, tmp_path, fmt, dtype):
    ids = [f"sample-{i}" for i in range(10)]
    full = sym(10, 2)
    write_partials(ssu, tmp_path, ids, full, [0, 2, 5])
    out = str(tmp_path / "merged.h5")
    r = run(SSU, "--mode", "merge-partial", "--partial-pattern", str(tmp_path / "m.start*.partial"), "-o", out,
            "--format", fmt, "--pcoa", "0")
    assert r.returncode == 0, r.stderr
    h = h5_read(out)
    assert h["format"] == "BDSM" and h["version"] == "2020.12" and h["order"] == ids and h["dtype"] == dtype
    assert np.array_equal(np.asarray(h["matrix"]), full.astype(dtype).astype(np.float64))
    assert "pcoa_eigvals" not in h


# ----------------------------------------------------------------- GPU tests

@pytest.mark.gpu
def test_one_off_ascii_on_reference_inputs(ref_trees, goldens, tmp_path):
    tre = tmp_path / "test.tre"
    tre.write_text(ref_trees["test"])
    out = tmp_path / "out.dm"
    r = run(SSU, "-i", os.path.join(GOLDEN, "test.biom"), "-t", str(tre), "-m", "unweighted_fp64", "-o", str(out))
    assert r.returncode == 0, r.stderr
    ids, m = read_txt_matrix(out)
    assert ids == [f"Sample{i}" for i in range(1, 7)]
    cf = m[np.triu_indices(6, 1)]                       # squareform order of test/capi_test.c:25
    assert np.max(np.abs(cf - np.array(goldens["condensed_unweighted"]["values"]))) < 1e-8
    assert np.array_equal(m, m.T) and np.all(np.diag(m) == 0)


@pytest.mark.gpu
def test_partial_then_merge_equals_one_off(ref_trees, tmp_path):
    tre = tmp_path / "crawford.tre"
    tre.write_text(ref_trees["crawford"])
    biom = os.path.join(GOLDEN, "crawford.biom")
    base = ["-i", biom, "-t", str(tre), "-m", "weighted_normalized_fp64"]
    r = run(SSU, *base, "-o", str(tmp_path / "full.txt"))
    assert r.returncode == 0, r.stderr
    r = run(SSU, "--mode", "partial-report", "-i", biom, "--n-partials", "2", "--report-bare")
    for ln in r.stdout.splitlines():
        a, b = ln.split("\t")
        r = run(SSU, *base, "--mode", "partial", "--start", a, "--stop", b, "-o", str(tmp_path / f"c.start{a}.partial"))
        assert r.returncode == 0, r.stderr
    r = run(SSU, "--mode", "merge-partial", "--partial-pattern", str(tmp_path / "c.start*.partial"),
            "-o", str(tmp_path / "merged.txt"))
    assert r.returncode == 0, r.stderr
    assert open(tmp_path / "merged.txt").read() == open(tmp_path / "full.txt").read()


@pytest.mark.gpu
def test_one_off_hdf5(ref_trees, tmp_path):
    tre = tmp_path / "test.tre"
    tre.write_text(ref_trees["test"])
    out = str(tmp_path / "out.h5")
    r = run(SSU, "-i", os.path.join(GOLDEN, "test.biom"), "-t", str(tre), "-m", "weighted_normalized", "-o", out,
            "--format", "hdf5", "--pcoa", "0")
    assert r.returncode == 0, r.stderr
    h = h5_read(out)
    assert h["format"] == "BDSM" and h["dtype"] == "float32" and len(h["order"]) == 6


@pytest.mark.gpu
def test_faithpd_on_reference_inputs(ref_trees, goldens, tmp_path):
    tre = tmp_path / "test.tre"
    tre.write_text(ref_trees["test"])
    out = tmp_path / "faith.txt"
    r = run(FAITHPD, "-i", os.path.join(GOLDEN, "test.biom"), "-t", str(tre), "-o", str(out))
    assert r.returncode == 0, r.stderr
    lines = out.read_text().splitlines()
    assert lines[0] == "#SampleID\tfaith_pd"
    assert [float(x.split("\t")[1]) for x in lines[1:]] == [float(v) for v in goldens["faith_test_tre"]["values"]]


@pytest.mark.gpu
def test_merge_partial_hdf5_with_pcoa(ssu, O, tmp_path):
    """--pcoa 3 (the default is 10) adds the PCoA datasets to the merged matrix
    (src/su.cpp:172-237); the eigenvalues match the exact decomposition."""
    ids = [f"s{i}" for i in range(12)]
    pts = np.random.default_rng(3).normal(size=(12, 3)) * np.array([3.0, 1.5, 0.5])
    full = np.sqrt(((pts[:, None, :] - pts[None, :, :]) ** 2).sum(-1))
    write_partials(ssu, tmp_path, ids, full, [0, 3, 6])
    out = str(tmp_path / "merged.h5")
    r = run(SSU, "--mode", "merge-partial", "--partial-pattern", str(tmp_path / "m.start*.partial"), "-o", out,
            "--format", "hdf5_fp64", "--pcoa", "3")
    assert r.returncode == 0, r.stderr
    h = h5_read(out)
    w, coords, prop = O.pcoa_exact(full, 3)
    assert np.max(np.abs(np.array(h["pcoa_eigvals"]) - w)) < 1e-9
    assert np.max(np.abs(np.abs(np.array(h["pcoa_samples"])) - np.abs(coords))) < 1e-9
    assert np.max(np.abs(np.array(h["pcoa_proportion_explained"]) - prop)) < 1e-9
