"""Gower centring, PCoA and PERMANOVA on the device (src/skbio_alt.cpp:144-233).

Against the reference's own vectors (src/test_ska.cpp and
src/test_ska_pcoa_big.dat, copied to tests/golden/ska_vectors.json) with the
reference's tolerances, against the oracle's numpy restatement (exact
eigendecomposition, pseudo-F), and through the file entry points and the CLI.
The permutation p-values follow this library's random stream, so they are held
to the reference's tolerance around its values (which come from a different
generator), and to a permutation-free bound where the statistic is extreme.
"""
import ctypes as C
import json
import os
import subprocess
import sys

import numpy as np
import pytest

from conftest import GOLDEN

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
SSU = os.path.join(HERE, "..", "unifrac-binaries_amd", "bin", "ssu")
H5PY = "/opt/conda/bin/python3.9"


def say(*a):
    print("[gpu-case]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def _need_device(ssu):
    if not ssu.device_available():
        pytest.fail("no HIP device visible: the HIP path cannot run (no CPU fallback exists)")


@pytest.fixture(scope="module")
def ska():
    return json.load(open(os.path.join(GOLDEN, "ska_vectors.json")))


def h5_all(path):
    if not os.path.exists(H5PY):
        pytest.skip("no h5py interpreter to read the HDF5 output back")
    code = ("import h5py, json, sys\n"
            "f = h5py.File(sys.argv[1], 'r')\n"
            "out = {}\n"
            "for k in f:\n"
            "    v = f[k][()]\n"
            "    if hasattr(v, 'tolist'): v = v.tolist()\n"
            "    if isinstance(v, list): v = [x.decode() if isinstance(x, bytes) else x for x in v]\n"
            "    elif isinstance(v, bytes): v = v.decode()\n"
            "    out[k] = v\n"
            "print(json.dumps(out))\n")
    r = subprocess.run([H5PY, "-c", code, path], capture_output=True, text=True, timeout=120)
    if r.returncode != 0:
        pytest.skip(f"h5py unavailable: {r.stderr[-300:]}")
    return json.loads(r.stdout)


@pytest.mark.parametrize("fp32", [False, True])
def test_center_matrix_reference_vectors(ssu, ska, fp32):
    """src/test_ska.cpp:7-60 and the centred crawford matrix of :62-140."""
    say("center", fp32)
    c = ska["center_mat"]
    got = ssu.center_matrix(np.array(c["matrix"]).reshape(6, 6), fp32=fp32)
    assert np.max(np.abs(got - np.array(c["expected"]).reshape(6, 6))) < 1e-6
    p = ska["pcoa"]
    got = ssu.center_matrix(np.array(p["matrix"]).reshape(9, 9), fp32=fp32)
    assert np.max(np.abs(got - np.array(p["centered"]).reshape(9, 9))) < 1e-5


@pytest.mark.parametrize("fp32", [False, True])
def test_pcoa_reference_vectors(ssu, ska, fp32):
    """src/test_ska.cpp:142-245: 9 samples, 5 dimensions; eigenvalues,
    coordinates up to sign and proportion explained within 1e-5."""
    say("pcoa crawford", fp32)
    p = ska["pcoa"]
    e, s, prop = ssu.pcoa(np.array(p["matrix"]).reshape(9, 9), 5, fp32=fp32)
    assert np.max(np.abs(e - np.array(p["eigenvalues"]))) < 1e-5
    assert np.max(np.abs(np.abs(s) - np.abs(np.array(p["samples"]).reshape(9, 5)))) < 1e-5
    assert np.max(np.abs(prop - np.array(p["proportion"]))) < 1e-5


@pytest.mark.parametrize("fp32", [False, True])
def test_pcoa_big_reference_vectors(ssu, ska, fp32):
    """src/test_ska.cpp:247-420 on src/test_ska_pcoa_big.dat (57 samples of EMP,
    7 dimensions), with the reference's tolerances: the last three eigenvalues
    are close together and may come back in any order."""
    say("pcoa big", fp32)
    b = ska["pcoa_big"]
    n, dims = b["n"], b["dims"]
    e, s, prop = ssu.pcoa(np.array(b["matrix"]).reshape(n, n), dims, fp32=fp32)
    exp_e, exp_s, exp_p = np.array(b["eigenvalues"]), np.array(b["samples"]).reshape(n, dims), np.array(b["proportion"])
    for i in range(dims):
        assert abs(e[i] - exp_e[i]) < (0.01 if i < 4 else 0.1), i
        assert abs(prop[i] - exp_p[i]) < (0.001 if i < 4 else 0.01), i
    tol = 0.1 if fp32 else 0.05
    for i in range(n):
        for d in range(dims):
            if d < 4:
                assert abs(abs(s[i, d]) - abs(exp_s[i, d])) < tol, (i, d)
            else:
                assert min(abs(abs(s[i, d]) - abs(exp_s[i, k])) for k in (4, 5, 6)) < 0.15, (i, d)


def test_pcoa_against_exact_decomposition(ssu, O):
    """Euclidean distances of points with a decaying spectrum: the randomized
    top eigenpairs agree with the oracle's full decomposition."""
    rng = np.random.default_rng(11)
    pts = rng.normal(size=(400, 6)) * np.array([8.0, 4.0, 2.0, 1.0, 0.5, 0.25])
    m = np.sqrt(((pts[:, None, :] - pts[None, :, :]) ** 2).sum(-1))
    say("pcoa vs exact", m.shape)
    e, s, prop = ssu.pcoa(m, 4)
    w, coords, p = O.pcoa_exact(m, 4)
    assert np.max(np.abs(e - w) / w) < 1e-6
    assert np.max(np.abs(np.abs(s) - np.abs(coords))) < 1e-5 * np.max(np.abs(coords))
    assert np.max(np.abs(prop - p)) < 1e-6
    # more dimensions than samples: the extra ones are zero
    e, s, prop = ssu.pcoa(m[:5, :5], 8)
    assert np.all(e[5:] == 0) and np.all(s[:, 5:] == 0)


@pytest.mark.parametrize("key", ["ties", "noties", "unequal"])
@pytest.mark.parametrize("fp32", [False, True])
def test_permanova_reference_vectors(ssu, ska, key, fp32):
    """src/test_ska.cpp test_permanova_*: pseudo-F within 1e-5, p within 0.05
    of the reference's value, for both groupings."""
    v = ska[f"permanova_{key}"]
    n = len(v["grouping"])
    m = np.array(v["matrix_fp32" if fp32 else "matrix_fp64"]).reshape(n, n)
    for gk in ("grouping", "grouping_swapped"):
        say("permanova", key, fp32, gk)
        f, p = ssu.permanova(m, v[gk], 999, fp32=fp32)
        assert abs(f - v["stat"]) < 1e-5
        assert abs(p - v["pvalue"]) < 0.05


def test_permanova_against_oracle_and_extremes(ssu, O):
    """Pseudo-F equals the oracle's on a random matrix with three groups; a
    perfectly separated grouping gets the smallest p-value 1/(n_perm+1)."""
    rng = np.random.default_rng(5)
    a = rng.random((150, 150))
    m = a + a.T
    np.fill_diagonal(m, 0)
    g = rng.integers(0, 3, 150)
    say("permanova random")
    f, p = ssu.permanova(m, g, 199)
    assert abs(f - O.permanova_f(m, g)) <= 1e-9 * abs(f)
    assert 0.0 < p <= 1.0
    # separated clusters
    pts = np.concatenate([rng.normal(size=(40, 2)), rng.normal(size=(40, 2)) + 20.0])
    d = np.sqrt(((pts[:, None, :] - pts[None, :, :]) ** 2).sum(-1))
    f, p = ssu.permanova(d, [0] * 40 + [1] * 40, 99)
    assert abs(f - O.permanova_f(d, [0] * 40 + [1] * 40)) <= 1e-9 * f
    assert p == pytest.approx(1.0 / 100)
    # one group: undefined
    f, p = ssu.permanova(d, [0] * 80, 9)
    assert np.isnan(f) and np.isnan(p)


def _grouping_file(path, ids, columns):
    with open(path, "w") as fh:
        fh.write("#SampleID\t" + "\t".join(columns) + "\n")
        for i, sid in enumerate(ids):
            fh.write(sid + "\t" + "\t".join(vals[i] for vals in columns.values()) + "\n")
        fh.write("not_a_sample\t" + "\t".join("x" for _ in columns) + "\n")


def test_compute_permanova_from_grouping_file(ssu, O, tmp_path):
    """src/api.cpp:951-976 with a TSV grouping file (src/tsv.cpp:120-178):
    groups numbered in order of first appearance; rows of other samples ignored."""
    rng = np.random.default_rng(8)
    n = 40
    a = rng.random((n, n))
    m = a + a.T
    np.fill_diagonal(m, 0)
    ids = [f"S{i}" for i in range(n)]
    cols = {"body_site": [["gut", "skin", "oral"][i % 3] for i in range(n)],
            "sex": [["f", "m"][(i // 3) % 2] for i in range(n)]}
    tsv = str(tmp_path / "meta.tsv")
    _grouping_file(tsv, ids, cols)
    idp = (C.c_char_p * n)(*[s.encode() for s in ids])
    res = ssu.MatFullFP64(n, 0, m.ctypes.data_as(C.POINTER(C.c_double)), idp)
    names = (C.c_char_p * 2)(b"body_site", b"sex")
    f = (C.c_double * 2)()
    p = (C.c_double * 2)()
    g = (C.c_uint * 2)()
    say("compute_permanova")
    st = ssu.lib().compute_permanova_fp64(tsv.encode(), 2, names, C.byref(res), 99, f, p, g)
    assert st == ssu.OKAY
    assert list(g) == [3, 2]
    exp0 = O.permanova_f(m, [i % 3 for i in range(n)])
    exp1 = O.permanova_f(m, [(i // 3) % 2 for i in range(n)])
    assert abs(f[0] - exp0) <= 1e-9 * abs(exp0) and abs(f[1] - exp1) <= 1e-9 * abs(exp1)
    bad = (C.c_char_p * 1)(b"no_such_column")
    assert ssu.lib().compute_permanova_fp64(tsv.encode(), 1, bad, C.byref(res), 99, f, p, g) == ssu.GROUPING_MISSING


def test_file_outputs_with_pcoa_and_permanova(ssu, O, ref_trees, tmp_path):
    """unifrac_to_file_v3 with PCoA and PERMANOVA (src/api.cpp:989-1117), the
    CLI's one-off with -g/-c, and multi mode with both (src/api.cpp:1444-1563)."""
    tre = tmp_path / "test.tre"
    tre.write_text(ref_trees["test"])
    biom = os.path.join(GOLDEN, "test.biom")
    ids = [f"Sample{i}" for i in range(1, 7)]
    tsv = str(tmp_path / "meta.tsv")
    _grouping_file(tsv, ids, {"grp": ["a", "a", "b", "b", "c", "c"]})
    out = str(tmp_path / "o.h5")
    say("unifrac_to_file_v3 pcoa+permanova")
    st = ssu.lib().unifrac_to_file_v3(biom.encode(), str(tre).encode(), out.encode(), b"unweighted_fp64", False,
                                      C.c_double(1.0), False, True, 1, b"hdf5", 0, True, 3, 99, tsv.encode(), b"grp",
                                      None)
    assert st == ssu.OKAY, ssu.last_error()
    h = h5_all(out)
    m = np.array(h["matrix"])
    w, coords, prop = O.pcoa_exact(m, 3)
    assert h["pcoa_method"] in ("FSVD", ["FSVD"])
    assert np.max(np.abs(np.array(h["pcoa_eigvals"]) - w)) < 1e-9
    assert np.max(np.abs(np.abs(np.array(h["pcoa_samples"])) - np.abs(coords))) < 1e-9
    assert h["stat_methods"] == ["PERMANOVA"] and h["stat_test_names"] == ["pseudo-F"]
    assert h["stat_grouping_names"] == ["grp"] and h["stat_n_groups"] == [3] and h["stat_n_permutations"] == [99]
    assert abs(h["stat_values"][0] - O.permanova_f(m, [0, 0, 1, 1, 2, 2])) < 1e-9
    say("cli one-off pcoa+permanova")
    out2 = str(tmp_path / "cli.h5")
    r = subprocess.run([SSU, "-i", biom, "-t", str(tre), "-m", "unweighted", "-o", out2, "--format", "hdf5_fp32",
                        "--pcoa", "4", "-g", tsv, "-c", "grp"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    h2 = h5_all(out2)
    assert len(h2["pcoa_eigvals"]) == 4 and h2["stat_n_permutations"] == [999]
    say("multi pcoa+permanova")
    out3 = str(tmp_path / "multi.h5")
    st = ssu.lib().unifrac_multi_to_file_v3(biom.encode(), str(tre).encode(), out3.encode(), b"unweighted", False,
                                            C.c_double(1.0), False, True, 1, b"hdf5", 3, 5, True, 2, 49,
                                            tsv.encode(), b"grp", None)
    assert st == ssu.OKAY, ssu.last_error()
    h3 = h5_all(out3)
    kept = h3["order"]
    assert all(f"pcoa_eigvals:{i}" in h3 and f"matrix:{i}" in h3 for i in range(3))
    assert h3["stat_grouping_names"] == ["grp:0", "grp:1", "grp:2"]
    for i in range(3):
        mi = np.array(h3[f"matrix:{i}"], np.float64)
        g = [["a", "a", "b", "b", "c", "c"][ids.index(s)] for s in kept]
        if len(set(g)) > 1 and len(set(g)) < len(g):
            assert abs(h3["stat_values"][i] - O.permanova_f(mi, g)) < 1e-4 * max(1.0, abs(h3["stat_values"][i]))
