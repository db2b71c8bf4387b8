"""ORACLE -- test infrastructure only.

ctypes front-end for

* ``liboracle.so``: the clean-room FP64 CPU restatement in ``ssu_oracle.c``;
* numpy restatements of the distance-matrix statistics the reference takes
  from its third-party skbb library (``src/skbio_alt.cpp:144-233``): Gower
  centring, exact PCoA (a full symmetric eigendecomposition, the limit the
  randomized FSVD approximates) and the PERMANOVA pseudo-F;
* ``_ref/libssu_ref.so``: the reference implementation's own CPU code
  (``src/tree.cpp``, ``src/biom_inmem.cpp``, ``src/unifrac_internal.cpp`` and
  the ``run_*_T`` kernels of ``src/unifrac_task_impl.hpp``) behind
  ``ref_harness.cpp``; present only where it was built.

Only tests, ``__graft_entry__.smoke()`` and bench.py's CPU-baseline leg import
this module.  Nothing in the product path does.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_PATH_V3 = os.path.join(HERE, "_ref", "libssu_ref.so")
REF_PATH_V4 = os.path.join(HERE, "_ref", "libssu_ref_v4.so")


def _host_has_v4() -> bool:
    """x86-64-v4 needs AVX-512 F, BW, CD, DQ and VL."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("flags"):
                    flags = set(line.split(":", 1)[1].split())
                    return {"avx512f", "avx512bw", "avx512cd", "avx512dq", "avx512vl"} <= flags
    except OSError:
        pass
    return False


# the AVX-512 build where the host runs it, else the AVX2 one
REF_ISA = "x86-64-v4" if os.path.exists(REF_PATH_V4) and _host_has_v4() else "x86-64-v3"
REF_PATH = REF_PATH_V4 if REF_ISA == "x86-64-v4" else REF_PATH_V3

METHODS = {
    "unweighted": 0,
    "unweighted_unnormalized": 1,
    "weighted_normalized": 2,
    "weighted_unnormalized": 3,
    "generalized": 4,
}

_lib = None
_ref = None


def build(force: bool = False) -> None:
    """Compile liboracle.so (and _ref when the reference sources are present)."""
    targets = ["liboracle"]
    if os.path.isdir("/root/reference/src"):
        targets.append("ref")
    if force or not os.path.exists(LIB_PATH) or (len(targets) > 1 and not os.path.exists(REF_PATH)):
        subprocess.run(["make", "-s", "-C", HERE] + targets, check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = C.CDLL(LIB_PATH)
        _lib.orc_unifrac_stripes.restype = C.c_int
        _lib.orc_faith_pd.restype = C.c_int
        _lib.orc_postorder.restype = C.c_int
        _lib.orc_newick_to_bp.restype = C.c_int
    return _lib


def ref_available() -> bool:
    return os.path.exists(REF_PATH)


def ref():
    global _ref
    if _ref is None:
        _ref = C.CDLL(REF_PATH)
        for f in ("ref_unifrac_stripes", "ref_faith_pd", "ref_postorder", "ref_newick_to_bp", "ref_sample_counts"):
            getattr(_ref, f).restype = C.c_int
    return _ref


# ---------------------------------------------------------------- marshalling

def _cstrs(strs):
    arr = (C.c_char_p * max(1, len(strs)))()
    keep = [s.encode() if isinstance(s, str) else s for s in strs]
    for i, s in enumerate(keep):
        arr[i] = s
    return arr, keep


class Tree:
    """Balanced-parentheses tree: structure (uint8), lengths (float64), names (list[str])."""

    def __init__(self, structure, lengths, names):
        self.structure = np.ascontiguousarray(structure, dtype=np.uint8)
        self.lengths = np.ascontiguousarray(lengths, dtype=np.float64)
        self.names = list(names)
        self._cn, self._keep = _cstrs(self.names)

    @property
    def n_parens(self):
        return int(self.structure.size)

    def args(self):
        return (C.c_int(self.n_parens), self.structure.ctypes.data_as(C.c_void_p),
                self.lengths.ctypes.data_as(C.c_void_p), self._cn)


class Table:
    """Observation-major CSR table (the reference's support_biom_t layout)."""

    def __init__(self, obs_ids, sample_ids, indptr, indices, data):
        self.obs_ids = list(obs_ids)
        self.sample_ids = list(sample_ids)
        self.indptr = np.ascontiguousarray(indptr, dtype=np.uint32)
        self.indices = np.ascontiguousarray(indices, dtype=np.uint32)
        self.data = np.ascontiguousarray(data, dtype=np.float64)
        self._co, self._k1 = _cstrs(self.obs_ids)
        self._cs, self._k2 = _cstrs(self.sample_ids)

    @property
    def n_obs(self):
        return len(self.obs_ids)

    @property
    def n_samples(self):
        return len(self.sample_ids)

    def dense(self):
        m = np.zeros((self.n_obs, self.n_samples))
        for i in range(self.n_obs):
            a, b = self.indptr[i], self.indptr[i + 1]
            m[i, self.indices[a:b]] = self.data[a:b]
        return m


def newick_to_tree(newick: str) -> Tree:
    L = lib()
    n = C.c_int()
    s = C.POINTER(C.c_uint8)()
    le = C.POINTER(C.c_double)()
    nm = C.POINTER(C.c_char_p)()
    rc = L.orc_newick_to_bp(newick.encode(), C.byref(n), C.byref(s), C.byref(le), C.byref(nm))
    if rc != 0:
        raise ValueError("unbalanced newick")
    k = n.value
    structure = np.array([s[i] for i in range(k)], dtype=np.uint8)
    lengths = np.array([le[i] for i in range(k)], dtype=np.float64)
    names = [nm[i].decode() for i in range(k)]
    L.orc_free_bp(n, s, le, nm)
    return Tree(structure, lengths, names)


def stripes(tree: Tree, table: Table, method: str, variance_adjust=False, alpha=1.0,
            bypass_tips=False, normalize=True, start=0, stop=None, threads=0) -> np.ndarray:
    n = table.n_samples
    if stop is None:
        stop = (n + 1) // 2
    out = np.zeros((stop - start, n), dtype=np.float64)
    rc = lib().orc_unifrac_stripes(
        *tree.args(), C.c_int(table.n_obs), C.c_int(n), table._co,
        table.indptr.ctypes.data_as(C.c_void_p), table.indices.ctypes.data_as(C.c_void_p),
        table.data.ctypes.data_as(C.c_void_p), C.c_int(METHODS[method]), C.c_int(int(variance_adjust)),
        C.c_double(alpha), C.c_int(int(bypass_tips)), C.c_int(int(normalize)), C.c_int(threads),
        C.c_uint(start), C.c_uint(stop), out.ctypes.data_as(C.c_void_p))
    assert rc == 0
    return out


def stripes_to_matrix(st: np.ndarray, n: int) -> np.ndarray:
    m = np.zeros((n, n), dtype=np.float64)
    st = np.ascontiguousarray(st, dtype=np.float64)
    lib().orc_stripes_to_matrix(st.ctypes.data_as(C.c_void_p), C.c_uint(n), C.c_uint(st.shape[0]),
                                m.ctypes.data_as(C.c_void_p))
    return m


def stripes_to_condensed(st: np.ndarray, n: int, start=0) -> np.ndarray:
    cf = np.full(n * (n - 1) // 2, np.nan)
    st = np.ascontiguousarray(st, dtype=np.float64)
    lib().orc_stripes_to_condensed(st.ctypes.data_as(C.c_void_p), C.c_uint(n), C.c_uint(start),
                                   C.c_uint(start + st.shape[0]), cf.ctypes.data_as(C.c_void_p))
    return cf


def matrix(tree, table, method, **kw) -> np.ndarray:
    n = table.n_samples
    return stripes_to_matrix(stripes(tree, table, method, **kw), n)


def faith_pd(tree: Tree, table: Table) -> np.ndarray:
    out = np.zeros(table.n_samples)
    rc = lib().orc_faith_pd(*tree.args(), C.c_int(table.n_obs), C.c_int(table.n_samples), table._co,
                            table.indptr.ctypes.data_as(C.c_void_p), table.indices.ctypes.data_as(C.c_void_p),
                            table.data.ctypes.data_as(C.c_void_p), out.ctypes.data_as(C.c_void_p))
    assert rc == 0
    return out


def postorder(tree: Tree, table: Table):
    m = tree.n_parens // 2 + 1
    parent = np.zeros(m, np.int32)
    length = np.zeros(m)
    leaf = np.zeros(m, np.int32)
    row = np.zeros(m, np.int32)
    k = lib().orc_postorder(*tree.args(), C.c_int(table.n_obs), table._co,
                            parent.ctypes.data_as(C.c_void_p), length.ctypes.data_as(C.c_void_p),
                            leaf.ctypes.data_as(C.c_void_p), row.ctypes.data_as(C.c_void_p))
    return parent[:k], length[:k], leaf[:k], row[:k]


# ------------------------------------------------------ reference harness side

def ref_newick_to_tree(newick: str, cap: int | None = None) -> Tree:
    R = ref()
    cap = cap or (4 * len(newick) + 16)
    s = np.zeros(cap, np.uint8)
    le = np.zeros(cap)
    bufs = [C.create_string_buffer(256) for _ in range(cap)]
    arr = (C.c_char_p * cap)(*[C.cast(b, C.c_char_p) for b in bufs])
    k = R.ref_newick_to_bp(newick.encode(), C.c_int(cap), s.ctypes.data_as(C.c_void_p),
                           le.ctypes.data_as(C.c_void_p), arr, C.c_int(256))
    assert k >= 0
    return Tree(s[:k], le[:k], [bufs[i].value.decode() for i in range(k)])


def _ref_table_args(table: Table):
    return (C.c_int(table.n_obs), C.c_int(table.n_samples), table._co, table._cs,
            table.indptr.ctypes.data_as(C.c_void_p), table.indices.ctypes.data_as(C.c_void_p),
            table.data.ctypes.data_as(C.c_void_p))


def ref_stripes(tree: Tree, table: Table, method: str, fp32=False, variance_adjust=False, alpha=1.0,
                bypass_tips=False, normalize=True, start=0, stop=None) -> np.ndarray:
    n = table.n_samples
    if stop is None:
        stop = (n + 1) // 2
    out = np.zeros((stop - start, n))
    rc = ref().ref_unifrac_stripes(*tree.args(), *_ref_table_args(table), C.c_int(METHODS[method]),
                                   C.c_int(int(fp32)), C.c_int(int(variance_adjust)), C.c_double(alpha),
                                   C.c_int(int(bypass_tips)), C.c_int(int(normalize)), C.c_uint(start),
                                   C.c_uint(stop), out.ctypes.data_as(C.c_void_p))
    assert rc == 0
    return out


def ref_faith_pd(tree: Tree, table: Table) -> np.ndarray:
    out = np.zeros(table.n_samples)
    rc = ref().ref_faith_pd(*tree.args(), *_ref_table_args(table), out.ctypes.data_as(C.c_void_p))
    assert rc == 0
    return out


def ref_postorder(tree: Tree, table: Table):
    m = tree.n_parens // 2 + 1
    parent = np.zeros(m, np.int32)
    length = np.zeros(m)
    leaf = np.zeros(m, np.int32)
    row = np.zeros(m, np.int32)
    k = ref().ref_postorder(*tree.args(), *_ref_table_args(table), parent.ctypes.data_as(C.c_void_p),
                            length.ctypes.data_as(C.c_void_p), leaf.ctypes.data_as(C.c_void_p),
                            row.ctypes.data_as(C.c_void_p))
    return parent[:k], length[:k], leaf[:k], row[:k]


def ref_sample_counts(table: Table) -> np.ndarray:
    out = np.zeros(table.n_samples)
    ref().ref_sample_counts(*_ref_table_args(table), out.ctypes.data_as(C.c_void_p))
    return out


# ------------------------------------------------ distance-matrix statistics
# The published algorithms (scikit-bio's ``center_distance_matrix``, ``pcoa``
# and ``permanova``) that skbb implements for src/skbio_alt.cpp:144-233; the
# reference's vectors for them are in src/test_ska.cpp.


def center_matrix(m: np.ndarray) -> np.ndarray:
    """B = -1/2 (I - 11'/n) D∘D (I - 11'/n) (src/skbio_alt.hpp:22-27)."""
    a = -0.5 * np.asarray(m, np.float64) ** 2
    r = a.mean(axis=1)
    c = a.mean(axis=0)
    return a - r[:, None] - c[None, :] + a.mean()


def pcoa_exact(m: np.ndarray, dims: int):
    """Eigenvalues (largest first), coordinates v * sqrt(lambda) and proportion
    lambda / trace(B) from a full eigendecomposition of the centred matrix."""
    b = center_matrix(m)
    w, v = np.linalg.eigh(b)
    order = np.argsort(-np.abs(w))[:dims]
    w, v = w[order], v[:, order]
    return w, v * np.sqrt(np.maximum(w, 0.0)), w / np.trace(b)


def permanova_f(m: np.ndarray, grouping) -> float:
    """Pseudo-F over the upper triangle (scikit-bio ``_compute_f_stat``):
    s_T = sum d^2 / n, s_W = sum_g sum_{i<j in g} d^2 / n_g,
    F = ((s_T - s_W) / (G - 1)) / (s_W / (n - G))."""
    d = np.asarray(m, np.float64)
    g = np.asarray(grouping)
    n = d.shape[0]
    iu = np.triu_indices(n, 1)
    d2 = d[iu] ** 2
    s_t = d2.sum() / n
    same = g[iu[0]] == g[iu[1]]
    groups, sizes = np.unique(g, return_counts=True)
    size_of = dict(zip(groups.tolist(), sizes.tolist()))
    s_w = sum(d2[same & (g[iu[0]] == k)].sum() / size_of[k] for k in groups.tolist())
    ng = len(groups)
    return ((s_t - s_w) / (ng - 1)) / (s_w / (n - ng))
