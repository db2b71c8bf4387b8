"""ctypes binding of libssu.so (include/ssu.h).

Mirrors the reference C API (biocore/unifrac-binaries src/api.hpp) closely
enough that tests read like the reference's own C tests: the same entry point
names, argument meanings and status codes.  Loading fails loudly when the
library has not been built.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libssu.so")

# src/status_enum.hpp:4-6
OKAY, TREE_MISSING, TABLE_MISSING, TABLE_EMPTY, UNKNOWN_METHOD, TABLE_AND_TREE_DO_NOT_OVERLAP, OUTPUT_ERROR, \
    INVALID_METHOD, GROUPING_MISSING = range(9)
READ_OKAY, WRITE_OKAY, OPEN_ERROR, READ_ERROR, MAGIC_INCOMPATIBLE, BAD_HEADER, UNEXPECTED_END, WRITE_ERROR = range(8)
MERGE_OKAY, INCOMPLETE_STRIPE_SET, SAMPLE_ID_CONSISTENCY, SQUARE_MISMATCH, PARTIALS_MISMATCH, STRIPES_OVERLAP = range(6)


class SupportBiom(C.Structure):  # src/api.hpp:145-154
    _fields_ = [("obs_ids", C.POINTER(C.c_char_p)), ("sample_ids", C.POINTER(C.c_char_p)),
                ("indices", C.POINTER(C.c_uint32)), ("indptr", C.POINTER(C.c_uint32)),
                ("data", C.POINTER(C.c_double)), ("n_obs", C.c_int), ("n_samples", C.c_int), ("nnz", C.c_int)]


class SupportBPTree(C.Structure):  # src/api.hpp:163-168
    _fields_ = [("structure", C.POINTER(C.c_bool)), ("lengths", C.POINTER(C.c_double)),
                ("names", C.POINTER(C.c_char_p)), ("n_parens", C.c_int)]


class MatFullFP64(C.Structure):  # src/api.hpp:54-59
    _fields_ = [("n_samples", C.c_uint32), ("flags", C.c_uint32), ("matrix", C.POINTER(C.c_double)),
                ("sample_ids", C.POINTER(C.c_char_p))]


class MatFullFP32(C.Structure):  # src/api.hpp:67-72
    _fields_ = [("n_samples", C.c_uint32), ("flags", C.c_uint32), ("matrix", C.POINTER(C.c_float)),
                ("sample_ids", C.POINTER(C.c_char_p))]


class Mat(C.Structure):  # src/api.hpp:40-46
    _fields_ = [("n_samples", C.c_uint), ("cf_size", C.c_uint), ("is_upper_triangle", C.c_bool),
                ("condensed_form", C.POINTER(C.c_double)), ("sample_ids", C.POINTER(C.c_char_p))]


class RVec(C.Structure):  # src/api.hpp:82-86
    _fields_ = [("n_samples", C.c_uint), ("values", C.POINTER(C.c_double)), ("sample_ids", C.POINTER(C.c_char_p))]


class PartialMat(C.Structure):  # src/api.hpp:99-107
    _fields_ = [("n_samples", C.c_uint32), ("sample_ids", C.POINTER(C.c_char_p)),
                ("stripes", C.POINTER(C.POINTER(C.c_double))), ("stripe_start", C.c_uint32),
                ("stripe_stop", C.c_uint32), ("stripe_total", C.c_uint32), ("is_upper_triangle", C.c_bool)]


class PartialDynMat(C.Structure):  # src/api.hpp:122-132
    _fields_ = [("n_samples", C.c_uint32), ("sample_ids", C.POINTER(C.c_char_p)),
                ("offsets", C.POINTER(C.c_uint64)), ("stripes", C.POINTER(C.POINTER(C.c_double))),
                ("stripe_start", C.c_uint32), ("stripe_stop", C.c_uint32), ("stripe_total", C.c_uint32),
                ("is_upper_triangle", C.c_bool), ("filename", C.c_char_p)]


# every symbol include/ssu.h declares
EXPORTS = [
    "ssu_set_random_seed", "read_bptree_opaque", "load_bptree_opaque", "convert_bptree_opaque",
    "destroy_bptree_opaque", "get_bptree_opaque_els", "destroy_mat", "destroy_mat_full_fp64",
    "destroy_mat_full_fp32", "destroy_partial_mat", "destroy_partial_dyn_mat", "destroy_results_vec",
    "one_off_v3", "one_off_wtree_v3", "one_off", "one_off_wtree", "one_off_matrix_inmem_v3",
    "one_off_matrix_inmem_v2", "one_off_inmem", "one_off_matrix_inmem_fp32_v3", "one_off_matrix_inmem_fp32_v2",
    "one_off_inmem_fp32", "one_off_matrix_v3", "one_off_matrix_v3t", "one_off_matrix_v2", "one_off_matrix_v2t",
    "one_off_matrix", "one_off_matrix_fp32_v3", "one_off_matrix_fp32_v3t", "one_off_matrix_fp32_v2",
    "one_off_matrix_fp32_v2t", "one_off_matrix_fp32", "one_dense_pair_v3t", "one_dense_pair_v3",
    "one_dense_pair_v2t", "one_dense_pair_v2", "faith_pd_one_off", "partial_v3", "partial", "write_mat",
    "write_mat_from_matrix", "write_mat_from_matrix_fp32", "write_vec", "write_partial", "read_partial",
    "read_partial_header", "read_partial_one_stripe", "validate_partial", "merge_partial_to_mmap_matrix",
    "merge_partial_to_mmap_matrix_fp32", "write_mat_from_matrix_hdf5_fp64", "write_mat_from_matrix_hdf5_fp32",
    "ssu_faith_pd_inmem", "write_mat_from_matrix_hdf5_fp64_v2", "write_mat_from_matrix_hdf5_fp32_v2",
    "unifrac_to_file_v3", "unifrac_to_txt_file_v3", "unifrac_to_file_v2", "unifrac_to_file",
    "unifrac_multi_to_file_v3", "unifrac_multi_to_file_v2", "compute_permanova_fp64", "compute_permanova_fp32",
    "ssu_center_matrix_fp64", "ssu_center_matrix_fp32", "ssu_center_matrix_fp64_noalloc", "ssu_pcoa_fp64",
    "ssu_pcoa_fp32", "ssu_permanova_fp64",
    "ssu_permanova_fp32",
    "ssu_plan_create_inmem", "ssu_plan_execute", "ssu_plan_execute_stripes", "ssu_plan_copy_stripes",
    "ssu_plan_result_bytes", "ssu_plan_copy_matrix",
    "ssu_plan_copy_rows",
    "ssu_plan_timings", "ssu_plan_pair_nodes", "ssu_plan_kernel_terms", "ssu_plan_sparse_terms", "ssu_plan_refined_pairs", "ssu_plan_split_theta", "ssu_last_call_timings", "ssu_release_device_memory", "ssu_choose_split",
    "ssu_plan_main_kernel", "ssu_plan_destroy", "ssu_postorder",
    "ssu_newick_to_bp", "ssu_biom_n_samples", "ssu_device_available", "ssu_last_error",
    "ssu_rarefy_create_inmem", "ssu_rarefy_execute", "ssu_rarefy_n_samples", "ssu_rarefy_n_obs", "ssu_rarefy_nnz",
    "ssu_rarefy_kept", "ssu_rarefy_copy_counts", "ssu_rarefy_copy_matrix", "ssu_rarefy_last_ms",
    "ssu_rarefy_graph_used", "ssu_rarefy_pair_nodes", "ssu_rarefy_refined_pairs", "ssu_rarefy_destroy",
    "ssu_rarefy_group_create", "ssu_rarefy_group_execute", "ssu_rarefy_group_last_ms", "ssu_rarefy_group_destroy",
]

_lib = None


def lib():
    """The loaded libssu.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it first (__graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        for name in EXPORTS:
            getattr(L, name)  # fail loudly on a missing export
        status_fns = [n for n in EXPORTS if n.startswith(("one_off", "one_dense", "faith", "partial", "ssu_faith",
                                                          "ssu_plan_create", "ssu_plan_execute", "ssu_plan_copy"))]
        for n in status_fns:
            getattr(L, n).restype = C.c_int
        L.ssu_plan_pair_nodes.restype = C.c_uint64
        L.ssu_plan_kernel_terms.restype = C.c_uint64
        L.ssu_plan_sparse_terms.restype = C.c_uint64
        L.ssu_plan_refined_pairs.restype = C.c_uint64
        L.ssu_plan_split_theta.restype = C.c_double
        L.ssu_plan_main_kernel.restype = C.c_char_p
        L.ssu_last_error.restype = C.c_char_p
        L.ssu_postorder.restype = C.c_int
        L.ssu_newick_to_bp.restype = C.c_int
        L.get_bptree_opaque_els.restype = C.c_int
        L.read_bptree_opaque.restype = C.c_int
        for n in ("write_partial", "read_partial", "read_partial_header", "read_partial_one_stripe",
                  "validate_partial", "merge_partial_to_mmap_matrix", "merge_partial_to_mmap_matrix_fp32"):
            getattr(L, n).restype = C.c_int
        for n in ("compute_permanova_fp64", "compute_permanova_fp32", "ssu_center_matrix_fp64",
                  "ssu_center_matrix_fp32", "ssu_center_matrix_fp64_noalloc", "ssu_pcoa_fp64", "ssu_pcoa_fp32", "ssu_permanova_fp64",
                  "ssu_permanova_fp32"):
            getattr(L, n).restype = C.c_int
        for n in ("unifrac_to_file_v3", "unifrac_to_txt_file_v3", "unifrac_to_file_v2", "unifrac_to_file",
                  "unifrac_multi_to_file_v3", "unifrac_multi_to_file_v2", "write_mat_from_matrix_hdf5_fp64_v2",
                  "write_mat_from_matrix_hdf5_fp32_v2"):
            getattr(L, n).restype = C.c_int
        L.write_mat_from_matrix_hdf5_fp64.restype = C.c_int
        L.write_mat_from_matrix_hdf5_fp32.restype = C.c_int
        L.ssu_biom_n_samples.restype = C.c_long
        L.read_partial_one_stripe.argtypes = [C.POINTER(PartialDynMat), C.c_uint32]
        for n in ("ssu_rarefy_create_inmem", "ssu_rarefy_execute", "ssu_rarefy_copy_counts", "ssu_rarefy_copy_matrix",
                  "ssu_rarefy_graph_used", "ssu_rarefy_group_create", "ssu_rarefy_group_execute"):
            getattr(L, n).restype = C.c_int
        L.ssu_rarefy_group_last_ms.restype = C.c_double
        L.ssu_rarefy_n_samples.restype = C.c_uint
        L.ssu_rarefy_n_obs.restype = C.c_uint
        L.ssu_rarefy_nnz.restype = C.c_uint64
        L.ssu_rarefy_pair_nodes.restype = C.c_uint64
        L.ssu_rarefy_refined_pairs.restype = C.c_uint64
        L.ssu_rarefy_last_ms.restype = C.c_double
        _lib = L
    return _lib


def device_available() -> bool:
    return bool(lib().ssu_device_available())


def last_error() -> str:
    return lib().ssu_last_error().decode()


def _cstrs(strs):
    keep = [s.encode() if isinstance(s, str) else s for s in strs]
    arr = (C.c_char_p * max(1, len(keep)))(*keep)
    return arr, keep


class Inputs:
    """support_biom_t + support_bptree_t built from numpy arrays (kept alive here)."""

    def __init__(self, structure, lengths, names, obs_ids, sample_ids, indptr, indices, data):
        self.structure = np.ascontiguousarray(structure, dtype=np.bool_)
        self.lengths = np.ascontiguousarray(lengths, dtype=np.float64)
        self.names_c, self._k1 = _cstrs(list(names))
        self.obs_c, self._k2 = _cstrs(list(obs_ids))
        self.samp_c, self._k3 = _cstrs(list(sample_ids))
        self.sample_ids = list(sample_ids)
        self.obs_ids = list(obs_ids)
        self.indptr = np.ascontiguousarray(indptr, dtype=np.uint32)
        self.indices = np.ascontiguousarray(indices, dtype=np.uint32)
        self.data = np.ascontiguousarray(data, dtype=np.float64)
        self.tree = SupportBPTree(self.structure.ctypes.data_as(C.POINTER(C.c_bool)),
                                  self.lengths.ctypes.data_as(C.POINTER(C.c_double)), self.names_c,
                                  int(self.structure.size))
        self.table = SupportBiom(self.obs_c, self.samp_c, self.indices.ctypes.data_as(C.POINTER(C.c_uint32)),
                                 self.indptr.ctypes.data_as(C.POINTER(C.c_uint32)),
                                 self.data.ctypes.data_as(C.POINTER(C.c_double)), len(self.obs_ids),
                                 len(self.sample_ids), int(self.data.size))

    @property
    def n_samples(self):
        return len(self.sample_ids)


def matrix_inmem(inp: Inputs, method: str, variance_adjust=False, alpha=1.0, bypass_tips=False,
                 normalize=True, fp32=None, subsample_depth=0, with_replacement=True, mmap_dir=None):
    """one_off_matrix_inmem_v3 / _fp32_v3 -> (status, n x n ndarray or None)."""
    L = lib()
    if fp32 is None:
        fp32 = not method.endswith("_fp64")
    md = mmap_dir.encode() if mmap_dir else None
    if fp32:
        res = C.POINTER(MatFullFP32)()
        st = L.one_off_matrix_inmem_fp32_v3(C.byref(inp.table), C.byref(inp.tree), method.encode(),
                                            C.c_bool(variance_adjust), C.c_double(alpha), C.c_bool(bypass_tips),
                                            C.c_bool(normalize), C.c_uint(1), C.c_uint(subsample_depth),
                                            C.c_bool(with_replacement), md, C.byref(res))
        if st != OKAY:
            if res:   # allocated before the call failed
                L.destroy_mat_full_fp32(C.byref(res))
            return st, None
        n = res.contents.n_samples
        m = np.ctypeslib.as_array(res.contents.matrix, shape=(n * n,)).reshape(n, n).copy()
        L.destroy_mat_full_fp32(C.byref(res))
    else:
        res = C.POINTER(MatFullFP64)()
        st = L.one_off_matrix_inmem_v3(C.byref(inp.table), C.byref(inp.tree), method.encode(),
                                       C.c_bool(variance_adjust), C.c_double(alpha), C.c_bool(bypass_tips),
                                       C.c_bool(normalize), C.c_uint(1), C.c_uint(subsample_depth),
                                       C.c_bool(with_replacement), md, C.byref(res))
        if st != OKAY:
            if res:   # allocated before the call failed
                L.destroy_mat_full_fp64(C.byref(res))
            return st, None
        n = res.contents.n_samples
        m = np.ctypeslib.as_array(res.contents.matrix, shape=(n * n,)).reshape(n, n).copy()
        L.destroy_mat_full_fp64(C.byref(res))
    return st, m


def choose_split(kind, n_samples, occupancy):
    """(threshold, modelled ms, modelled ms without a split) of the occupancy split.
    kind: 0 FP32 weighted, 1 FP64 weighted, 2 FP32 generalized, 3 unweighted Gram, 4 FP64 generalized,
    5 FP32 VAW, 6 FP64 VAW."""
    L = lib()
    L.ssu_choose_split.restype = C.c_uint
    occ = np.ascontiguousarray(occupancy, dtype=np.uint32)
    a, b = C.c_double(), C.c_double()
    thr = L.ssu_choose_split(C.c_int(kind), C.c_uint(n_samples), occ.ctypes.data_as(C.c_void_p), C.c_uint(occ.size),
                             C.byref(a), C.byref(b))
    return int(thr), a.value, b.value


def last_call_timings():
    """Wall-clock ms of this thread's last one_off_matrix_inmem* call, by phase."""
    ms = (C.c_double * 6)()
    k = lib().ssu_last_call_timings(ms, 6)
    keys = ["input", "tree", "plan", "execute", "copy", "total"]
    return {keys[i]: ms[i] for i in range(k)}


def faith_pd_inmem(inp: Inputs):
    L = lib()
    res = C.POINTER(RVec)()
    st = L.ssu_faith_pd_inmem(C.byref(inp.table), C.byref(inp.tree), C.byref(res))
    if st != OKAY:
        return st, None, None
    n = res.contents.n_samples
    vals = np.ctypeslib.as_array(res.contents.values, shape=(n,)).copy()
    ids = [res.contents.sample_ids[i].decode() for i in range(n)]
    L.destroy_results_vec(C.byref(res))
    return st, vals, ids


class Plan:
    """Device-resident plan (extension): upload once, execute, time, copy."""

    def __init__(self, inp: Inputs, method: str, variance_adjust=False, alpha=1.0, bypass_tips=False,
                 normalize=True, plain_kernel=0):
        L = lib()
        self._inp = inp
        self.h = C.c_void_p()
        st = L.ssu_plan_create_inmem(C.byref(inp.table), C.byref(inp.tree), method.encode(),
                                     C.c_bool(variance_adjust), C.c_double(alpha), C.c_bool(bypass_tips),
                                     C.c_bool(normalize), C.c_int(plain_kernel), C.byref(self.h))
        if st != OKAY:
            raise RuntimeError(f"ssu_plan_create_inmem failed: status {st}: {last_error()}")
        self.n = inp.n_samples

    def execute(self):
        st = lib().ssu_plan_execute(self.h)
        if st != OKAY:
            raise RuntimeError(f"ssu_plan_execute failed: {last_error()}")

    def execute_stripes(self, start, stop):
        """compute only the pairs of stripes [start, stop)"""
        st = lib().ssu_plan_execute_stripes(self.h, C.c_uint(start), C.c_uint(stop))
        if st != OKAY:
            raise RuntimeError(f"ssu_plan_execute_stripes failed: {last_error()}")

    def stripes(self, start, stop):
        """stripes [start, stop) of the result, (stop - start) x n doubles"""
        out = np.zeros((stop - start, self.n))
        st = lib().ssu_plan_copy_stripes(self.h, C.c_uint(start), C.c_uint(stop), out.ctypes.data_as(C.c_void_p))
        if st != OKAY:
            raise RuntimeError(last_error())
        return out

    def timings(self):
        ms = (C.c_double * 8)()
        k = lib().ssu_plan_timings(self.h, ms, 8)
        keys = ["embed", "prep", "main", "combine", "total", "sparse"]
        return {keys[i]: ms[i] for i in range(min(k, 6))}

    def pair_nodes(self) -> int:
        return int(lib().ssu_plan_pair_nodes(self.h))

    def kernel_terms(self) -> int:
        return int(lib().ssu_plan_kernel_terms(self.h))

    def sparse_terms(self) -> int:
        return int(lib().ssu_plan_sparse_terms(self.h))

    def refined_pairs(self) -> int:
        return int(lib().ssu_plan_refined_pairs(self.h))

    def split_theta(self) -> float:
        return float(lib().ssu_plan_split_theta(self.h))

    def main_kernel(self) -> str:
        return lib().ssu_plan_main_kernel(self.h).decode()

    def matrix(self):
        L = lib()
        eb = L.ssu_plan_result_bytes(self.h)
        out = np.empty((self.n, self.n), dtype=np.float64 if eb == 8 else np.float32)
        st = L.ssu_plan_copy_matrix(self.h, out.ctypes.data_as(C.c_void_p))
        if st != OKAY:
            raise RuntimeError(last_error())
        return out

    def rows(self, row0, nrows):
        """rows [row0, row0 + nrows) of the result"""
        L = lib()
        eb = L.ssu_plan_result_bytes(self.h)
        out = np.empty((nrows, self.n), dtype=np.float64 if eb == 8 else np.float32)
        st = L.ssu_plan_copy_rows(self.h, C.c_uint(row0), C.c_uint(nrows), out.ctypes.data_as(C.c_void_p))
        if st != OKAY:
            raise RuntimeError(last_error())
        return out

    def close(self):
        if self.h:
            lib().ssu_plan_destroy(C.byref(self.h))
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _mat_ptr(m, dtype):
    a = np.ascontiguousarray(m, dtype=dtype)
    return a, a.ctypes.data_as(C.c_void_p)


def center_matrix(m, fp32=False):
    """Gower-centred matrix on the device (ssu_center_matrix_fp64 / _fp32)."""
    dt = np.float32 if fp32 else np.float64
    a, pa = _mat_ptr(m, dt)
    out = np.empty_like(a)
    fn = lib().ssu_center_matrix_fp32 if fp32 else lib().ssu_center_matrix_fp64
    if fn(pa, C.c_uint(a.shape[0]), out.ctypes.data_as(C.c_void_p)) != OKAY:
        raise RuntimeError(last_error())
    return out


def pcoa(m, dims, fp32=False):
    """(eigenvalues, samples n x dims, proportion_explained) by randomized FSVD on the device."""
    dt = np.float32 if fp32 else np.float64
    a, pa = _mat_ptr(m, dt)
    n = a.shape[0]
    e = np.zeros(dims, dt)
    s = np.zeros((n, dims), dt)
    p = np.zeros(dims, dt)
    fn = lib().ssu_pcoa_fp32 if fp32 else lib().ssu_pcoa_fp64
    if fn(pa, C.c_uint(n), C.c_uint(dims), e.ctypes.data_as(C.c_void_p), s.ctypes.data_as(C.c_void_p),
          p.ctypes.data_as(C.c_void_p)) != OKAY:
        raise RuntimeError(last_error())
    return e, s, p


def permanova(m, grouping, n_perm=999, fp32=False):
    """(pseudo-F, p-value) on the device."""
    dt = np.float32 if fp32 else np.float64
    a, pa = _mat_ptr(m, dt)
    g = np.ascontiguousarray(grouping, dtype=np.uint32)
    f = (C.c_float if fp32 else C.c_double)()
    p = (C.c_float if fp32 else C.c_double)()
    fn = lib().ssu_permanova_fp32 if fp32 else lib().ssu_permanova_fp64
    if fn(pa, C.c_uint(a.shape[0]), g.ctypes.data_as(C.c_void_p), C.c_uint(n_perm), C.byref(f), C.byref(p)) != OKAY:
        raise RuntimeError(last_error())
    return f.value, p.value


class Rarefaction:
    """Device rarefaction replicates (extension): n_reps subsampled tables of
    one input at one depth, each through every listed method; the first
    execute runs eagerly and captures a HIP graph that later executes replay."""

    def __init__(self, inp: Inputs, methods, depth, n_reps, with_replacement=True, seed=0, variance_adjust=False,
                 alpha=1.0, normalize=True):
        L = lib()
        self._inp = inp
        self.methods = list(methods)
        arr, self._keep = _cstrs(self.methods)
        self.h = C.c_void_p()
        st = L.ssu_rarefy_create_inmem(C.byref(inp.table), C.byref(inp.tree), arr, C.c_int(len(self.methods)),
                                       C.c_bool(variance_adjust), C.c_double(alpha), C.c_bool(normalize),
                                       C.c_uint(depth), C.c_bool(with_replacement), C.c_uint64(seed),
                                       C.c_uint(n_reps), C.byref(self.h))
        if st != OKAY:
            raise RuntimeError(f"ssu_rarefy_create_inmem failed: status {st}: {last_error()}")
        self.n = int(L.ssu_rarefy_n_samples(self.h))
        self.n_obs = int(L.ssu_rarefy_n_obs(self.h))
        self.nnz = int(L.ssu_rarefy_nnz(self.h))
        self.n_reps = n_reps

    def execute(self):
        if lib().ssu_rarefy_execute(self.h) != OKAY:
            raise RuntimeError(f"ssu_rarefy_execute failed: {last_error()}")

    def kept(self):
        """(sample_index, obs_index, indptr, indices) of the kept table."""
        si = np.zeros(self.n, np.uint32)
        oi = np.zeros(self.n_obs, np.uint32)
        ip = np.zeros(self.n_obs + 1, np.uint32)
        ix = np.zeros(max(self.nnz, 1), np.uint32)
        lib().ssu_rarefy_kept(self.h, *(a.ctypes.data_as(C.c_void_p) for a in (si, oi, ip, ix)))
        return si, oi, ip, ix[:self.nnz]

    def counts(self, rep):
        out = np.zeros(max(self.nnz, 1))
        if lib().ssu_rarefy_copy_counts(self.h, C.c_uint(rep), out.ctypes.data_as(C.c_void_p)) != OKAY:
            raise RuntimeError(last_error())
        return out[:self.nnz]

    def matrix(self, method, rep):
        k = self.methods.index(method) if isinstance(method, str) else int(method)
        out = np.empty((self.n, self.n), np.float64 if self.methods[k].endswith("_fp64") else np.float32)
        if lib().ssu_rarefy_copy_matrix(self.h, C.c_int(k), C.c_uint(rep), out.ctypes.data_as(C.c_void_p)) != OKAY:
            raise RuntimeError(last_error())
        return out

    def last_ms(self) -> float:
        return float(lib().ssu_rarefy_last_ms(self.h))

    def graph_used(self) -> bool:
        return bool(lib().ssu_rarefy_graph_used(self.h))

    def pair_nodes(self) -> int:
        return int(lib().ssu_rarefy_pair_nodes(self.h))

    def refined_pairs(self) -> int:
        return int(lib().ssu_rarefy_refined_pairs(self.h))

    def close(self):
        if self.h:
            lib().ssu_rarefy_destroy(C.byref(self.h))
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class RarefactionGroup:
    """Several Rarefaction batches in one HIP graph launch (extension): each
    batch's captured graph is a child node of one parent graph, so the tables
    run side by side.  The batches must stay open while the group is."""

    def __init__(self, batches):
        self.batches = list(batches)
        arr = (C.c_void_p * len(self.batches))(*[b.h.value for b in self.batches])
        self.h = C.c_void_p()
        st = lib().ssu_rarefy_group_create(arr, C.c_int(len(self.batches)), C.byref(self.h))
        if st != OKAY:
            raise RuntimeError(f"ssu_rarefy_group_create failed: status {st}: {last_error()}")

    def execute(self):
        if lib().ssu_rarefy_group_execute(self.h) != OKAY:
            raise RuntimeError(f"ssu_rarefy_group_execute failed: {last_error()}")

    def last_ms(self) -> float:
        return float(lib().ssu_rarefy_group_last_ms(self.h))

    def close(self):
        if self.h:
            lib().ssu_rarefy_group_destroy(C.byref(self.h))
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def postorder(structure, lengths, names, obs_ids):
    """Host-only sheared/collapsed postorder arrays (no device needed)."""
    L = lib()
    st = np.ascontiguousarray(structure, dtype=np.bool_)
    le = np.ascontiguousarray(lengths, dtype=np.float64)
    nm, k1 = _cstrs(list(names))
    tr = SupportBPTree(st.ctypes.data_as(C.POINTER(C.c_bool)), le.ctypes.data_as(C.POINTER(C.c_double)), nm,
                       int(st.size))
    ob, k2 = _cstrs(list(obs_ids))
    m = st.size // 2 + 1
    parent = np.zeros(m, np.int32)
    length = np.zeros(m)
    leaf = np.zeros(m, np.int32)
    row = np.zeros(m, np.int32)
    k = L.ssu_postorder(C.byref(tr), C.c_int(len(obs_ids)), ob, parent.ctypes.data_as(C.c_void_p),
                        length.ctypes.data_as(C.c_void_p), leaf.ctypes.data_as(C.c_void_p),
                        row.ctypes.data_as(C.c_void_p))
    if k < 0:
        raise ValueError("invalid tree")
    return parent[:k], length[:k], leaf[:k], row[:k]


def newick_to_bp(newick: str):
    L = lib()
    cap = 4 * len(newick) + 16
    s = np.zeros(cap, np.uint8)
    le = np.zeros(cap)
    bufs = [C.create_string_buffer(256) for _ in range(cap)]
    arr = (C.c_char_p * cap)(*[C.cast(b, C.c_char_p) for b in bufs])
    k = L.ssu_newick_to_bp(newick.encode(), C.c_int(cap), s.ctypes.data_as(C.c_void_p),
                           le.ctypes.data_as(C.c_void_p), arr, C.c_int(256))
    if k < 0:
        raise ValueError("newick too large for buffer")
    return s[:k].astype(bool), le[:k], [bufs[i].value.decode() for i in range(k)]
