#include "ssu_biom.hpp"

#include <dlfcn.h>

#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <vector>

namespace ssu {

namespace {

// Minimal HDF5 1.10 C ABI, resolved with dlsym (hid_t is int64_t since 1.10).
typedef int64_t hid_t;
typedef int herr_t;
typedef int htri_t;
typedef unsigned long long hsize_t;
constexpr hid_t P_DEFAULT = 0, S_ALL = 0;
constexpr unsigned F_ACC_RDONLY = 0u;
constexpr unsigned F_ACC_TRUNC = 2u;
constexpr size_t T_VARIABLE = (size_t)-1;

struct H5 {
    bool ok = false;
    std::string why;
    herr_t (*open)();
    hid_t (*Fopen)(const char*, unsigned, hid_t);
    herr_t (*Fclose)(hid_t);
    hid_t (*Dopen2)(hid_t, const char*, hid_t);
    herr_t (*Dclose)(hid_t);
    hid_t (*Dget_space)(hid_t);
    hid_t (*Dget_type)(hid_t);
    herr_t (*Dread)(hid_t, hid_t, hid_t, hid_t, hid_t, void*);
    herr_t (*Dvlen_reclaim)(hid_t, hid_t, hid_t, void*);
    int (*Sget_simple_extent_ndims)(hid_t);
    int (*Sget_simple_extent_dims)(hid_t, hsize_t*, hsize_t*);
    herr_t (*Sclose)(hid_t);
    htri_t (*Tis_variable_str)(hid_t);
    size_t (*Tget_size)(hid_t);
    hid_t (*Tcopy)(hid_t);
    herr_t (*Tset_size)(hid_t, size_t);
    herr_t (*Tclose)(hid_t);
    int (*Tget_cset)(hid_t);
    herr_t (*Tset_cset)(hid_t, int);
    herr_t (*Eset_auto2)(hid_t, void*, void*);
    hid_t (*Fcreate)(const char*, unsigned, hid_t, hid_t);
    hid_t (*Screate_simple)(int, const hsize_t*, const hsize_t*);
    hid_t (*Dcreate2)(hid_t, const char*, hid_t, hid_t, hid_t, hid_t, hid_t);
    herr_t (*Dwrite)(hid_t, hid_t, hid_t, hid_t, hid_t, const void*);
    hid_t* NATIVE_DOUBLE_g;
    hid_t* NATIVE_UINT32_g;
    hid_t* C_S1_g;
    hid_t* FORTRAN_S1_g;
    hid_t* IEEE_F64LE_g;
    hid_t* IEEE_F32LE_g;
    hid_t* STD_U32LE_g;
};

H5& h5() {
    static H5 h;
    static std::once_flag once;
    std::call_once(once, [] {
        const char* names[] = {"libhdf5.so", "libhdf5.so.103", "/opt/conda/lib/libhdf5.so"};
        void* lib = nullptr;
        for (const char* n : names)
            if ((lib = dlopen(n, RTLD_NOW | RTLD_LOCAL))) break;
        if (!lib) { h.why = "libhdf5 not found (needed for .biom files)"; return; }
        bool good = true;
        auto sym = [&](auto& fp, const char* name) {
            void* p = dlsym(lib, name);
            if (!p) good = false;
            fp = reinterpret_cast<std::remove_reference_t<decltype(fp)>>(p);
        };
        sym(h.open, "H5open"); sym(h.Fopen, "H5Fopen"); sym(h.Fclose, "H5Fclose");
        sym(h.Dopen2, "H5Dopen2"); sym(h.Dclose, "H5Dclose"); sym(h.Dget_space, "H5Dget_space");
        sym(h.Dget_type, "H5Dget_type"); sym(h.Dread, "H5Dread"); sym(h.Dvlen_reclaim, "H5Dvlen_reclaim");
        sym(h.Sget_simple_extent_ndims, "H5Sget_simple_extent_ndims");
        sym(h.Sget_simple_extent_dims, "H5Sget_simple_extent_dims"); sym(h.Sclose, "H5Sclose");
        sym(h.Tis_variable_str, "H5Tis_variable_str"); sym(h.Tget_size, "H5Tget_size");
        sym(h.Tcopy, "H5Tcopy"); sym(h.Tset_size, "H5Tset_size"); sym(h.Tclose, "H5Tclose");
        sym(h.Tget_cset, "H5Tget_cset"); sym(h.Tset_cset, "H5Tset_cset");
        sym(h.Eset_auto2, "H5Eset_auto2");
        sym(h.Fcreate, "H5Fcreate"); sym(h.Screate_simple, "H5Screate_simple"); sym(h.Dcreate2, "H5Dcreate2");
        sym(h.Dwrite, "H5Dwrite");
        sym(h.NATIVE_DOUBLE_g, "H5T_NATIVE_DOUBLE_g"); sym(h.NATIVE_UINT32_g, "H5T_NATIVE_UINT32_g");
        sym(h.C_S1_g, "H5T_C_S1_g"); sym(h.FORTRAN_S1_g, "H5T_FORTRAN_S1_g");
        sym(h.IEEE_F64LE_g, "H5T_IEEE_F64LE_g"); sym(h.IEEE_F32LE_g, "H5T_IEEE_F32LE_g");
        sym(h.STD_U32LE_g, "H5T_STD_U32LE_g");
        if (!good || h.open() < 0) { h.why = "libhdf5 lacks the 1.10 C API"; return; }
        h.Eset_auto2(0 /* H5E_DEFAULT */, nullptr, nullptr);  // silence the error stack printer
        h.ok = true;
    });
    return h;
}

struct Handle {
    hid_t id;
    herr_t (*close)(hid_t);
    ~Handle() { if (id >= 0 && close) close(id); }
};

bool dims1(H5& h, hid_t ds, hsize_t& n) {
    Handle sp{h.Dget_space(ds), h.Sclose};
    if (sp.id < 0 || h.Sget_simple_extent_ndims(sp.id) != 1) return false;
    return h.Sget_simple_extent_dims(sp.id, &n, nullptr) == 1;
}

template <class T>
bool read_numeric(H5& h, hid_t file, const char* path, hid_t memtype, std::vector<T>& out) {
    Handle ds{h.Dopen2(file, path, P_DEFAULT), h.Dclose};
    if (ds.id < 0) return false;
    hsize_t n = 0;
    if (!dims1(h, ds.id, n)) return false;
    out.resize(n);
    return n == 0 || h.Dread(ds.id, memtype, S_ALL, S_ALL, P_DEFAULT, out.data()) >= 0;
}

bool read_strings(H5& h, hid_t file, const char* path, std::vector<std::string>& out) {
    Handle ds{h.Dopen2(file, path, P_DEFAULT), h.Dclose};
    if (ds.id < 0) return false;
    hsize_t n = 0;
    if (!dims1(h, ds.id, n)) return false;
    out.clear();
    out.reserve(n);
    if (n == 0) return true;
    Handle ft{h.Dget_type(ds.id), h.Tclose};
    // the memory type takes the file's character set (ASCII or UTF-8): HDF5
    // converts between strings of one set only
    const int cset = h.Tget_cset(ft.id);
    if (h.Tis_variable_str(ft.id) > 0) {
        Handle mt{h.Tcopy(*h.C_S1_g), h.Tclose};
        h.Tset_size(mt.id, T_VARIABLE);
        if (cset >= 0) h.Tset_cset(mt.id, cset);
        std::vector<char*> buf(n, nullptr);
        if (h.Dread(ds.id, mt.id, S_ALL, S_ALL, P_DEFAULT, buf.data()) < 0) return false;
        for (auto* s : buf) out.emplace_back(s ? s : "");
        Handle sp{h.Dget_space(ds.id), h.Sclose};
        h.Dvlen_reclaim(mt.id, sp.id, P_DEFAULT, buf.data());
    } else {
        const size_t w = h.Tget_size(ft.id);
        Handle mt{h.Tcopy(*h.C_S1_g), h.Tclose};
        h.Tset_size(mt.id, w);
        if (cset >= 0) h.Tset_cset(mt.id, cset);
        std::vector<char> buf(n * w);
        if (h.Dread(ds.id, mt.id, S_ALL, S_ALL, P_DEFAULT, buf.data()) < 0) return false;
        for (hsize_t i = 0; i < n; ++i) out.emplace_back(std::string(&buf[i * w], strnlen(&buf[i * w], w)));
    }
    return true;
}

// one fixed-length (Fortran, unterminated) string, as src/api.cpp:1167-1189 stores it
bool write_string(H5& h, hid_t file, const char* name, const std::string& v) {
    Handle ft{h.Tcopy(*h.FORTRAN_S1_g), h.Tclose};
    Handle mt{h.Tcopy(*h.C_S1_g), h.Tclose};
    if (ft.id < 0 || mt.id < 0 || h.Tset_size(ft.id, v.size()) < 0 || h.Tset_size(mt.id, v.size() + 1) < 0)
        return false;
    const hsize_t one = 1;
    Handle sp{h.Screate_simple(1, &one, nullptr), h.Sclose};
    Handle ds{h.Dcreate2(file, name, ft.id, sp.id, P_DEFAULT, P_DEFAULT, P_DEFAULT), h.Dclose};
    return ds.id >= 0 && h.Dwrite(ds.id, mt.id, S_ALL, S_ALL, P_DEFAULT, v.c_str()) >= 0;
}

// variable-length C strings (src/api.cpp:1192-1216)
bool write_strings(H5& h, hid_t file, const char* name, uint32_t n, const char* const* v) {
    Handle t{h.Tcopy(*h.C_S1_g), h.Tclose};
    if (t.id < 0 || h.Tset_size(t.id, T_VARIABLE) < 0) return false;
    const hsize_t dims = n;
    Handle sp{h.Screate_simple(1, &dims, nullptr), h.Sclose};
    Handle ds{h.Dcreate2(file, name, t.id, sp.id, P_DEFAULT, P_DEFAULT, P_DEFAULT), h.Dclose};
    return ds.id >= 0 && (n == 0 || h.Dwrite(ds.id, t.id, S_ALL, S_ALL, P_DEFAULT, v) >= 0);
}

// n-dimensional numeric dataset (src/api.cpp:1218-1262)
bool write_array(H5& h, hid_t file, const char* name, int rank, const hsize_t* dims, hid_t type, const void* v) {
    Handle sp{h.Screate_simple(rank, dims, nullptr), h.Sclose};
    Handle ds{h.Dcreate2(file, name, type, sp.id, P_DEFAULT, P_DEFAULT, P_DEFAULT), h.Dclose};
    uint64_t count = 1;
    for (int r = 0; r < rank; ++r) count *= dims[r];
    return sp.id >= 0 && ds.id >= 0 && (count == 0 || h.Dwrite(ds.id, type, S_ALL, S_ALL, P_DEFAULT, v) >= 0);
}

}  // namespace

std::unique_ptr<H5Writer> H5Writer::create(const char* filename, std::string* err) {
    H5& h = h5();
    if (!h.ok) { if (err) *err = h.why; return nullptr; }
    std::unique_ptr<H5Writer> w(new H5Writer());
    w->file_ = h.Fcreate(filename, F_ACC_TRUNC, P_DEFAULT, P_DEFAULT);
    if (w->file_ < 0) { if (err) *err = std::string("cannot create ") + filename; return nullptr; }
    return w;
}

H5Writer::~H5Writer() {
    if (file_ >= 0) h5().Fclose(file_);
}

bool H5Writer::string(const char* name, const std::string& v) { return write_string(h5(), file_, name, v); }

bool H5Writer::strings(const char* name, uint64_t n, const char* const* v) {
    return write_strings(h5(), file_, name, (uint32_t)n, v);
}

bool H5Writer::reals(const char* name, uint64_t n, const void* v, bool fp64) {
    H5& h = h5();
    const hsize_t dims[1] = {n};
    return write_array(h, file_, name, 1, dims, fp64 ? *h.IEEE_F64LE_g : *h.IEEE_F32LE_g, v);
}

bool H5Writer::matrix(const char* name, uint64_t rows, uint64_t cols, const void* v, bool fp64) {
    H5& h = h5();
    const hsize_t dims[2] = {rows, cols};
    return write_array(h, file_, name, 2, dims, fp64 ? *h.IEEE_F64LE_g : *h.IEEE_F32LE_g, v);
}

bool H5Writer::u32s(const char* name, uint64_t n, const uint32_t* v) {
    H5& h = h5();
    const hsize_t dims[1] = {n};
    return write_array(h, file_, name, 1, dims, *h.STD_U32LE_g, v);
}

bool biom_n_samples(const char* filename, uint64_t& n, std::string* err) {
    H5& h = h5();
    if (!h.ok) { if (err) *err = h.why; return false; }
    Handle f{h.Fopen(filename, F_ACC_RDONLY, P_DEFAULT), h.Fclose};
    if (f.id < 0) { if (err) *err = std::string("cannot open BIOM file ") + filename; return false; }
    Handle ds{h.Dopen2(f.id, "/sample/ids", P_DEFAULT), h.Dclose};
    hsize_t m = 0;
    if (ds.id < 0 || !dims1(h, ds.id, m)) { if (err) *err = std::string("no /sample/ids in ") + filename; return false; }
    n = m;
    return true;
}

bool write_matrix_hdf5(const char* filename, uint32_t n, const char* const* ids, const void* matrix, bool fp64,
                       bool save_dist, std::string* err) {
    auto w = H5Writer::create(filename, err);
    if (!w) return false;
    bool ok = w->string("format", "BDSM") && w->string("version", "2020.12") && w->strings("order", n, ids);
    if (ok && save_dist) ok = w->matrix("matrix", n, n, matrix, fp64);
    if (!ok && err) *err = std::string("HDF5 write failed for ") + filename;
    return ok;
}

bool load_biom(const char* filename, Table& out, std::string* err) {
    H5& h = h5();
    if (!h.ok) { if (err) *err = h.why; return false; }
    Handle f{h.Fopen(filename, F_ACC_RDONLY, P_DEFAULT), h.Fclose};
    if (f.id < 0) { if (err) *err = std::string("cannot open BIOM file ") + filename; return false; }
    std::vector<std::string> oids, sids;
    std::vector<uint32_t> indptr, indices;
    std::vector<double> data;
    if (!read_strings(h, f.id, "/observation/ids", oids) || !read_strings(h, f.id, "/sample/ids", sids) ||
        !read_numeric(h, f.id, "/observation/matrix/indptr", *h.NATIVE_UINT32_g, indptr) ||
        !read_numeric(h, f.id, "/observation/matrix/indices", *h.NATIVE_UINT32_g, indices) ||
        !read_numeric(h, f.id, "/observation/matrix/data", *h.NATIVE_DOUBLE_g, data)) {
        if (err) *err = std::string("malformed BIOM file ") + filename;
        return false;
    }
    if (indptr.size() != oids.size() + 1 || indices.size() != data.size() ||
        (!indptr.empty() && indptr.back() > data.size())) {
        if (err) *err = std::string("inconsistent BIOM CSR in ") + filename;
        return false;
    }
    for (uint32_t s : indices)
        if (s >= sids.size()) { if (err) *err = "BIOM sample index out of range"; return false; }
    out = Table::own(std::move(oids), std::move(sids), std::move(indptr), std::move(indices), std::move(data));
    return true;
}

}  // namespace ssu
