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
    herr_t (*Eset_auto2)(hid_t, void*, void*);
    hid_t* NATIVE_DOUBLE_g;
    hid_t* NATIVE_UINT32_g;
    hid_t* C_S1_g;
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
        sym(h.Eset_auto2, "H5Eset_auto2");
        sym(h.NATIVE_DOUBLE_g, "H5T_NATIVE_DOUBLE_g"); sym(h.NATIVE_UINT32_g, "H5T_NATIVE_UINT32_g");
        sym(h.C_S1_g, "H5T_C_S1_g");
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
    if (h.Tis_variable_str(ft.id) > 0) {
        Handle mt{h.Tcopy(*h.C_S1_g), h.Tclose};
        h.Tset_size(mt.id, T_VARIABLE);
        std::vector<char*> buf(n, nullptr);
        if (h.Dread(ds.id, mt.id, S_ALL, S_ALL, P_DEFAULT, buf.data()) < 0) return false;
        for (auto* s : buf) out.emplace_back(s ? s : "");
        Handle sp{h.Dget_space(ds.id), h.Sclose};
        h.Dvlen_reclaim(mt.id, sp.id, P_DEFAULT, buf.data());
    } else {
        const size_t w = h.Tget_size(ft.id);
        Handle mt{h.Tcopy(*h.C_S1_g), h.Tclose};
        h.Tset_size(mt.id, w);
        std::vector<char> buf(n * w);
        if (h.Dread(ds.id, mt.id, S_ALL, S_ALL, P_DEFAULT, buf.data()) < 0) return false;
        for (hsize_t i = 0; i < n; ++i) out.emplace_back(std::string(&buf[i * w], strnlen(&buf[i * w], w)));
    }
    return true;
}

}  // namespace

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
