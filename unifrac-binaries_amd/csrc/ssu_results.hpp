// Host-side result containers shared by the C ABI translation units:
// sample-id copies, full matrices in malloc'd or mmap'd memory, last-error text.
#pragma once

#include <fcntl.h>
#include <sys/mman.h>
#include <unistd.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

namespace ssu {

constexpr uint32_t MMAP_FD_MASK = 0x0fff;  // src/api.cpp:28-29
constexpr uint32_t MMAP_FLAG = 0x1000;

// Records the message returned by ssu_last_error() and prints it to stderr.
void set_error(const std::string& e);
const std::string& last_error();

inline char** dup_ids(const std::vector<std::string>& ids) {
    char** out = (char**)malloc(sizeof(char*) * (ids.empty() ? 1 : ids.size()));
    for (size_t i = 0; i < ids.size(); ++i) out[i] = strdup(ids[i].c_str());
    return out;
}

// src/api.cpp:270-327 (initialize_mat_full_no_biom_T): malloc, or an unlinked mmap'd file under mmap_dir
template <class TReal, class TMat>
inline bool init_full(TMat*& r, const std::vector<std::string>& ids, const char* mmap_dir) {
    r = (TMat*)malloc(sizeof(TMat));
    r->n_samples = (uint32_t)ids.size();
    r->flags = 0;
    r->sample_ids = dup_ids(ids);
    const uint64_t n = r->n_samples;
    const uint64_t msize = sizeof(TReal) * n * n;
    if (mmap_dir && mmap_dir[0] == 0) mmap_dir = nullptr;
    if (!mmap_dir) {
        r->matrix = (TReal*)malloc(msize ? msize : sizeof(TReal));
    } else {
        std::string tmpl = std::string(mmap_dir) + "/su_mmap_XXXXXX";
        int fd = mkostemp(&tmpl[0], O_NOATIME);
        r->matrix = nullptr;
        if (fd >= 0) {
            unlink(tmpl.c_str());
            if (ftruncate(fd, (off_t)msize) == 0) {
                void* p = mmap(nullptr, msize ? msize : 1, PROT_READ | PROT_WRITE, MAP_SHARED | MAP_NORESERVE, fd, 0);
                if (p != MAP_FAILED) {
                    r->matrix = (TReal*)p;
                    r->flags = ((uint32_t)fd & MMAP_FD_MASK) | MMAP_FLAG;
                } else {
                    close(fd);
                }
            } else {
                close(fd);
            }
        }
    }
    return r->matrix != nullptr;
}

template <class TReal, class TMat>
inline void destroy_full(TMat** result) {
    if (!result || !*result) return;
    TMat* r = *result;
    for (uint32_t i = 0; i < r->n_samples; ++i) free(r->sample_ids[i]);
    free(r->sample_ids);
    if (r->matrix) {
        if ((r->flags & MMAP_FLAG) == 0) {
            free(r->matrix);
        } else {
            const uint64_t n = r->n_samples;
            munmap(r->matrix, (sizeof(TReal) * n * n) > 0 ? sizeof(TReal) * n * n : 1);
            close((int)(r->flags & MMAP_FD_MASK));
        }
    }
    free(r);
    *result = nullptr;
}

}  // namespace ssu
