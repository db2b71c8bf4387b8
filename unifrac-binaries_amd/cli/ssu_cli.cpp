// ssu: command-line front end over libssu.so, with the options and modes of
// the reference's src/su.cpp (one-off, partial, merge-partial, check-partial,
// partial-report, multi).  Output formats: ascii (the default) and, when the HDF5 C
// library can be loaded at run time, hdf5 / hdf5_fp32 / hdf5_fp64 /
// hdf5_nodist, with PCoA (--pcoa, default 10) and PERMANOVA (-g/-c/--permanova).
#include <glob.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <string>
#include <vector>

#include "ssu.h"

namespace {

void usage() {
    std::cout
        << "usage: ssu -i <biom> -o <out.dm> -m [METHOD] -t <newick> [-a alpha] [-f]  [--vaw]\n"
           "    [--mode MODE] [--start starting-stripe] [--stop stopping-stripe] [--partial-pattern <glob>]\n"
           "    [--n-partials number_of_partitions] [--report-bare] [--format|-r out-mode]\n"
           "    [--normalize-sample-counts true|false] [--n-substeps n] [--pcoa dims] [--diskbuf path]\n"
           "\n"
           "    -i\t\tThe input BIOM table.\n"
           "    -t\t\tThe input phylogeny in newick.\n"
           "    -m\t\tThe method, [unweighted | weighted_normalized | weighted_unnormalized | unweighted_unnormalized |\n"
           "      \t\t  generalized], each also with an _fp64 or _fp32 suffix.\n"
           "    -o\t\tThe output distance matrix.\n"
           "    -a\t\t[OPTIONAL] Generalized UniFrac alpha, default is 1.\n"
           "    -f\t\t[OPTIONAL] Bypass tips.\n"
           "    --vaw\t[OPTIONAL] Variance adjusted.\n"
           "    --mode\t[OPTIONAL] one-off (default) | partial | partial-report | merge-partial | check-partial | multi\n"
           "    --start\t[OPTIONAL] If mode==partial, the starting stripe.\n"
           "    --stop\t[OPTIONAL] If mode==partial, the stopping stripe.\n"
           "    --partial-pattern\t[OPTIONAL] If mode==merge-partial or check-partial, a glob pattern for partial outputs.\n"
           "    --n-partials\t[OPTIONAL] If mode==partial-report, the number of partitions to compute.\n"
           "    --report-bare\t[OPTIONAL] If mode==partial-report, produce barebones output.\n"
           "    --n-substeps\t[OPTIONAL] Accepted for compatibility; the device pipeline does not need it.\n"
           "    --normalize-sample-counts\t[OPTIONAL] true (default) or false (absolute quant mode).\n"
           "    --format|-r\t[OPTIONAL] ascii (default) | hdf5 | hdf5_fp32 | hdf5_fp64 | hdf5_nodist\n"
           "    --subsample-depth\t[OPTIONAL] Depth of subsampling of the input BIOM before computing UniFrac.\n"
           "    --subsample-replacement\t[OPTIONAL] with (default) or without.\n"
           "    --n-subsamples\t[OPTIONAL] If mode==multi, the number of subsampled matrices (default: 100).\n"
           "    -g\t\t[OPTIONAL] Grouping (TSV) file for PERMANOVA.\n"
           "    -c\t\t[OPTIONAL] The column(s) to use for grouping, multiple values comma separated.\n"
           "    --permanova\t[OPTIONAL] Number of PERMANOVA permutations (default: 999 with -g and -c, none if 0).\n"
           "    --pcoa\t[OPTIONAL] Number of PCoA dimensions (default: 10, none if 0).\n"
           "    --seed\t[OPTIONAL] Seed for the subsampling generator.\n"
           "    --diskbuf\t[OPTIONAL] Directory for a disk-backed result matrix.\n"
           "\n"
           "Runs on one AMD Instinct GPU (gfx950).\n";
}

// src/su.cpp:93-101: messages indexed by ComputeStatus
const char* compute_message(int st) {
    static const char* msgs[9] = {"No error.",
                                  "The tree file cannot be found.",
                                  "The table file cannot be found.",
                                  "The table file contains an empty table.",
                                  "An unknown method was requested.",
                                  "Table observation IDs are not a subset of the tree tips. This error can also be "
                                  "triggered if a node name contains a single quote (this is unlikely).",
                                  "Error creating the output.",
                                  "The requested method is not supported.",
                                  "The grouping file cannot be found or does not have the necessary data."};
    return (st >= 0 && st < 9) ? msgs[st] : "Unknown error.";
}

void err(const std::string& msg) {
    std::cerr << "ERROR: " << msg << std::endl << std::endl;
    usage();
}

struct Args {
    std::vector<std::string> tok;
    Args(int argc, char** argv) : tok(argv + 1, argv + argc) {}
    // the token after the flag, or "" (src/cmd.hpp semantics)
    std::string get(const char* flag) const {
        for (size_t i = 0; i + 1 < tok.size(); ++i)
            if (tok[i] == flag) return tok[i + 1];
        return "";
    }
    bool has(const char* flag) const {
        for (const auto& t : tok)
            if (t == flag) return true;
        return false;
    }
};

enum Format { format_invalid, format_ascii, format_hdf5_fp32, format_hdf5_fp64, format_hdf5_nodist };

bool fp64_method(const std::string& m) { return m.size() > 5 && m.compare(m.size() - 5, 5, "_fp64") == 0; }

// src/su.cpp:612-640
Format get_format(const std::string& f, const std::string& method, const std::string& mode) {
    if (f.empty()) return (mode != "multi" && mode != "multiple") ? format_ascii : format_hdf5_nodist;
    if (f == "ascii") return format_ascii;
    if (f == "hdf5_fp32") return format_hdf5_fp32;
    if (f == "hdf5_fp64") return format_hdf5_fp64;
    if (f == "hdf5_nodist") return format_hdf5_nodist;
    if (f == "hdf5") return fp64_method(method) ? format_hdf5_fp64 : format_hdf5_fp32;
    return format_invalid;
}

std::vector<std::string> glob_files(const std::string& pat) {
    glob_t g;
    std::vector<std::string> out;
    if (glob(pat.c_str(), GLOB_TILDE, nullptr, &g) == 0)
        for (size_t i = 0; i < g.gl_pathc; ++i) out.emplace_back(g.gl_pathv[i]);
    globfree(&g);
    return out;
}

int write_full(const std::string& out, Format fmt, unsigned pcoa, mat_full_fp64_t* m64, mat_full_fp32_t* m32) {
    IOStatus st;
    if (fmt == format_ascii) {
        st = m64 ? write_mat_from_matrix(out.c_str(), m64) : write_mat_from_matrix_fp32(out.c_str(), m32);
    } else {
        const int save_dist = fmt != format_hdf5_nodist;
        st = m64 ? write_mat_from_matrix_hdf5_fp64(out.c_str(), m64, pcoa, save_dist)
                 : write_mat_from_matrix_hdf5_fp32(out.c_str(), m32, pcoa, save_dist);
    }
    if (st != write_okay) {
        err("Unable to write; err " + std::to_string((int)st) + (ssu_last_error()[0] ? std::string(": ") + ssu_last_error() : ""));
        return EXIT_FAILURE;
    }
    return EXIT_SUCCESS;
}

const char* format_name(Format f) {
    switch (f) {
    case format_hdf5_fp32: return "hdf5_fp32";
    case format_hdf5_fp64: return "hdf5_fp64";
    case format_hdf5_nodist: return "hdf5_nodist";
    default: return "ascii";
    }
}

// src/su.cpp:399-470
int mode_one_off(const Args& a, Format fmt, const std::string& method, unsigned depth, bool with_replacement,
                 unsigned pcoa, unsigned perms, const std::string& grouping, const std::string& columns, bool vaw,
                 double alpha, bool bypass, bool normalize, unsigned nsub, const std::string& diskbuf) {
    const std::string table = a.get("-i"), tree = a.get("-t"), out = a.get("-o");
    if (out.empty()) { err("output filename missing"); return EXIT_FAILURE; }
    if (table.empty()) { err("table filename missing"); return EXIT_FAILURE; }
    if (tree.empty()) { err("tree filename missing"); return EXIT_FAILURE; }
    if (perms > 0 && grouping.empty()) { err("grouping filename missing"); return EXIT_FAILURE; }
    if (perms > 0 && columns.empty()) { err("grouping columns missing"); return EXIT_FAILURE; }
    if (method.empty()) { err("method missing"); return EXIT_FAILURE; }
    const char* mmap_dir = diskbuf.empty() ? nullptr : diskbuf.c_str();
    ComputeStatus st;
    if (fmt == format_ascii) {
        if (depth > 0) { err("Subsampling not supported with ASCII output."); return EXIT_FAILURE; }
        st = unifrac_to_txt_file_v3(table.c_str(), tree.c_str(), out.c_str(), method.c_str(), vaw, alpha, bypass,
                                    normalize, nsub, mmap_dir);
    } else {
        st = unifrac_to_file_v3(table.c_str(), tree.c_str(), out.c_str(), method.c_str(), vaw, alpha, bypass,
                                normalize, nsub, format_name(fmt), depth, with_replacement, pcoa, perms,
                                perms > 0 ? grouping.c_str() : nullptr, perms > 0 ? columns.c_str() : nullptr,
                                mmap_dir);
    }
    if (st != okay) fprintf(stderr, "Compute failed in one_off: %s\n", compute_message(st));
    return st == okay ? EXIT_SUCCESS : EXIT_FAILURE;
}

// src/su.cpp:472-545
int mode_multi(const Args& a, Format fmt, const std::string& method, unsigned n_subsamples, unsigned depth,
               bool with_replacement, unsigned pcoa, unsigned perms, const std::string& grouping,
               const std::string& columns, bool vaw, double alpha, bool bypass, bool normalize, unsigned nsub,
               const std::string& diskbuf) {
    const std::string table = a.get("-i"), tree = a.get("-t"), out = a.get("-o");
    if (out.empty()) { err("output filename missing"); return EXIT_FAILURE; }
    if (table.empty()) { err("table filename missing"); return EXIT_FAILURE; }
    if (tree.empty()) { err("tree filename missing"); return EXIT_FAILURE; }
    if (perms > 0 && grouping.empty()) { err("grouping filename missing"); return EXIT_FAILURE; }
    if (perms > 0 && columns.empty()) { err("grouping columns missing"); return EXIT_FAILURE; }
    if (method.empty()) { err("method missing"); return EXIT_FAILURE; }
    if (depth < 1) { err("subsample_depth cannot be 0."); return EXIT_FAILURE; }
    if (n_subsamples < 1) { err("n_subsamples cannot be 0."); return EXIT_FAILURE; }
    if (fmt == format_ascii) { err("ASCII format not supported in multi mode"); return EXIT_FAILURE; }
    const ComputeStatus st = unifrac_multi_to_file_v3(
        table.c_str(), tree.c_str(), out.c_str(), method.c_str(), vaw, alpha, bypass, normalize, nsub,
        format_name(fmt), n_subsamples, depth, with_replacement, pcoa, perms, perms > 0 ? grouping.c_str() : nullptr,
        perms > 0 ? columns.c_str() : nullptr, diskbuf.empty() ? nullptr : diskbuf.c_str());
    if (st != okay) fprintf(stderr, "Compute failed in multi: %s\n", compute_message(st));
    return st == okay ? EXIT_SUCCESS : EXIT_FAILURE;
}

// src/su.cpp:354-397
int mode_partial(const Args& a, const std::string& method, bool vaw, double alpha, bool bypass, bool normalize,
                 unsigned nsub, int start, int stop) {
    const std::string table = a.get("-i"), tree = a.get("-t"), out = a.get("-o");
    if (out.empty()) { err("output filename missing"); return EXIT_FAILURE; }
    if (table.empty()) { err("table filename missing"); return EXIT_FAILURE; }
    if (tree.empty()) { err("tree filename missing"); return EXIT_FAILURE; }
    if (method.empty()) { err("method missing"); return EXIT_FAILURE; }
    if (start < 0) { err("Starting stripe must be >= 0"); return EXIT_FAILURE; }
    if (stop <= start) {
        err("In '--mode partial', the stop and start stripes must be specified, and the stop stripe must be > start stripe");
        return EXIT_FAILURE;
    }
    partial_mat_t* result = nullptr;
    const ComputeStatus st = partial_v3(table.c_str(), tree.c_str(), method.c_str(), vaw, alpha, bypass, normalize, nsub,
                                        (unsigned)start, (unsigned)stop, &result);
    if (st != okay || !result) {
        fprintf(stderr, "Compute failed in partial: %s\n", compute_message(st));
        return EXIT_FAILURE;
    }
    const IOStatus io = write_partial(out.c_str(), result);
    destroy_partial_mat(&result);
    if (io != write_okay) {
        fprintf(stderr, "Write failed: %s\n", io == open_error ? "could not open output" : "unknown error");
        return EXIT_FAILURE;
    }
    return EXIT_SUCCESS;
}

bool load_partials(const std::string& pattern, std::vector<partial_dyn_mat_t*>& mats) {
    for (const auto& f : glob_files(pattern)) {
        partial_dyn_mat_t* m = nullptr;
        const IOStatus io = read_partial_header(f.c_str(), &m);
        if (io != read_okay) {
            err("Unable to parse file (" + f + "); err " + std::to_string((int)io));
            for (auto* p : mats) destroy_partial_dyn_mat(&p);
            mats.clear();
            return false;
        }
        mats.push_back(m);
    }
    return true;
}

const char* PATTERN_HELP =
    "Partial file pattern missing. For instance, if your partial results\n"
    "are named 'ssu.unweighted.start0.partial', 'ssu.unweighted.start10.partial', \n"
    "etc, then a pattern of 'ssu.unweighted.start*.partial' would make sense";

// src/su.cpp:239-300
int mode_merge_partial(const Args& a, Format fmt, unsigned pcoa, const std::string& diskbuf) {
    const std::string out = a.get("-o"), pattern = a.get("--partial-pattern");
    if (out.empty()) { err("output filename missing"); return EXIT_FAILURE; }
    if (pattern.empty()) { err(PATTERN_HELP); return EXIT_FAILURE; }
    std::vector<partial_dyn_mat_t*> mats;
    if (!load_partials(pattern, mats)) return EXIT_FAILURE;
    const char* mmap_dir = diskbuf.empty() ? nullptr : diskbuf.c_str();
    int rc = EXIT_FAILURE;
    MergeStatus ms;
    if (fmt == format_hdf5_fp32) {
        mat_full_fp32_t* m = nullptr;
        ms = merge_partial_to_mmap_matrix_fp32(mats.data(), (int)mats.size(), mmap_dir, &m);
        if (ms == merge_okay) rc = write_full(out, fmt, pcoa, nullptr, m);
        if (m) destroy_mat_full_fp32(&m);
    } else {
        mat_full_fp64_t* m = nullptr;
        ms = merge_partial_to_mmap_matrix(mats.data(), (int)mats.size(), mmap_dir, &m);
        if (ms == merge_okay) rc = write_full(out, fmt, pcoa, m, nullptr);
        if (m) destroy_mat_full_fp64(&m);
    }
    if (ms != merge_okay) err("Unable to complete merge; err " + std::to_string((int)ms));
    for (auto* p : mats) destroy_partial_dyn_mat(&p);
    return rc;
}

// src/su.cpp:302-352
int mode_check_partial(const Args& a) {
    const std::string pattern = a.get("--partial-pattern");
    if (pattern.empty()) { err(PATTERN_HELP); return EXIT_FAILURE; }
    const auto files = glob_files(pattern);
    std::cout << "Processing " << files.size() << " partial files." << std::endl;
    std::vector<partial_dyn_mat_t*> mats;
    if (!load_partials(pattern, mats)) return EXIT_FAILURE;
    const int status = validate_partial(mats.data(), (int)mats.size());
    if (status == merge_okay) {
        std::cout << "Partials are ready to be merged" << std::endl;
        const uint64_t n = mats[0]->n_samples;
        std::cout << "n_samples = " << n << std::endl;
        std::cout << "Matrix RAM needs: fp32 = " << (4.0 * n * n) / (1024.0 * 1024 * 1024)
                  << "GB, fp64 = " << (8.0 * n * n) / (1024.0 * 1024 * 1024) << "GB" << std::endl;
    }
    for (auto* p : mats) destroy_partial_dyn_mat(&p);
    return status;
}

// src/su.cpp:122-170: stripe ranges of near-equal size, the larger ones first
int mode_partial_report(const Args& a, unsigned npartials, bool bare) {
    const std::string table = a.get("-i");
    if (table.empty()) { err("table filename missing"); return EXIT_FAILURE; }
    if (npartials < 1) { err("--n-partials cannot be < 1"); return EXIT_FAILURE; }
    const long n_samples = ssu_biom_n_samples(table.c_str());
    if (n_samples < 0) {
        err(std::string("cannot read the sample ids of ") + table + ": " + ssu_last_error());
        return EXIT_FAILURE;
    }
    const unsigned total = (unsigned)((n_samples + 1) / 2);
    if (!bare) {
        std::cout << "Total samples: " << n_samples << std::endl;
        std::cout << "Total stripes: " << total << std::endl;
    }
    const unsigned full = (total + npartials - 1) / npartials, small = total / npartials;
    unsigned n_full = total % npartials;
    if (n_full == 0) n_full = npartials;
    unsigned start = 0;
    for (unsigned p = 0; p < npartials; ++p) {
        const unsigned stop = start + (p < n_full ? full : small);
        if (bare)
            std::cout << start << "\t" << stop << std::endl;
        else
            std::cout << "Partition " << p << ", suggested start and stop: " << start << ", " << stop << std::endl;
        start = stop;
    }
    return EXIT_SUCCESS;
}

}  // namespace

int main(int argc, char** argv) {
    const Args a(argc, argv);
    if (a.has("-h") || a.has("--help") || argc == 1) {
        usage();
        return EXIT_SUCCESS;
    }
    const std::string method = a.get("-m"), mode = a.get("--mode");
    const std::string nsub_s = a.get("--n-substeps"), alpha_s = a.get("-a");
    const unsigned nsub = nsub_s.empty() ? 1 : (unsigned)atoi(nsub_s.c_str());
    const double alpha = alpha_s.empty() ? 1.0 : atof(alpha_s.c_str());
    const bool vaw = a.has("--vaw"), bare = a.has("--report-bare"), bypass = a.has("-f");
    const std::string start_s = a.get("--start"), stop_s = a.get("--stop"), np_s = a.get("--n-partials");
    const int start = start_s.empty() ? 0 : atoi(start_s.c_str());
    const int stop = stop_s.empty() ? 0 : atoi(stop_s.c_str());
    const long npart = np_s.empty() ? 1 : atol(np_s.c_str());
    if (npart < 1) { err("--n-partials cannot be < 1"); return EXIT_FAILURE; }
    if (npart > 1000000000) { err("--n-partials cannot be > 1G"); return EXIT_FAILURE; }

    const std::string norm_s = a.get("--normalize-sample-counts");
    bool normalize = true;
    if (norm_s == "false") normalize = false;
    else if (!norm_s.empty() && norm_s != "true") {
        err("Invalid normalize-sample-counts, must be true or false");
        return EXIT_FAILURE;
    }
    std::string fmt_s = a.get("--format");
    if (fmt_s.empty()) fmt_s = a.get("-r");
    const Format fmt = get_format(fmt_s, method, mode);
    if (fmt == format_invalid) {
        err("Invalid format, must be one of ascii|hdf5|hdf5_fp32|hdf5_fp64|hdf5_nodist");
        return EXIT_FAILURE;
    }
    const std::string pcoa_s = a.get("--pcoa");
    const unsigned pcoa = pcoa_s.empty() ? 10 : (unsigned)atoi(pcoa_s.c_str());
    // src/su.cpp:700-711: 999 permutations by default when a grouping file and columns are given
    const std::string grouping = a.get("-g"), columns = a.get("-c"), perm_s = a.get("--permanova");
    const unsigned perms = !perm_s.empty() ? (unsigned)atoi(perm_s.c_str())
                                           : ((grouping.empty() || columns.empty()) ? 0u : 999u);
    const std::string seed_s = a.get("--seed");
    if (!seed_s.empty()) ssu_set_random_seed((unsigned)atoi(seed_s.c_str()));
    const bool multi = mode == "multi" || mode == "multiple";
    const std::string depth_s = a.get("--subsample-depth");
    if (depth_s.empty() && multi) { err("--subsample-depth required in multi mode."); return EXIT_FAILURE; }
    const unsigned depth = depth_s.empty() ? 0 : (unsigned)atoi(depth_s.c_str());
    const std::string nsubs_s = a.get("--n-subsamples");
    if (!nsubs_s.empty() && !multi) { err("--n-subsamples only allowed in multi mode."); return EXIT_FAILURE; }
    const unsigned n_subsamples = nsubs_s.empty() ? 100 : (unsigned)atoi(nsubs_s.c_str());
    const std::string repl = a.get("--subsample-replacement");
    bool with_replacement = true;
    if (repl == "without") with_replacement = false;
    else if (!repl.empty() && repl != "with") {
        err("Invalid --subsample-replacement argument, must be 'with' or 'without'.");
        return EXIT_FAILURE;
    }
    const std::string diskbuf = a.get("--diskbuf");

    if (mode.empty() || mode == "one-off")
        return mode_one_off(a, fmt, method, depth, with_replacement, pcoa, perms, grouping, columns, vaw, alpha, bypass,
                            normalize, nsub, diskbuf);
    if (mode == "partial") {
        if (depth > 0) { err("Cannot subsample in partial mode."); return EXIT_FAILURE; }
        return mode_partial(a, method, vaw, alpha, bypass, normalize, nsub, start, stop);
    }
    if (mode == "merge-partial") return mode_merge_partial(a, fmt, pcoa, diskbuf);
    if (mode == "check-partial") return mode_check_partial(a);
    if (mode == "partial-report") return mode_partial_report(a, (unsigned)npart, bare);
    if (multi)
        return mode_multi(a, fmt, method, n_subsamples, depth, with_replacement, pcoa, perms, grouping, columns, vaw,
                          alpha, bypass, normalize, nsub, diskbuf);
    err("Unknown mode. Valid options are: one-off, partial, merge-partial, check-partial, partial-report, multi");
    return EXIT_SUCCESS;
}
