// faithpd: Faith's phylogenetic diversity of every sample of a BIOM table
// (the reference's src/faithpd.cpp front end) over libssu.so.
#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <string>
#include <vector>

#include "ssu.h"

namespace {

void usage() {
    std::cout << "usage: faithpd -i <biom> -t <newick> -o <out.txt>\n"
                 "\n"
                 "    -i\t\tThe input BIOM table.\n"
                 "    -t\t\tThe input phylogeny in newick.\n"
                 "    -o\t\tThe output series.\n"
                 "\n"
                 "Runs on one AMD Instinct GPU (gfx950).\n";
}

void err(const std::string& msg) {
    std::cerr << "ERROR: " << msg << std::endl << std::endl;
    usage();
}

// src/faithpd.cpp:23-29: messages indexed by ComputeStatus
const char* compute_message(int st) {
    static const char* msgs[7] = {"No error.",
                                  "The tree file cannot be found.",
                                  "The table file cannot be found.",
                                  "The table file contains an empty table.",
                                  "An unknown method was requested.",
                                  "Table observation IDs are not a subset of the tree tips. This error can also be "
                                  "triggered if a node name contains a single quote (this is unlikely).",
                                  "Error creating the output."};
    return (st >= 0 && st < 7) ? msgs[st] : "Unknown error.";
}

std::string opt(const std::vector<std::string>& tok, const char* flag) {
    for (size_t i = 0; i + 1 < tok.size(); ++i)
        if (tok[i] == flag) return tok[i + 1];
    return "";
}

}  // namespace

int main(int argc, char** argv) {
    const std::vector<std::string> tok(argv + 1, argv + argc);
    for (const auto& t : tok)
        if (t == "-h" || t == "--help") { usage(); return EXIT_SUCCESS; }
    if (argc == 1) { usage(); return EXIT_SUCCESS; }
    const std::string table = opt(tok, "-i"), tree = opt(tok, "-t"), out = opt(tok, "-o");
    if (out.empty()) { err("output filename missing"); return EXIT_FAILURE; }
    if (table.empty()) { err("table filename missing"); return EXIT_FAILURE; }
    if (tree.empty()) { err("tree filename missing"); return EXIT_FAILURE; }
    r_vec* result = nullptr;
    const ComputeStatus st = faith_pd_one_off(table.c_str(), tree.c_str(), &result);
    if (st != okay || !result) {
        fprintf(stderr, "Compute failed in faith_pd_one_off: %s\n", compute_message(st));
        return EXIT_FAILURE;
    }
    const IOStatus io = write_vec(out.c_str(), result);
    destroy_results_vec(&result);
    return io == write_okay ? EXIT_SUCCESS : EXIT_FAILURE;
}
