#include "ssu_tree.hpp"

#include <cctype>
#include <cstdlib>
#include <cstring>
#include <algorithm>
#include <mutex>
#include <thread>

namespace ssu {

namespace {

bool structure_char(char c) { return c == '(' || c == ')' || c == ',' || c == ';'; }

// Parentheses from newick text, src/tree.cpp:316-363.  Text inside quotes is
// skipped.  "(x)" and a last child after ',' each add a tip "10" ahead of the
// closing ")"; a ")" that follows ")" only closes.
std::vector<uint8_t> newick_parens(const char* nw) {
    std::vector<uint8_t> s;
    s.reserve(strlen(nw) / 2 + 16);
    char last = '\0';
    bool maybe_single = false, quoted = false;
    for (const char* p = nw; *p; ++p) {
        const char c = *p;
        if (c == '\'') quoted = !quoted;
        if (quoted) continue;
        switch (c) {
        case '(':
            s.push_back(1);
            last = c;
            maybe_single = true;
            break;
        case ')':
            if (maybe_single || last == ',') {
                s.push_back(1); s.push_back(0); s.push_back(0);
                maybe_single = false;
            } else {
                s.push_back(0);
            }
            last = c;
            break;
        case ',':
            if (last != ')') { s.push_back(1); s.push_back(0); }
            maybe_single = false;
            last = c;
            break;
        default:
            break;
        }
    }
    return s;
}

// Next newick token, src/tree.cpp:455-490.  Newlines vanish, quote marks are
// dropped, and a structure character that ends a non-empty token is eaten
// together with it.
std::string next_token(const char*& cur, const char* end) {
    std::string tok;
    bool quoted = false;
    do {
        const char c = *cur++;
        if (c == '\n') continue;
        const bool q = (c == '\'');
        if (quoted && q) { quoted = false; continue; }
        if (!quoted && q) { quoted = true; continue; }
        if (structure_char(c) && !quoted) {
            if (tok.empty()) tok.push_back(c);
            break;
        }
        tok.push_back(c);
    } while (cur != end);
    return tok;
}

// Branch lengths go through a float parse, as std::stof does in src/tree.cpp:440-449.
double parse_length(const std::string& s) { return (double)std::strtof(s.c_str(), nullptr); }

}  // namespace

void PhyloTree::index() {
    const uint32_t n = nparens();
    partner.assign(n, 0);
    parent_open.assign(n, -1);
    std::vector<uint32_t> stack;
    stack.reserve(64);
    valid_ = true;
    for (uint32_t i = 0; i < n; ++i) {
        if (open[i]) {
            parent_open[i] = stack.empty() ? -1 : (int32_t)stack.back();
            stack.push_back(i);
        } else {
            if (stack.empty()) { valid_ = false; return; }
            const uint32_t o = stack.back();
            stack.pop_back();
            partner[i] = o;
            partner[o] = i;
            parent_open[i] = parent_open[o];
        }
    }
    if (!stack.empty()) valid_ = false;
}

PhyloTree PhyloTree::from_newick(const char* nw) {
    PhyloTree t;
    t.open = newick_parens(nw);
    const uint32_t n = t.nparens();
    t.length.assign(n, 0.0);
    t.name.assign(n, std::string());
    t.index();
    if (!t.valid_) return t;

    // Token -> node assignment, src/tree.cpp:384-425.  structure_idx walks the
    // parentheses; "lag" re-synchronises after ")" runs whose characters the
    // tokenizer swallowed.
    const char* cur = nw;
    const char* end = nw + strlen(nw);
    while (end != cur && std::isspace((unsigned char)end[-1])) --end;
    char last = '\0';
    uint32_t idx = 0, lag = 0;
    while (cur != end) {
        std::string tok = next_token(cur, end);
        if (tok.size() == 1 && structure_char(tok[0])) {
            if (tok[0] == '(') {
                ++idx;
            } else if (tok[0] == ')' || tok[0] == ',') {
                ++idx;
                if (last == ')') ++lag;
            }
        } else {
            idx += lag;
            lag = 0;
            if (idx < n) {
                const uint32_t o = t.open_of(idx);
                // src/tree.cpp:432-448: name before the last ':', float length after it
                const size_t colon = tok.find_last_of(':');
                if (colon == 0) {
                    t.length[o] = parse_length(tok.substr(1));
                    t.name[o].clear();
                } else if (colon != std::string::npos) {
                    t.name[o] = tok.substr(0, colon);
                    t.length[o] = parse_length(tok.substr(colon + 1));
                } else {
                    t.name[o] = tok;
                    t.length[o] = 0.0;
                }
                idx += t.is_leaf(o) ? 2 : 1;
            }
        }
        last = tok.empty() ? '\0' : tok[0];
    }
    return t;
}

PhyloTree PhyloTree::from_arrays(const bool* structure, const double* lengths, const char* const* names, int n) {
    PhyloTree t;
    t.open.resize(n > 0 ? n : 0);
    t.length.resize(t.open.size());
    t.name.resize(t.open.size());
    for (int i = 0; i < n; ++i) {
        t.open[i] = structure[i] ? 1 : 0;
        t.length[i] = lengths[i];
        t.name[i] = names[i] ? names[i] : "";
    }
    t.index();
    return t;
}

PhyloTree PhyloTree::from_vectors(std::vector<uint8_t> o, std::vector<double> l, std::vector<std::string> nm) {
    PhyloTree t;
    t.open = std::move(o);
    t.length = std::move(l);
    t.name = std::move(nm);
    t.index();
    return t;
}

std::unordered_set<std::string> PhyloTree::tip_names() const {
    std::unordered_set<std::string> s;
    for (uint32_t i = 0; i < nparens(); ++i)
        if (is_leaf(i)) s.insert(name[i]);
    return s;
}

size_t PhyloTree::n_tips() const { return tip_names().size(); }

// src/tree.cpp:126-143 (mask)
PhyloTree PhyloTree::masked(const std::vector<uint8_t>& keep, const std::vector<double>& lengths) const {
    std::vector<uint8_t> o;
    std::vector<double> l;
    std::vector<std::string> nm;
    size_t c = 0;
    for (uint8_t k : keep) c += k;
    o.reserve(c); l.reserve(c); nm.reserve(c);
    for (uint32_t i = 0; i < nparens(); ++i)
        if (keep[i]) { o.push_back(open[i]); l.push_back(lengths[i]); nm.push_back(name[i]); }
    return from_vectors(std::move(o), std::move(l), std::move(nm));
}

PhyloTree PhyloTree::shear(const std::unordered_set<std::string>& keep) const {
    std::vector<uint8_t> mask(nparens(), 0);
    for (uint32_t i = 0; i < nparens(); ++i) {
        if (is_leaf(i) && keep.count(name[i])) {
            mask[i] = 1;
            mask[i + 1] = 1;
            int32_t p = parent_open[i];
            while (p != -1 && !mask[p]) {
                mask[p] = 1;
                mask[partner[p]] = 1;
                p = parent_open[p];
            }
        }
    }
    return masked(mask, length);
}

PhyloTree PhyloTree::collapse() const {
    std::vector<uint8_t> mask(nparens(), 0);
    std::vector<double> nl(length);
    for (uint32_t cur = 0; cur < nparens(); ++cur) {
        if (!open[cur]) continue;  // preorder visits open positions left to right
        if (is_leaf(cur) || cur == 0) {
            mask[cur] = 1;
            mask[partner[cur]] = 1;
        } else {
            const uint32_t first = cur + 1;
            const uint32_t last = open_of(partner[cur] - 1);
            if (first == last) {
                nl[first] = nl[first] + nl[cur];
            } else {
                mask[cur] = 1;
                mask[partner[cur]] = 1;
            }
        }
    }
    return masked(mask, nl);
}

PostorderNodes postorder_nodes(const PhyloTree& t, const std::unordered_map<std::string, uint32_t>& obs_index) {
    PostorderNodes p;
    const uint32_t n = t.nparens();
    const uint32_t total = n / 2;
    p.m = total > 0 ? total - 1 : 0;
    std::vector<int32_t> post_of_open(n, -1);
    uint32_t k = 0;
    for (uint32_t i = 0; i < n; ++i)
        if (!t.open[i]) post_of_open[t.partner[i]] = (int32_t)k++;
    p.parent.assign(p.m, (int32_t)p.m);
    p.length.assign(p.m, 0.0);
    p.is_leaf.assign(p.m, 0);
    p.obs_row.assign(p.m, -1);
    p.height.assign(p.m, 0);
    std::vector<int32_t> h(total, 0);
    for (uint32_t i = 0; i < n; ++i) {
        if (t.open[i]) continue;
        const uint32_t o = t.partner[i];
        const uint32_t kk = (uint32_t)post_of_open[o];
        if (kk >= p.m) continue;  // root
        const int32_t po = t.parent_open[o];
        const int32_t pk = po >= 0 ? post_of_open[po] : (int32_t)p.m;
        p.parent[kk] = pk;
        p.length[kk] = t.length[o];
        p.is_leaf[kk] = t.is_leaf(o) ? 1 : 0;
        if (p.is_leaf[kk]) {
            auto it = obs_index.find(t.name[o]);
            if (it != obs_index.end()) p.obs_row[kk] = (int32_t)it->second;
        }
        p.height[kk] = h[kk];
        if (pk >= 0 && (uint32_t)pk < total && h[pk] < h[kk] + 1) h[pk] = h[kk] + 1;
    }
    return p;
}

namespace {
inline uint64_t name_hash(const char* s) {   // FNV-1a
    uint64_t h = 1469598103934665603ull;
    for (; *s; ++s) h = (h ^ (unsigned char)*s) * 1099511628211ull;
    return h ^ (h >> 29);
}

// fn(t, begin, end) over `parts` consecutive ranges of [0, n), range 0 on the calling thread
template <class F>
void split_range(uint64_t n, unsigned parts, F fn) {
    const uint64_t step = (n + parts - 1) / std::max(parts, 1u);
    std::vector<std::thread> pool;
    for (unsigned t = 1; t < parts; ++t) {
        const uint64_t a = std::min(n, step * t), b = std::min(n, a + step);
        pool.emplace_back([=] { fn(t, a, b); });
    }
    fn(0, 0, std::min(n, step));
    for (auto& th : pool) th.join();
}

unsigned threads_for(uint64_t n) {
    return n < (1u << 16) ? 1u : std::max(1u, std::min(8u, std::thread::hardware_concurrency()));
}

// Scratch arrays of the postorder passes, handed from call to call (their pages
// stay mapped: first-touch faults of fresh arrays cost as much as a pass).  The
// list keeps one set per caller that ran at the same time, 9 bytes per tree node
// each, until the process ends.
struct PassScratch {
    std::vector<uint32_t> kc;
    std::vector<int32_t> row;
    std::vector<uint8_t> kept;
};
std::mutex g_scratch_mu;
std::vector<PassScratch> g_scratch;
struct ScratchLease {
    PassScratch s;
    ScratchLease() {
        std::lock_guard<std::mutex> g(g_scratch_mu);
        if (!g_scratch.empty()) { s = std::move(g_scratch.back()); g_scratch.pop_back(); }
    }
    ~ScratchLease() {
        std::lock_guard<std::mutex> g(g_scratch_mu);
        g_scratch.push_back(std::move(s));
    }
};
}  // namespace

NameIndex::NameIndex(const char* const* ids, uint32_t n) : ids_(ids) {
    const unsigned threads = threads_for(n);
    std::vector<uint64_t> hv(n);
    std::vector<uint32_t> count(PARTS, 0);
    split_range(n, threads, [&](unsigned, uint64_t a, uint64_t b) {
        for (uint64_t i = a; i < b; ++i) hv[i] = name_hash(ids[i] ? ids[i] : "");
    });
    for (uint32_t i = 0; i < n; ++i) ++count[hv[i] >> 58];
    uint32_t total = 0;
    for (uint32_t p = 0; p < PARTS; ++p) {
        uint32_t cap = 8;
        while (cap < 2 * (uint64_t)count[p] + 8) cap <<= 1;
        base_[p] = total;
        mask_[p] = cap - 1;
        total += cap;
    }
    slot_.assign(total, Slot{0, -1});
    std::vector<uint32_t> distinct(threads, 0);
    // thread t fills tables t, t + threads, ...: every id of a table in row order
    split_range(threads, threads, [&](unsigned, uint64_t t0, uint64_t t1) {
        for (uint64_t t = t0; t < t1; ++t)
            for (uint32_t i = 0; i < n; ++i) {
                const uint32_t p = (uint32_t)(hv[i] >> 58);
                if (p % threads != t) continue;
                Slot* tab = slot_.data() + base_[p];
                const char* id = ids[i] ? ids[i] : "";
                for (uint32_t h = (uint32_t)hv[i] & mask_[p];; h = (h + 1) & mask_[p]) {
                    Slot& s = tab[h];
                    if (s.row < 0) { s = Slot{hv[i], (int32_t)i}; ++distinct[t]; break; }
                    if (s.hash == hv[i]) {
                        const char* o = ids[s.row] ? ids[s.row] : "";
                        if (std::strcmp(o, id) == 0) { s.row = (int32_t)i; break; }   // the last row wins
                    }
                }
            }
    });
    for (uint32_t d : distinct) distinct_ += d;
}

int32_t NameIndex::find(const char* name) const {
    const char* nm = name ? name : "";
    const uint64_t hv = name_hash(nm);
    const uint32_t p = (uint32_t)(hv >> 58);
    const Slot* tab = slot_.data() + base_[p];
    for (uint32_t h = (uint32_t)hv & mask_[p];; h = (h + 1) & mask_[p]) {
        const Slot& s = tab[h];
        if (s.row < 0) return -1;
        if (s.hash != hv) continue;
        const char* o = ids_[s.row] ? ids_[s.row] : "";
        if (std::strcmp(o, nm) == 0) return s.row;
    }
}

bool synced_postorder(const bool* structure, const double* lengths, const char* const* names, uint32_t n,
                      const NameIndex& rows, PostorderNodes& out, uint32_t& matched) {
    // pass 1, closes in postorder: kept nodes (leaves named by a row, and their
    // ancestors) and each node's number of kept children
    ScratchLease lease;
    lease.s.kc.assign(n, 0);
    lease.s.row.assign(n, -1);
    lease.s.kept.assign(n, 0);
    uint32_t* const kc = lease.s.kc.data();
    int32_t* const row = lease.s.row.data();
    uint8_t* const kept = lease.s.kept.data();
    std::vector<uint32_t> stack;
    stack.reserve(64);
    std::vector<uint8_t> seen;
    matched = 0;
    // the leaves' rows first, on several threads for large trees (a leaf: an open
    // followed by its close, src/tree.cpp:264-266)
    split_range(n, threads_for(n), [&](unsigned, uint64_t a, uint64_t b) {
        for (uint64_t i = a; i < b; ++i)
            if (structure[i] && i + 1 < n && !structure[i + 1]) row[i] = rows.find(names[i]);
    });
    for (uint32_t i = 0; i < n; ++i) {
        if (structure[i]) { stack.push_back(i); continue; }
        if (stack.empty()) return false;
        const uint32_t o = stack.back();
        stack.pop_back();
        bool keep;
        if (i == o + 1) {   // a leaf
            const int32_t r = row[o];
            keep = r >= 0;
            if (keep) {
                if (seen.size() <= (size_t)r) seen.resize((size_t)r + 1, 0);
                if (!seen[r]) { seen[r] = 1; ++matched; }
            }
        } else {
            keep = kc[o] > 0;
        }
        kept[o] = keep;
        if (keep && !stack.empty()) ++kc[stack.back()];
    }
    if (!stack.empty()) return false;
    // pass 2, opens in preorder: drop unkept nodes and collapse the kept internal
    // nodes (not the root) with one kept child, whose lengths accumulate top-down
    // into that child, l_child + (l_parent + ...) (src/tree.cpp:165-191); each
    // kept node is numbered at its close, which is the postorder of the sheared,
    // collapsed tree (as postorder_nodes), and takes the closed nodes pending
    // since its open as its children
    uint32_t total = 0;   // kept, uncollapsed nodes
    for (uint32_t i = 0; i < n; ++i)
        if (structure[i] && kept[i] && (i == 0 || !(kc[i] == 1 && !(i + 1 < n && !structure[i + 1])))) ++total;
    // (a kept node inside a dropped subtree does not exist: kept means a kept leaf below)
    PostorderNodes& p = out;
    p.m = total > 0 ? total - 1 : 0;
    p.parent.assign(p.m, (int32_t)p.m);
    p.length.assign(p.m, 0.0);
    p.is_leaf.assign(p.m, 0);
    p.obs_row.assign(p.m, -1);
    p.height.assign(p.m, 0);
    std::vector<int32_t> h(total + 1, 0);
    struct Frame { bool emit; bool pass; double chain; double len; uint32_t kids0; };
    std::vector<Frame> fs;
    fs.reserve(64);
    std::vector<uint32_t> kids;   // numbers of closed nodes whose parent is still open
    kids.reserve(64);
    uint32_t k = 0;
    for (uint32_t i = 0; i < n; ++i) {
        if (structure[i]) {
            const bool has_chain = !fs.empty() && fs.back().pass;
            const double up = has_chain ? fs.back().chain : 0.0;
            Frame f{false, false, 0.0, 0.0, 0};
            if (kept[i] && (fs.empty() || fs.back().emit || fs.back().pass)) {
                const bool leaf = i + 1 < n && !structure[i + 1];
                const bool collapse = !leaf && i != 0 && kc[i] == 1;
                const double l = has_chain ? lengths[i] + up : lengths[i];
                if (collapse) {
                    f.pass = true;
                    f.chain = l;
                } else {
                    f.emit = true;
                    f.len = l;
                    f.kids0 = (uint32_t)kids.size();
                }
            }
            fs.push_back(f);
            continue;
        }
        const Frame f = fs.back();
        fs.pop_back();
        if (!f.emit) continue;
        const uint32_t kk = k++;
        int32_t hk = 0;
        for (size_t c = f.kids0; c < kids.size(); ++c) {
            const uint32_t ch = kids[c];
            p.parent[ch] = (int32_t)kk;
            hk = std::max(hk, h[ch] + 1);
        }
        kids.resize(f.kids0);
        kids.push_back(kk);
        h[kk] = hk;
        if (kk >= p.m) continue;   // the root
        const bool leaf = structure[i - 1];   // its own open just before (an emitted close has i > 0)
        p.length[kk] = f.len;
        p.is_leaf[kk] = leaf ? 1 : 0;
        if (leaf) p.obs_row[kk] = row[i - 1];
        p.height[kk] = hk;
    }
    return true;
}

}  // namespace ssu
