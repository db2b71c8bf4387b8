// LZ4 block codec from the format specification.
//
// A block is a sequence of (token, literals, match) records.  Token high
// nibble: literal count (15 = more count bytes follow, each 255 continues);
// low nibble: match length - 4 (same extension).  A match is a 2-byte
// little-endian back offset (1..65535).  The final record has literals only.
// End-of-block rules: the last 5 bytes are literals and no match starts in
// the last 12 bytes.
#include "ssu_lz4.hpp"

#include <cstring>
#include <vector>

namespace ssu {

namespace {

constexpr size_t MIN_MATCH = 4;
constexpr size_t LAST_LITERALS = 5;
constexpr size_t MF_LIMIT = 12;
constexpr uint32_t MAX_OFFSET = 65535;
constexpr int HASH_BITS = 16;

inline uint32_t read32(const uint8_t* p) {
    uint32_t v;
    memcpy(&v, p, 4);
    return v;
}

inline uint32_t hash4(uint32_t v) { return (v * 2654435761u) >> (32 - HASH_BITS); }

// writes a length continuation (after the 15 in the token): 255, 255, ..., rest
inline bool put_len(uint8_t*& o, const uint8_t* end, size_t len) {
    while (len >= 255) {
        if (o >= end) return false;
        *o++ = 255;
        len -= 255;
    }
    if (o >= end) return false;
    *o++ = (uint8_t)len;
    return true;
}

inline bool emit(uint8_t*& o, const uint8_t* end, const uint8_t* lit, size_t nlit, size_t offset, size_t mlen) {
    if (o >= end) return false;
    uint8_t* token = o++;
    const size_t lcode = nlit < 15 ? nlit : 15;
    size_t mcode = 0;
    if (mlen) mcode = (mlen - MIN_MATCH) < 15 ? (mlen - MIN_MATCH) : 15;
    *token = (uint8_t)((lcode << 4) | mcode);
    if (nlit >= 15 && !put_len(o, end, nlit - 15)) return false;
    if ((size_t)(end - o) < nlit) return false;
    memcpy(o, lit, nlit);
    o += nlit;
    if (!mlen) return true;
    if (end - o < 2) return false;
    *o++ = (uint8_t)(offset & 0xff);
    *o++ = (uint8_t)(offset >> 8);
    if (mlen - MIN_MATCH >= 15 && !put_len(o, end, mlen - MIN_MATCH - 15)) return false;
    return true;
}

}  // namespace

size_t lz4_compress(const uint8_t* src, size_t n, uint8_t* dst, size_t cap) {
    uint8_t* o = dst;
    const uint8_t* const oend = dst + cap;
    size_t anchor = 0;
    if (n > MF_LIMIT) {
        std::vector<int64_t> table((size_t)1 << HASH_BITS, -1);
        const size_t match_limit = n - LAST_LITERALS;   // matches end before this
        const size_t start_limit = n - MF_LIMIT;        // matches start before this
        size_t ip = 0;
        while (ip < start_limit) {
            const uint32_t seq = read32(src + ip);
            const uint32_t h = hash4(seq);
            const int64_t cand = table[h];
            table[h] = (int64_t)ip;
            if (cand >= 0 && ip - (size_t)cand <= MAX_OFFSET && read32(src + cand) == seq) {
                size_t len = MIN_MATCH;
                while (ip + len < match_limit && src[cand + len] == src[ip + len]) ++len;
                if (!emit(o, oend, src + anchor, ip - anchor, ip - (size_t)cand, len)) return 0;
                // index a position inside the match so nearby repeats are found
                if (len > 2 && ip + len - 2 < start_limit) table[hash4(read32(src + ip + len - 2))] = (int64_t)(ip + len - 2);
                ip += len;
                anchor = ip;
            } else {
                ++ip;
            }
        }
    }
    if (!emit(o, oend, src + anchor, n - anchor, 0, 0)) return 0;
    return (size_t)(o - dst);
}

long lz4_decompress(const uint8_t* src, size_t n, uint8_t* dst, size_t cap) {
    const uint8_t* i = src;
    const uint8_t* const iend = src + n;
    uint8_t* o = dst;
    uint8_t* const oend = dst + cap;
    if (n == 0) return -1;
    for (;;) {
        if (i >= iend) return -1;
        const uint8_t token = *i++;
        size_t nlit = token >> 4;
        if (nlit == 15) {
            uint8_t b;
            do {
                if (i >= iend) return -1;
                b = *i++;
                nlit += b;
            } while (b == 255);
        }
        if ((size_t)(iend - i) < nlit || (size_t)(oend - o) < nlit) return -1;
        memcpy(o, i, nlit);
        i += nlit;
        o += nlit;
        if (i == iend) break;                      // last record: literals only
        if (iend - i < 2) return -1;
        const size_t offset = (size_t)i[0] | ((size_t)i[1] << 8);
        i += 2;
        if (offset == 0 || offset > (size_t)(o - dst)) return -1;
        size_t mlen = token & 15;
        if (mlen == 15) {
            uint8_t b;
            do {
                if (i >= iend) return -1;
                b = *i++;
                mlen += b;
            } while (b == 255);
        }
        mlen += MIN_MATCH;
        if ((size_t)(oend - o) < mlen) return -1;
        const uint8_t* m = o - offset;
        for (size_t k = 0; k < mlen; ++k) o[k] = m[k];   // overlapping copies repeat the pattern
        o += mlen;
    }
    return (long)(o - dst);
}

}  // namespace ssu
