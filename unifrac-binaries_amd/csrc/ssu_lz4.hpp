// LZ4 block format (the public LZ4 block specification): the compression used
// by the partial-result files (src/api.cpp:1972-2055 writes each stripe and
// the sample-id block as one LZ4 block).  This is an independent codec for
// that format; any conforming LZ4 block decoder reads its output and it reads
// any conforming block.
#pragma once

#include <cstddef>
#include <cstdint>

namespace ssu {

// Worst-case compressed size of n input bytes (same bound as LZ4_compressBound).
inline size_t lz4_bound(size_t n) { return n + n / 255 + 16; }

// Compress src[0..n) into dst (capacity cap >= lz4_bound(n)).  Returns the
// compressed size, or 0 if cap is too small.
size_t lz4_compress(const uint8_t* src, size_t n, uint8_t* dst, size_t cap);

// Decompress one block of n bytes into dst (capacity cap).  Returns the
// decompressed size, or -1 when the block is malformed or does not fit.
long lz4_decompress(const uint8_t* src, size_t n, uint8_t* dst, size_t cap);

}  // namespace ssu
