// Stable LSD radix sort of u32 keys with u32 values (see sort.hip).
#pragma once

#include "common.hpp"

namespace flip2 {

size_t sort_workspace_bytes(uint32_t n, uint32_t beginBit, uint32_t endBit);

// Sorts (keysIn, valsIn) by bits [beginBit, endBit) of the key into
// (keysOut, valsOut).  valsIn == nullptr means values 0..n-1 (the sort
// permutation).  keysAlt/valsAlt are ping-pong buffers (unused for a single
// pass).  Inputs are not modified and must not alias any output buffer.
// passEvents (optional, passes + 1 events): recorded after the histogram and
// offset kernels and after each digit pass, for per-kernel timing.
// itemsPerThread: keys per lane of a digit-pass tile (16, 24 or 32; 0 picks).
void sort_pairs_u32(const uint32_t* keysIn, const uint32_t* valsIn, uint32_t* keysOut, uint32_t* valsOut,
                    uint32_t* keysAlt, uint32_t* valsAlt, uint32_t n, uint32_t beginBit, uint32_t endBit,
                    void* ws, size_t wsBytes, hipStream_t stream, hipEvent_t* passEvents = nullptr,
                    int itemsPerThread = 0);
uint32_t sort_passes(uint32_t beginBit, uint32_t endBit);

// Number of bits needed to represent maxKey (0 -> 0).
inline uint32_t bits_for(uint64_t maxKey) {
    uint32_t b = 0;
    while (b < 32 && (maxKey >> b) != 0) ++b;
    return b;
}

}  // namespace flip2
