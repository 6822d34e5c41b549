// scan.hpp -- device-wide exclusive prefix sum (uint32).
#pragma once
#include "common.hpp"

namespace ppf {

// Scratch (uint32 words) that exclusive_scan_u32 needs for n items.
size_t scan_scratch_words(size_t n);

// out[i] = sum(in[0..i)), n items, deterministic. `in` and `out` may alias.
hipError_t exclusive_scan_u32(const uint32_t *in, uint32_t *out, size_t n,
                              uint32_t *scratch, hipStream_t s);

} // namespace ppf
