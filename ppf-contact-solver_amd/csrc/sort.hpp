// sort.hpp -- stable device radix sort of (uint32 key, uint32 value) pairs.
#pragma once
#include "common.hpp"
#include <utility>

namespace ppf {

// Scratch words radix_sort_pairs needs for n pairs.
size_t radix_sort_scratch_words(size_t n);

// Stable ascending sort by the low `key_bits` bits of the keys, in place.
hipError_t radix_sort_pairs(uint32_t *keys, uint32_t *vals, size_t n,
                            uint32_t key_bits, uint32_t *scratch, hipStream_t s);

} // namespace ppf
