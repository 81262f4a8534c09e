// Stable LSD radix sort of (uint32 key, int32 value) pairs, 8-bit digits.
#pragma once
#include "halma_common.h"

namespace halma {

size_t radix_workspace_bytes(int64_t n);

// Sorts by the low `bits` bits of keys (keys above are ignored).  vals_in may be null:
// the values are then the input positions 0..n-1.  Results in keys_out / vals_out.
int radix_sort_pairs(const uint32_t* keys_in, const int32_t* vals_in, uint32_t* keys_out,
                     int32_t* vals_out, int64_t n, int bits, void* ws, hipStream_t stream);

}  // namespace halma
