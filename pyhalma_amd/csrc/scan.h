// Device-wide exclusive prefix sums (reduce-then-scan, two or three kernels; deterministic).
#pragma once
#include "halma_common.h"

namespace halma {

// Workspace (bytes) for scanning n elements of 8 bytes.
size_t scan_workspace_bytes(int64_t n);

// out[i] = sum_{k<i} in[i] (int64); *total (device, may be null) = sum of all.
// in and out may alias.
int exclusive_scan_i64(const int64_t* in, int64_t* out, int64_t n, int64_t* total,
                       void* ws, hipStream_t stream);

// Same for int32 values accumulated in int64 (out is int64).
int exclusive_scan_i32_to_i64(const int32_t* in, int64_t* out, int64_t n, int64_t* total,
                              void* ws, hipStream_t stream);

}  // namespace halma
