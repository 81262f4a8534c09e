// Exclusive prefix sums: per-block totals, then a per-block scan that adds its block's
// offset (the sum of the totals before it, added up by the block itself for up to 4096
// blocks, else taken from a single-block scan of the totals).  Block = 256 threads x 8
// items.
#include "scan.h"

namespace halma {

namespace {
constexpr int kThreads = 256;
constexpr int kItems = 8;
constexpr int kTile = kThreads * kItems;

template <typename T>
__device__ __forceinline__ int64_t ld(const T* in, int64_t i, int64_t n) {
    return i < n ? static_cast<int64_t>(in[i]) : 0;
}

// A thread's kItems consecutive values: 16-byte loads when the array is 16-byte aligned
// (VEC, checked by the host) and the thread's items are all in range, else one by one.
template <bool VEC, typename T>
__device__ __forceinline__ void ld_items(const T* in, int64_t base, int64_t n,
                                         int64_t (&v)[kItems]) {
    if (VEC && base + kItems <= n) {
        constexpr int kPer = 16 / sizeof(T);
        const int4* q = reinterpret_cast<const int4*>(in + base);
#pragma unroll
        for (int c = 0; c < kItems / kPer; ++c) {
            const int4 w = q[c];
            if constexpr (sizeof(T) == 4) {
                v[4 * c + 0] = w.x;
                v[4 * c + 1] = w.y;
                v[4 * c + 2] = w.z;
                v[4 * c + 3] = w.w;
            } else {
                v[2 * c + 0] = static_cast<int64_t>((static_cast<uint64_t>(static_cast<uint32_t>(w.y)) << 32) |
                                                    static_cast<uint32_t>(w.x));
                v[2 * c + 1] = static_cast<int64_t>((static_cast<uint64_t>(static_cast<uint32_t>(w.w)) << 32) |
                                                    static_cast<uint32_t>(w.z));
            }
        }
    } else {
#pragma unroll
        for (int k = 0; k < kItems; ++k) v[k] = ld(in, base + k, n);
    }
}

// The thread's kItems outputs from the running value: 16-byte stores when aligned and in range.
template <bool VEC>
__device__ __forceinline__ void st_items(int64_t* out, int64_t base, int64_t n, int64_t run,
                                         const int64_t (&v)[kItems]) {
    if (VEC && base + kItems <= n) {
        longlong2* q = reinterpret_cast<longlong2*>(out + base);
#pragma unroll
        for (int c = 0; c < kItems / 2; ++c) {
            longlong2 w;
            w.x = run;
            run += v[2 * c];
            w.y = run;
            run += v[2 * c + 1];
            q[c] = w;
        }
    } else {
#pragma unroll
        for (int k = 0; k < kItems; ++k) {
            if (base + k < n) out[base + k] = run;
            run += v[k];
        }
    }
}

// block-wide inclusive scan of one value per thread; returns inclusive, sets *block_total
__device__ int64_t block_incl_scan(int64_t v, int64_t* sh, int64_t* block_total) {
    const int lane = __lane_id();
    const int w = threadIdx.x / kWave;
    int64_t x = v;
    for (int s = 1; s < kWave; s <<= 1) {
        const int64_t y = __shfl_up(x, s, kWave);
        if (lane >= s) x += y;
    }
    if (lane == kWave - 1) sh[w] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
        int64_t acc = 0;
        for (int k = 0; k < kThreads / kWave; ++k) {
            const int64_t t = sh[k];
            sh[k] = acc;
            acc += t;
        }
        sh[kThreads / kWave] = acc;
    }
    __syncthreads();
    x += sh[w];
    *block_total = sh[kThreads / kWave];
    __syncthreads();
    return x;
}

template <bool VEC, typename T>
__global__ __launch_bounds__(kThreads) void k_block_sums(const T* __restrict__ in, int64_t n,
                                                         int64_t* __restrict__ sums) {
    __shared__ int64_t sh[kThreads / kWave + 1];
    const int64_t base = static_cast<int64_t>(blockIdx.x) * kTile + threadIdx.x * kItems;
    int64_t vals[kItems];
    ld_items<VEC>(in, base, n, vals);
    int64_t v = 0;
#pragma unroll
    for (int k = 0; k < kItems; ++k) v += vals[k];
    int64_t tot;
    block_incl_scan(v, sh, &tot);
    if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

// single block: exclusive scan of nb block sums in place, total to *total
__global__ __launch_bounds__(kThreads) void k_scan_sums(int64_t* __restrict__ sums, int64_t nb,
                                                        int64_t* __restrict__ total) {
    __shared__ int64_t sh[kThreads / kWave + 1];
    int64_t carry = 0;
    for (int64_t base = 0; base < nb; base += kThreads) {
        const int64_t i = base + threadIdx.x;
        const int64_t v = i < nb ? sums[i] : 0;
        int64_t tot;
        const int64_t incl = block_incl_scan(v, sh, &tot);
        if (i < nb) sums[i] = carry + incl - v;
        carry += tot;
    }
    if (threadIdx.x == 0 && total) *total = carry;
}

// With FOLD (at most kFoldMax blocks) each block adds up the block sums before its own
// instead of reading their scan, and the last block writes the total: no single-block
// scan between the two passes.
constexpr int64_t kFoldMax = 4096;

template <bool VEC, bool FOLD, typename T>
__global__ __launch_bounds__(kThreads) void k_scan_apply(const T* __restrict__ in, int64_t n,
                                                         const int64_t* __restrict__ sums,
                                                         int64_t* __restrict__ out,
                                                         int64_t* __restrict__ total) {
    __shared__ int64_t sh[kThreads / kWave + 1];
    const int64_t base = static_cast<int64_t>(blockIdx.x) * kTile + threadIdx.x * kItems;
    int64_t vals[kItems];
    ld_items<VEC>(in, base, n, vals);
    int64_t off;
    if (FOLD) {
        int64_t part = 0;
        for (int64_t j0 = threadIdx.x; j0 < blockIdx.x; j0 += 4 * kThreads) {
            int64_t q[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int64_t j = j0 + u * kThreads;
                q[u] = j < blockIdx.x ? sums[j] : 0;
            }
            part += (q[0] + q[1]) + (q[2] + q[3]);
        }
        block_incl_scan(part, sh, &off);
    } else {
        off = sums[blockIdx.x];
    }
    int64_t v = 0;
#pragma unroll
    for (int k = 0; k < kItems; ++k) v += vals[k];
    int64_t tot;
    const int64_t incl = block_incl_scan(v, sh, &tot);
    st_items<VEC>(out, base, n, off + incl - v, vals);
    if (FOLD && total && threadIdx.x == 0 && blockIdx.x == gridDim.x - 1) *total = off + tot;
}

// single block over the whole input, tile by tile with a carry (short inputs: one
// launch instead of three)
constexpr int64_t kSingleTiles = 8;

template <typename T>
__global__ __launch_bounds__(kThreads) void k_scan_single(const T* __restrict__ in, int64_t n,
                                                          int64_t* __restrict__ out,
                                                          int64_t* __restrict__ total) {
    __shared__ int64_t sh[kThreads / kWave + 1];
    int64_t carry = 0;
    for (int64_t t0 = 0; t0 < n; t0 += kTile) {
        const int64_t base = t0 + threadIdx.x * kItems;
        int64_t vals[kItems];
        int64_t v = 0;
#pragma unroll
        for (int k = 0; k < kItems; ++k) {
            vals[k] = ld(in, base + k, n);
            v += vals[k];
        }
        int64_t tot;
        const int64_t incl = block_incl_scan(v, sh, &tot);
        int64_t run = carry + incl - v;
#pragma unroll
        for (int k = 0; k < kItems; ++k) {
            if (base + k < n) out[base + k] = run;
            run += vals[k];
        }
        carry += tot;
    }
    if (threadIdx.x == 0 && total) *total = carry;
}

template <typename T>
int scan_impl(const T* in, int64_t* out, int64_t n, int64_t* total, void* ws,
              hipStream_t stream) {
    const int64_t nb = n > 0 ? (n + kTile - 1) / kTile : 0;
    int64_t* sums = static_cast<int64_t*>(ws);
    if (nb > 0 && nb <= kSingleTiles) {
        k_scan_single<T><<<1, kThreads, 0, stream>>>(in, n, out, total);
        HALMA_LAUNCH_CHECK();
        return 0;
    }
    if (nb == 0) {
        if (total) {
            k_scan_sums<<<1, kThreads, 0, stream>>>(sums, 0, total);
            HALMA_LAUNCH_CHECK();
        }
        return 0;
    }
    if (nb > INT32_MAX) return set_error(HALMA_E_LIMIT, "scan too large");
    const bool vec = reinterpret_cast<uintptr_t>(in) % 16 == 0 &&
                     reinterpret_cast<uintptr_t>(out) % 16 == 0;
    const unsigned g = static_cast<unsigned>(nb);
    if (vec)
        k_block_sums<true, T><<<g, kThreads, 0, stream>>>(in, n, sums);
    else
        k_block_sums<false, T><<<g, kThreads, 0, stream>>>(in, n, sums);
    HALMA_LAUNCH_CHECK();
    if (nb <= kFoldMax) {
        if (vec)
            k_scan_apply<true, true, T><<<g, kThreads, 0, stream>>>(in, n, sums, out, total);
        else
            k_scan_apply<false, true, T><<<g, kThreads, 0, stream>>>(in, n, sums, out, total);
        HALMA_LAUNCH_CHECK();
        return 0;
    }
    k_scan_sums<<<1, kThreads, 0, stream>>>(sums, nb, total);
    HALMA_LAUNCH_CHECK();
    if (vec)
        k_scan_apply<true, false, T><<<g, kThreads, 0, stream>>>(in, n, sums, out, nullptr);
    else
        k_scan_apply<false, false, T><<<g, kThreads, 0, stream>>>(in, n, sums, out, nullptr);
    HALMA_LAUNCH_CHECK();
    return 0;
}

}  // namespace

size_t scan_workspace_bytes(int64_t n) {
    const int64_t nb = n > 0 ? (n + kTile - 1) / kTile : 0;
    return align256(static_cast<size_t>(nb + 1) * sizeof(int64_t));
}

int exclusive_scan_i64(const int64_t* in, int64_t* out, int64_t n, int64_t* total, void* ws,
                       hipStream_t stream) {
    return scan_impl<int64_t>(in, out, n, total, ws, stream);
}

int exclusive_scan_i32_to_i64(const int32_t* in, int64_t* out, int64_t n, int64_t* total,
                              void* ws, hipStream_t stream) {
    return scan_impl<int32_t>(in, out, n, total, ws, stream);
}

}  // namespace halma
