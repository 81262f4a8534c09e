// Exclusive prefix sums: per-block totals, a single-block scan of the totals, then a
// per-block scan that adds its block's offset.  Block = 256 threads x 8 items.
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

template <typename T>
__global__ __launch_bounds__(kThreads) void k_block_sums(const T* __restrict__ in, int64_t n,
                                                         int64_t* __restrict__ sums) {
    __shared__ int64_t sh[kThreads / kWave + 1];
    const int64_t base = static_cast<int64_t>(blockIdx.x) * kTile + threadIdx.x * kItems;
    int64_t v = 0;
#pragma unroll
    for (int k = 0; k < kItems; ++k) v += ld(in, base + k, n);
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

template <typename T>
__global__ __launch_bounds__(kThreads) void k_scan_apply(const T* __restrict__ in, int64_t n,
                                                         const int64_t* __restrict__ sums,
                                                         int64_t* __restrict__ out) {
    __shared__ int64_t sh[kThreads / kWave + 1];
    const int64_t base = static_cast<int64_t>(blockIdx.x) * kTile + threadIdx.x * kItems;
    int64_t vals[kItems];
    int64_t v = 0;
#pragma unroll
    for (int k = 0; k < kItems; ++k) {
        vals[k] = ld(in, base + k, n);
        v += vals[k];
    }
    int64_t tot;
    const int64_t incl = block_incl_scan(v, sh, &tot);
    int64_t run = sums[blockIdx.x] + incl - v;
#pragma unroll
    for (int k = 0; k < kItems; ++k) {
        if (base + k < n) out[base + k] = run;
        run += vals[k];
    }
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
    k_block_sums<T><<<static_cast<unsigned>(nb), kThreads, 0, stream>>>(in, n, sums);
    HALMA_LAUNCH_CHECK();
    k_scan_sums<<<1, kThreads, 0, stream>>>(sums, nb, total);
    HALMA_LAUNCH_CHECK();
    k_scan_apply<T><<<static_cast<unsigned>(nb), kThreads, 0, stream>>>(in, n, sums, out);
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
