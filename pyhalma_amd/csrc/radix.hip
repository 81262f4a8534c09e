// Stable LSD radix sort (8-bit digits).  Per pass: a per-block digit histogram, one
// exclusive scan over the digit-major [digit][block] table, then a stable scatter in
// which each block walks its tile in order, 256 keys per round; ranks among equal digits
// inside a wavefront come from eight ballots (lanes with the same digit), and the
// wavefronts of a round are offset by their predecessors' counts through LDS.
#include "radix.h"
#include "scan.h"

namespace halma {

namespace {

constexpr int kThreads = 256;
constexpr int kRounds = 8;
constexpr int kTile = kThreads * kRounds;
constexpr int kWaves = kThreads / kWave;

__global__ __launch_bounds__(kThreads) void k_radix_hist(const uint32_t* __restrict__ keys,
                                                         int64_t n, int shift, int64_t nb,
                                                         int32_t* __restrict__ hist) {
    __shared__ int h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    const int64_t base = static_cast<int64_t>(blockIdx.x) * kTile;
    for (int r = 0; r < kRounds; ++r) {
        const int64_t i = base + r * kThreads + threadIdx.x;
        if (i < n) atomicAdd(&h[(keys[i] >> shift) & 0xff], 1);
    }
    __syncthreads();
    hist[static_cast<int64_t>(threadIdx.x) * nb + blockIdx.x] = h[threadIdx.x];
}

__global__ __launch_bounds__(kThreads) void k_radix_scatter(
    const uint32_t* __restrict__ keys, const int32_t* __restrict__ vals, int64_t n, int shift,
    int64_t nb, const int64_t* __restrict__ offs, uint32_t* __restrict__ keys_out,
    int32_t* __restrict__ vals_out) {
    __shared__ int wcnt[kWaves][256];
    __shared__ int64_t run[256];
    const int tid = threadIdx.x;
    const int lane = __lane_id();
    const int w = tid / kWave;
    run[tid] = offs[static_cast<int64_t>(tid) * nb + blockIdx.x];
    const uint64_t lt = lane ? (~0ull >> (64 - lane)) : 0ull;
    const int64_t base = static_cast<int64_t>(blockIdx.x) * kTile;
    for (int r = 0; r < kRounds; ++r) {
#pragma unroll
        for (int k = 0; k < kWaves; ++k) wcnt[k][tid] = 0;
        __syncthreads();
        const int64_t i = base + r * kThreads + tid;
        const bool valid = i < n;
        uint32_t key = 0;
        int32_t val = 0;
        int digit = 0;
        if (valid) {
            key = keys[i];
            val = vals ? vals[i] : static_cast<int32_t>(i);
            digit = (key >> shift) & 0xff;
        }
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const bool bit = (digit >> b) & 1;
            const uint64_t m = __ballot(bit);
            peers &= bit ? m : ~m;
        }
        const int rank = __popcll(peers & lt);
        const bool last = valid && (peers >> lane) == 1ull;
        if (last) wcnt[w][digit] = rank + 1;
        __syncthreads();
        if (valid) {
            int64_t dest = run[digit] + rank;
            for (int k = 0; k < w; ++k) dest += wcnt[k][digit];
            keys_out[dest] = key;
            vals_out[dest] = val;
        }
        __syncthreads();
        int add = 0;
#pragma unroll
        for (int k = 0; k < kWaves; ++k) add += wcnt[k][tid];
        run[tid] += add;
        __syncthreads();
    }
}

}  // namespace

size_t radix_workspace_bytes(int64_t n) {
    const int64_t nb = n > 0 ? (n + kTile - 1) / kTile : 1;
    Carver c(nullptr);
    c.take<int32_t>(256 * nb);
    c.take<int64_t>(256 * nb);
    c.take<char>(scan_workspace_bytes(256 * nb));
    c.take<uint32_t>(n > 0 ? n : 1);
    c.take<int32_t>(n > 0 ? n : 1);
    return c.used;
}

int radix_sort_pairs(const uint32_t* keys_in, const int32_t* vals_in, uint32_t* keys_out,
                     int32_t* vals_out, int64_t n, int bits, void* ws, hipStream_t st) {
    if (n <= 0) return 0;
    if (n > INT32_MAX) return set_error(HALMA_E_LIMIT, "radix sort: too many keys");
    const int64_t nb = (n + kTile - 1) / kTile;
    Carver c(ws);
    int32_t* hist = c.take<int32_t>(256 * nb);
    int64_t* offs = c.take<int64_t>(256 * nb);
    void* sws = c.take<char>(scan_workspace_bytes(256 * nb));
    uint32_t* tk = c.take<uint32_t>(n);
    int32_t* tv = c.take<int32_t>(n);
    const int passes = bits <= 0 ? 1 : (bits + 7) / 8;
    // ping-pong so that the last pass lands in the output arrays
    const uint32_t* kin = keys_in;
    const int32_t* vin = vals_in;
    for (int p = 0; p < passes; ++p) {
        const bool to_out = ((passes - 1 - p) % 2) == 0;
        uint32_t* kout = to_out ? keys_out : tk;
        int32_t* vout = to_out ? vals_out : tv;
        k_radix_hist<<<static_cast<unsigned>(nb), kThreads, 0, st>>>(kin, n, 8 * p, nb, hist);
        HALMA_LAUNCH_CHECK();
        int rc = exclusive_scan_i32_to_i64(hist, offs, 256 * nb, nullptr, sws, st);
        if (rc) return rc;
        k_radix_scatter<<<static_cast<unsigned>(nb), kThreads, 0, st>>>(kin, vin, n, 8 * p, nb,
                                                                         offs, kout, vout);
        HALMA_LAUNCH_CHECK();
        kin = kout;
        vin = vout;
    }
    return 0;
}

}  // namespace halma
