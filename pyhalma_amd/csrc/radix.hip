// Stable LSD radix sort (8-bit digits).  Per pass: a per-block digit histogram, one
// exclusive scan over the digit-major [digit][block] table, then a stable scatter in
// which each block ranks its tile's keys by digit (256 keys per round; ranks among equal
// digits inside a wavefront from eight ballots, wavefronts of a round offset by their
// predecessors' counts through LDS), sorts the tile by digit in LDS and writes each
// digit's run to its place, consecutive threads to consecutive addresses.
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

// exclusive scan of one int per thread over the block (256 threads)
__device__ int block_excl_scan(int v, int* sh /* kWaves */) {
    const int lane = __lane_id(), w = threadIdx.x / kWave;
    int x = v;
    for (int d = 1; d < kWave; d <<= 1) {
        const int y = __shfl_up(x, d, kWave);
        if (lane >= d) x += y;
    }
    if (lane == kWave - 1) sh[w] = x;
    __syncthreads();
    int before = 0;
    for (int k = 0; k < w; ++k) before += sh[k];
    __syncthreads();
    return before + x - v;
}

// Stable scatter of one tile: every key's rank among the tile's keys of its digit (rounds
// in order, wavefronts in order, lanes by ballot), then the tile sorted by digit in LDS,
// then written out digit run by digit run, consecutive threads to consecutive addresses.
__global__ __launch_bounds__(kThreads) void k_radix_scatter(
    const uint32_t* __restrict__ keys, const int32_t* __restrict__ vals, int64_t n, int shift,
    int64_t nb, const int64_t* __restrict__ offs, uint32_t* __restrict__ keys_out,
    int32_t* __restrict__ vals_out) {
    __shared__ int wcnt[kWaves][256];
    __shared__ int tcnt[256];
    __shared__ int loff[256];
    __shared__ int scan_sh[kWaves];
    __shared__ int64_t run[256];
    __shared__ uint32_t sk[kTile];
    __shared__ int32_t sv[kTile];
    const int tid = threadIdx.x;
    const int lane = __lane_id();
    const int w = tid / kWave;
    run[tid] = offs[static_cast<int64_t>(tid) * nb + blockIdx.x];
    tcnt[tid] = 0;
    const uint64_t lt = lane ? (~0ull >> (64 - lane)) : 0ull;
    const int64_t base = static_cast<int64_t>(blockIdx.x) * kTile;
    uint32_t kk[kRounds];
    int32_t vv[kRounds];
    int rk[kRounds];
#pragma unroll
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
        if (valid && (peers >> lane) == 1ull) wcnt[w][digit] = rank + 1;
        __syncthreads();
        int lr = -1;
        if (valid) {
            lr = tcnt[digit] + rank;
            for (int k = 0; k < w; ++k) lr += wcnt[k][digit];
        }
        kk[r] = key;
        vv[r] = val;
        rk[r] = lr;
        __syncthreads();
        int add = 0;
#pragma unroll
        for (int k = 0; k < kWaves; ++k) add += wcnt[k][tid];
        tcnt[tid] += add;
    }
    __syncthreads();
    loff[tid] = block_excl_scan(tcnt[tid], scan_sh);
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kRounds; ++r)
        if (rk[r] >= 0) {
            const int pos = loff[(kk[r] >> shift) & 0xff] + rk[r];
            sk[pos] = kk[r];
            sv[pos] = vv[r];
        }
    __syncthreads();
    const int64_t left = n - base;
    const int tile_n = left < kTile ? static_cast<int>(left) : kTile;
#pragma unroll
    for (int r = 0; r < kRounds; ++r) {
        const int j = r * kThreads + tid;
        if (j < tile_n) {
            const uint32_t key = sk[j];
            const int d = (key >> shift) & 0xff;
            const int64_t dest = run[d] + (j - loff[d]);
            keys_out[dest] = key;
            vals_out[dest] = sv[j];
        }
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
