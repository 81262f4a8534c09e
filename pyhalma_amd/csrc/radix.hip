// Stable LSD radix sort (8-bit digits).  Per pass: a per-block digit histogram, an
// exclusive scan of each digit's row of the digit-major [digit][block] table (a block per
// digit, which also writes the digit's total), then a stable scatter (each block adds the
// exclusive scan of the 256 digit totals, taken in LDS, to its row offsets) in
// which each block ranks its tile's keys by digit (256 keys per round; ranks among equal
// digits inside a wavefront from eight ballots, wavefronts of a round offset by their
// predecessors' counts through LDS), sorts the tile by digit in LDS and writes each
// digit's run to its place, consecutive threads to consecutive addresses.
#include "radix.h"

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

// Exclusive scan of one digit's row hist[d][0 .. nb) into offs (a block per digit, 1024
// entries per round, four consecutive per thread); the row's total goes to tot[d].
__global__ __launch_bounds__(kThreads) void k_radix_rows(const int32_t* __restrict__ hist,
                                                         int64_t nb, int32_t* __restrict__ offs,
                                                         int32_t* __restrict__ tot) {
    __shared__ int scan_sh[kWaves];
    __shared__ int last;
    const int64_t row = static_cast<int64_t>(blockIdx.x) * nb;
    int run = 0;
    for (int64_t c = 0; c < nb; c += 4 * kThreads) {
        const int64_t j0 = c + 4 * threadIdx.x;
        int v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) v[u] = j0 + u < nb ? hist[row + j0 + u] : 0;
        const int local = (v[0] + v[1]) + (v[2] + v[3]);
        const int excl = block_excl_scan(local, scan_sh);
        int o = run + excl;
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            if (j0 + u < nb) offs[row + j0 + u] = o;
            o += v[u];
        }
        // the round's total: the last thread's exclusive prefix plus its own entries
        if (threadIdx.x == kThreads - 1) last = excl + local;
        __syncthreads();
        run += last;
        __syncthreads();
    }
    if (threadIdx.x == 0) tot[blockIdx.x] = run;
}

// Stable scatter of one tile: every key's rank among the tile's keys of its digit (rounds
// in order, wavefronts in order, lanes by ballot), then the tile sorted by digit in LDS,
// then written out digit run by digit run, consecutive threads to consecutive addresses.
__global__ __launch_bounds__(kThreads) void k_radix_scatter(
    const uint32_t* __restrict__ keys, const int32_t* __restrict__ vals, int64_t n, int shift,
    int64_t nb, const int32_t* __restrict__ offs, const int32_t* __restrict__ tot,
    uint32_t* __restrict__ keys_out, int32_t* __restrict__ vals_out) {
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
    {  // start of the digit's run: the totals of the smaller digits plus this block's row offset
        const int before = block_excl_scan(tot[tid], scan_sh);
        run[tid] = static_cast<int64_t>(before) + offs[static_cast<int64_t>(tid) * nb + blockIdx.x];
    }
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
    c.take<int32_t>(256 * nb);
    c.take<int32_t>(256);
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
    int32_t* offs = c.take<int32_t>(256 * nb);
    int32_t* tot = c.take<int32_t>(256);
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
        k_radix_rows<<<256, kThreads, 0, st>>>(hist, nb, offs, tot);
        HALMA_LAUNCH_CHECK();
        k_radix_scatter<<<static_cast<unsigned>(nb), kThreads, 0, st>>>(kin, vin, n, 8 * p, nb,
                                                                         offs, tot, kout, vout);
        HALMA_LAUNCH_CHECK();
        kin = kout;
        vin = vout;
    }
    return 0;
}

}  // namespace halma
