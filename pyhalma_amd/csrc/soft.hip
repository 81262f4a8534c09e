// Softened pair potential and its gradient (the potential-minimum centring and the
// differentiable halo potential energy).
//
//   S_i   = sum_j m_j (r_ij^2 + eps^2)^(-1/2)
//   a_i   = dS_i / dt_i = sum_j m_j (x_j - t_i) (r_ij^2 + eps^2)^(-3/2)
//   w_i   = sum_j m_j (r_ij^2 + eps^2)^(-3/2)
// over the sources j of test i's halo whose three coordinates all differ from the
// test's (the rule of particle_subroutines.f90:499-501, which also drops the self pair).
// Sources are FP32 (x, y, z, m).  Tests are either FP64 points (a centre), entering the
// pair terms as an FP32 (hi, lo) pair, dx = (x_j - hi) - lo, so a centre can move by less
// than an FP32 ulp of its box coordinate, or the halo's own sources (the potential
// energy; lo = 0).
//
// Two work shapes, chosen per halo from its own sizes:
//  * tile haloes (more than kFewTests tests): a wave holds 64 tests, one per lane, and
//    streams a chunk of the sources as wave-uniform scalar loads; FP32 sub-tiles of kSub
//    terms Kahan-added; one float2 {sum, compensation} partial per (test, chunk);
//  * few-test haloes (a centre per halo): a wave takes one test (wave-uniform) and a
//    chunk of kFewChunk sources, the lanes over the sources (coalesced 16-byte loads,
//    eight in flight per lane), each lane's terms in FP32 sub-tiles + Kahan, the lanes
//    folded by a fixed FP64 butterfly into one FP64 partial per (test, chunk).
// Every chunk size depends only on the halo's own source and test counts, and the merge
// adds a test's chunk partials in FP64 in chunk order whatever the launch shape (a
// thread per test, or a wavefront per test for calls with few tests): a test's result
// depends on its halo's inputs only.  Nothing of the N^2 pair terms is stored: the
// gradient kernels recompute every term.
#include <math.h>

#include <algorithm>
#include <vector>

#include "halma_common.h"
#include "staging.h"
#include "tuning.h"

namespace halma {

namespace {

constexpr int kSub = 8;
constexpr int kFewTests = 8;       // haloes with at most this many tests: lanes over sources
constexpr int kFewChunk = 4096;    // sources of a few-test item (64 per lane)
constexpr int kTileItems = 256;    // a tile halo's chunks aim at this many items ...
constexpr int kTileMinChunk = 1024;  // ... with chunks of at least this many sources

// pair quantities: bit 0 the potential S, bit 1 the gradient a and w
constexpr int kModePot = 1;
constexpr int kModeAcc = 2;
constexpr int kModeBoth = 3;
__host__ __device__ constexpr int n_planes(int mode) {
    return ((mode & kModePot) ? 1 : 0) + ((mode & kModeAcc) ? 4 : 0);
}

struct __attribute__((aligned(8))) SoftDesc {
    int64_t src_off, tst_off, part_off;
    int32_t src_len, tst_len, chunk, few;
};

struct TestPt {
    float hx, hy, hz, lx, ly, lz;
};

__device__ __forceinline__ TestPt split_test(double x, double y, double z) {
    TestPt t;
    t.hx = static_cast<float>(x);
    t.hy = static_cast<float>(y);
    t.hz = static_cast<float>(z);
    t.lx = static_cast<float>(x - static_cast<double>(t.hx));
    t.ly = static_cast<float>(y - static_cast<double>(t.hy));
    t.lz = static_cast<float>(z - static_cast<double>(t.hz));
    return t;
}

// Kahan-compensated addition: c holds the low part lost so far (s + c is the sum; the
// sequence of s is that of the textbook form, whose compensation is -c)
__device__ __forceinline__ void kahan(float& s, float& c, float v) {
    const float y = v + c;
    const float t = s + y;
    c = y - (t - s);
    s = t;
}

// One pair term added into the plain sub-tile sums ps (planes as n_planes(MODE): the
// potential first when present, then ax, ay, az, w).  FP64 tests (SELF false) subtract
// their lo part after the hi part.
template <int MODE, bool SELF>
__device__ __forceinline__ void pair_term(float* ps, float4 p, const TestPt& t, float e2) {
    const float dx = SELF ? p.x - t.hx : (p.x - t.hx) - t.lx;
    const float dy = SELF ? p.y - t.hy : (p.y - t.hy) - t.ly;
    const float dz = SELF ? p.z - t.hz : (p.z - t.hz) - t.lz;
    const float mn = fminf(fminf(fabsf(dx), fabsf(dy)), fabsf(dz));
    const float d2 = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, __builtin_fmaf(dx, dx, e2)));
    const float inv = mn != 0.0f ? __builtin_amdgcn_rsqf(d2) : 0.0f;
    constexpr int A = (MODE & kModePot) ? 1 : 0;
    if (MODE & kModePot) ps[0] = __builtin_fmaf(p.w, inv, ps[0]);
    if (MODE & kModeAcc) {
        const float w = p.w * (inv * (inv * inv));
        ps[A + 0] = __builtin_fmaf(w, dx, ps[A + 0]);
        ps[A + 1] = __builtin_fmaf(w, dy, ps[A + 1]);
        ps[A + 2] = __builtin_fmaf(w, dz, ps[A + 2]);
        ps[A + 3] += w;
    }
}

template <bool SELF>
__device__ __forceinline__ TestPt load_test(const float4* src, const double* tx,
                                            const double* ty, const double* tz,
                                            const SoftDesc& d, int i) {
    if (SELF) {
        const float4 q = src[d.src_off + i];
        TestPt t;
        t.hx = q.x;
        t.hy = q.y;
        t.hz = q.z;
        t.lx = t.ly = t.lz = 0.0f;
        return t;
    }
    const int64_t ti = d.tst_off + i;
    return split_test(tx[ti], ty[ti], tz[ti]);
}

// the next item of a persistent wave: its wave number first, then a shared counter
__device__ __forceinline__ int next_item(bool first, int k, int n_waves, int* head) {
    if (first) return k;
    int t = 0;
    if (__lane_id() == 0) t = atomicAdd(head, 1);
    return n_waves + __builtin_amdgcn_readfirstlane(t);
}

// Tile haloes: items {halo, 64-test tile, source chunk}.
template <int MODE, bool SELF>
__global__ __launch_bounds__(256) void k_soft_tile(const float4* __restrict__ src,
                                                   const double* __restrict__ tx,
                                                   const double* __restrict__ ty,
                                                   const double* __restrict__ tz,
                                                   const SoftDesc* __restrict__ halos,
                                                   const int4* __restrict__ items, int n_items,
                                                   float e2, float2* __restrict__ part,
                                                   int64_t n_part, int* __restrict__ head) {
    constexpr int NQ = n_planes(MODE);
    const int lane = __lane_id();
    const int n_waves = static_cast<int>(gridDim.x * (blockDim.x / kWave));
    const int wave = __builtin_amdgcn_readfirstlane(
        static_cast<int>(blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave));
    for (bool first = true;; first = false) {
        const int k = next_item(first, wave, n_waves, head);
        if (k >= n_items) break;
        const int4 it = uniform_load(items + k);
        const SoftDesc d = uniform_load(halos + it.x);
        const int nt = d.tst_len, ns = d.src_len;
        const int i0 = it.y * kWave;
        const int j0 = it.z * d.chunk;
        const int j1 = min(ns, j0 + d.chunk);
        if (i0 >= nt || j0 >= j1) continue;
        const int i = min(i0 + lane, nt - 1);
        const TestPt t = load_test<SELF>(src, tx, ty, tz, d, i);
        float as[NQ], ac[NQ];
#pragma unroll
        for (int q = 0; q < NQ; ++q) as[q] = ac[q] = 0.0f;
        const float4* sp = src + d.src_off;
        int j = j0;
        for (; j + kSub <= j1; j += kSub) {
            float ps[NQ];
#pragma unroll
            for (int q = 0; q < NQ; ++q) ps[q] = 0.0f;
#pragma unroll
            for (int s = 0; s < kSub; ++s) pair_term<MODE, SELF>(ps, uniform_load(sp + j + s), t, e2);
#pragma unroll
            for (int q = 0; q < NQ; ++q) kahan(as[q], ac[q], ps[q]);
        }
        if (j < j1) {
            float ps[NQ];
#pragma unroll
            for (int q = 0; q < NQ; ++q) ps[q] = 0.0f;
            for (; j < j1; ++j) pair_term<MODE, SELF>(ps, uniform_load(sp + j), t, e2);
#pragma unroll
            for (int q = 0; q < NQ; ++q) kahan(as[q], ac[q], ps[q]);
        }
        if (i0 + lane < nt) {
            const int64_t o = d.part_off + static_cast<int64_t>(it.z) * nt + i;
#pragma unroll
            for (int q = 0; q < NQ; ++q) part[q * n_part + o] = make_float2(as[q], ac[q]);
        }
    }
}

// Few-test haloes: items {halo, test, source chunk}; lanes over the chunk's sources.
template <int MODE, bool SELF>
__global__ __launch_bounds__(256) void k_soft_few(const float4* __restrict__ src,
                                                  const double* __restrict__ tx,
                                                  const double* __restrict__ ty,
                                                  const double* __restrict__ tz,
                                                  const SoftDesc* __restrict__ halos,
                                                  const int4* __restrict__ items, int n_items,
                                                  float e2, double* __restrict__ part,
                                                  int64_t n_part, int* __restrict__ head) {
    constexpr int NQ = n_planes(MODE);
    const int lane = __lane_id();
    const int n_waves = static_cast<int>(gridDim.x * (blockDim.x / kWave));
    const int wave = __builtin_amdgcn_readfirstlane(
        static_cast<int>(blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave));
    for (bool first = true;; first = false) {
        const int k = next_item(first, wave, n_waves, head);
        if (k >= n_items) break;
        const int4 it = uniform_load(items + k);
        const SoftDesc d = uniform_load(halos + it.x);
        const int ns = d.src_len;
        const int j0 = it.z * d.chunk;
        const int j1 = min(ns, j0 + d.chunk);
        if (j0 >= j1) continue;
        TestPt t;
        if (SELF) {
            const float4 q = uniform_load(src + d.src_off + it.y);
            t.hx = q.x;
            t.hy = q.y;
            t.hz = q.z;
            t.lx = t.ly = t.lz = 0.0f;
        } else {
            const int64_t ti = d.tst_off + it.y;
            t = split_test(uniform_load(tx + ti), uniform_load(ty + ti), uniform_load(tz + ti));
        }
        float as[NQ], ac[NQ];
#pragma unroll
        for (int q = 0; q < NQ; ++q) as[q] = ac[q] = 0.0f;
        const float4* sp = src + d.src_off;
        const int last = j1 - 1;
        for (int j = j0 + lane; j - lane < j1; j += kSub * kWave) {
            float4 p[kSub];
#pragma unroll
            for (int s = 0; s < kSub; ++s) p[s] = sp[min(j + s * kWave, last)];
            float ps[NQ];
#pragma unroll
            for (int q = 0; q < NQ; ++q) ps[q] = 0.0f;
#pragma unroll
            for (int s = 0; s < kSub; ++s) {
                if (j + s * kWave > last) p[s].w = 0.0f;  // beyond the chunk: no mass
                pair_term<MODE, SELF>(ps, p[s], t, e2);
            }
#pragma unroll
            for (int q = 0; q < NQ; ++q) kahan(as[q], ac[q], ps[q]);
        }
        // lanes a and a ^ m add the same two values: every lane ends with the same total
        double v[NQ];
#pragma unroll
        for (int q = 0; q < NQ; ++q) v[q] = static_cast<double>(as[q]) + static_cast<double>(ac[q]);
#pragma unroll
        for (int m = kWave / 2; m >= 1; m >>= 1)
#pragma unroll
            for (int q = 0; q < NQ; ++q) v[q] = v[q] + __shfl_xor(v[q], m, kWave);
        if (lane == 0) {
            const int64_t o = d.part_off + static_cast<int64_t>(it.z) * d.tst_len + it.y;
#pragma unroll
            for (int q = 0; q < NQ; ++q) part[q * n_part + o] = v[q];
        }
    }
}

// chunk c's partial of plane q for member p of halo d, in FP64
__device__ __forceinline__ double chunk_term(const SoftDesc& d, const void* part, int64_t n_part,
                                             int q, int c, int64_t p) {
    const int64_t o = q * n_part + d.part_off + static_cast<int64_t>(c) * d.tst_len + p;
    if (d.few) return static_cast<const double*>(part)[o];
    const float2 v = static_cast<const float2*>(part)[o];
    return static_cast<double>(v.x) + static_cast<double>(v.y);
}

template <int MODE>
__device__ __forceinline__ void write_out(int64_t g, const double* s, double* pot, double* acc,
                                          double* wsum) {
    constexpr int A = (MODE & kModePot) ? 1 : 0;
    if ((MODE & kModePot) && pot) pot[g] = s[0];
    if (MODE & kModeAcc) {
        if (acc) {
            acc[3 * g + 0] = s[A + 0];
            acc[3 * g + 1] = s[A + 1];
            acc[3 * g + 2] = s[A + 2];
        }
        if (wsum) wsum[g] = s[A + 3];
    }
}

// FP64 merge, a thread per test: the chunk partials added in chunk order, loads eight at
// a time.
template <int MODE>
__global__ __launch_bounds__(256) void k_soft_merge(const SoftDesc* __restrict__ halos,
                                                    const int* __restrict__ tst_halo,
                                                    int64_t n_tst, const void* __restrict__ part,
                                                    int64_t n_part, double* __restrict__ pot,
                                                    double* __restrict__ acc,
                                                    double* __restrict__ wsum) {
    constexpr int NQ = n_planes(MODE);
    const int64_t g = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (g >= n_tst) return;
    const SoftDesc d = halos[tst_halo[g]];
    const int64_t p = g - d.tst_off;
    const int nch = d.src_len > 0 ? (d.src_len + d.chunk - 1) / d.chunk : 0;
    double s[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) s[q] = 0.0;
    for (int c = 0; c < nch; c += 8) {
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            double v[8];
#pragma unroll
            for (int u = 0; u < 8; ++u)
                v[u] = c + u < nch ? chunk_term(d, part, n_part, q, c + u, p) : 0.0;
#pragma unroll
            for (int u = 0; u < 8; ++u)
                if (c + u < nch) s[q] = s[q] + v[u];
        }
    }
    write_out<MODE>(g, s, pot, acc, wsum);
}

// The same merge with a wavefront per test, for calls with few tests (a centre per halo
// against a hundred chunks): the lanes load 64 chunk partials at once into LDS, then
// every lane adds them in chunk order -- the same additions, in the same order, as
// k_soft_merge.
template <int MODE>
__global__ __launch_bounds__(256) void k_soft_merge_wave(
    const SoftDesc* __restrict__ halos, const int* __restrict__ tst_halo, int64_t n_tst,
    const void* __restrict__ part, int64_t n_part, double* __restrict__ pot,
    double* __restrict__ acc, double* __restrict__ wsum) {
    constexpr int NQ = n_planes(MODE);
    __shared__ double sh[256 / kWave][NQ][kWave];
    const int64_t g = (static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x) / kWave;
    const int lane = __lane_id();
    const int wv = threadIdx.x / kWave;
    if (g >= n_tst) return;  // wave-uniform
    const SoftDesc d = halos[tst_halo[g]];
    const int64_t p = g - d.tst_off;
    const int nch = d.src_len > 0 ? (d.src_len + d.chunk - 1) / d.chunk : 0;
    double s[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) s[q] = 0.0;
    for (int c0 = 0; c0 < nch; c0 += kWave) {
        const int c = c0 + lane;
#pragma unroll
        for (int q = 0; q < NQ; ++q) sh[wv][q][lane] = c < nch ? chunk_term(d, part, n_part, q, c, p) : 0.0;
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
        const int n = min(kWave, nch - c0);
        for (int u = 0; u < n; ++u)
#pragma unroll
            for (int q = 0; q < NQ; ++q) s[q] = s[q] + sh[wv][q][u];
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    }
    if (lane == 0) write_out<MODE>(g, s, pot, acc, wsum);
}

// tests per call up to which the merge takes a wavefront per test
constexpr int64_t kWaveMergeTests = 16384;

__global__ void k_soft_tst_halo(const SoftDesc* __restrict__ halos, int n_halo,
                                int* __restrict__ tst_halo) {
    const int h = blockIdx.x;
    if (h >= n_halo) return;
    const SoftDesc d = halos[h];
    for (int i = threadIdx.x; i < d.tst_len; i += blockDim.x) tst_halo[d.tst_off + i] = h;
}

__global__ void k_soft_clear(int* __restrict__ p) {
    if (threadIdx.x < 2) p[threadIdx.x * 16] = 0;
}

__global__ void k_soft_pack(const float* __restrict__ m, const float* __restrict__ x,
                            const float* __restrict__ y, const float* __restrict__ z, int64_t n,
                            float4* __restrict__ out) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) out[i] = make_float4(x[i], y[i], z[i], m[i]);
}

struct SoftPlan {
    std::vector<SoftDesc> halos;
    std::vector<int4> tile_items, few_items;  // largest items first
    int64_t n_src = 0, n_tst = 0, n_part = 0;
};

int plan_soft(int64_t H, const int64_t* src_off, const int64_t* tst_off, SoftPlan& plan) {
    plan = SoftPlan();
    plan.halos.resize(H);
    int64_t part = 0;
    std::vector<int64_t> tkey, fkey;
    for (int64_t h = 0; h < H; ++h) {
        const int64_t ns = src_off[h + 1] - src_off[h];
        const int64_t nt = tst_off[h + 1] - tst_off[h];
        if (ns < 0 || nt < 0 || ns > INT32_MAX || nt > INT32_MAX)
            return set_error(HALMA_E_ARG, "halo %lld: bad offsets", (long long)h);
        SoftDesc& d = plan.halos[h];
        d.src_off = src_off[h] - src_off[0];
        d.tst_off = tst_off[h] - tst_off[0];
        d.src_len = static_cast<int32_t>(ns);
        d.tst_len = static_cast<int32_t>(nt);
        d.few = nt <= kFewTests ? 1 : 0;
        int64_t chunk;
        if (d.few) {
            chunk = kFewChunk;
        } else {
            const int64_t tiles = (nt + kWave - 1) / kWave;
            int64_t nch = (kTileItems + tiles - 1) / tiles;
            nch = std::max<int64_t>(1, std::min<int64_t>(nch, ns / kTileMinChunk));
            chunk = ns > 0 ? (ns + nch - 1) / nch : 1;
            chunk = (chunk + kSub - 1) / kSub * kSub;
        }
        d.chunk = static_cast<int32_t>(std::max<int64_t>(chunk, 1));
        const int64_t nch = ns > 0 ? (ns + d.chunk - 1) / d.chunk : 0;
        d.part_off = part;
        part += nch * nt;
    }
    // items, haloes by descending chunk (the cost of one item), ties by index
    std::vector<int32_t> order(H);
    for (int64_t h = 0; h < H; ++h) order[h] = static_cast<int32_t>(h);
    std::stable_sort(order.begin(), order.end(), [&](int32_t a, int32_t b) {
        const int64_t ka = std::min(plan.halos[a].chunk, plan.halos[a].src_len);
        const int64_t kb = std::min(plan.halos[b].chunk, plan.halos[b].src_len);
        return ka > kb;
    });
    for (int32_t h : order) {
        const SoftDesc& d = plan.halos[h];
        if (d.tst_len == 0 || d.src_len == 0) continue;
        const int nch = (d.src_len + d.chunk - 1) / d.chunk;
        if (d.few) {
            for (int c = 0; c < nch; ++c)
                for (int i = 0; i < d.tst_len; ++i) plan.few_items.push_back(make_int4(h, i, c, 0));
        } else {
            const int tiles = (d.tst_len + kWave - 1) / kWave;
            for (int c = 0; c < nch; ++c)
                for (int ti = 0; ti < tiles; ++ti) plan.tile_items.push_back(make_int4(h, ti, c, 0));
        }
    }
    if (plan.tile_items.size() > static_cast<size_t>(INT32_MAX) ||
        plan.few_items.size() > static_cast<size_t>(INT32_MAX))
        return set_error(HALMA_E_LIMIT, "too many work items");
    plan.n_src = src_off[H] - src_off[0];
    plan.n_tst = tst_off[H] - tst_off[0];
    plan.n_part = part;
    return 0;
}

struct SoftWs {
    float4* src4;
    SoftDesc* halos;
    int4* tile_items;
    int4* few_items;
    char* part;  // n_planes planes of n_part 8-byte slots
    int* tst_halo;
    int* heads;  // [0] tile queue, [16] few-test queue
    size_t bytes;
};

SoftWs carve_soft(void* base, const SoftPlan& p, bool pack) {
    Carver c(base);
    SoftWs w;
    w.src4 = c.take<float4>(pack ? std::max<int64_t>(p.n_src, 1) : 1);
    w.halos = c.take<SoftDesc>(std::max<size_t>(p.halos.size(), 1));
    w.tile_items = c.take<int4>(std::max<size_t>(p.tile_items.size(), 1));
    w.few_items = c.take<int4>(std::max<size_t>(p.few_items.size(), 1));
    w.part = c.take<char>(8 * static_cast<size_t>(n_planes(kModeBoth)) *
                          static_cast<size_t>(std::max<int64_t>(p.n_part, 1)));
    w.tst_halo = c.take<int>(std::max<int64_t>(p.n_tst, 1));
    w.heads = c.take<int>(64);
    w.bytes = c.used;
    return w;
}

int check_offsets(int64_t H, const int64_t* off) {
    for (int64_t h = 0; h < H; ++h)
        if (off[h + 1] < off[h]) return set_error(HALMA_E_ARG, "offsets must not decrease");
    return 0;
}

template <int MODE, bool SELF>
int launch_mode(const SoftPlan& plan, const SoftWs& w, const float4* src4, const double* tx,
                const double* ty, const double* tz, float e2, int grid, double* pot,
                double* acc, double* wsum, hipStream_t st) {
    const int nt = static_cast<int>(plan.tile_items.size());
    const int nf = static_cast<int>(plan.few_items.size());
    k_soft_clear<<<1, 64, 0, st>>>(w.heads);
    HALMA_LAUNCH_CHECK();
    if (nt > 0) {
        const int g = std::max(1, std::min(grid, (nt + 3) / 4));
        k_soft_tile<MODE, SELF><<<g, 256, 0, st>>>(src4, tx, ty, tz, w.halos, w.tile_items, nt,
                                                   e2, reinterpret_cast<float2*>(w.part),
                                                   plan.n_part, w.heads);
        HALMA_LAUNCH_CHECK();
    }
    if (nf > 0) {
        const int g = std::max(1, std::min(grid, (nf + 3) / 4));
        k_soft_few<MODE, SELF><<<g, 256, 0, st>>>(src4, tx, ty, tz, w.halos, w.few_items, nf, e2,
                                                  reinterpret_cast<double*>(w.part), plan.n_part,
                                                  w.heads + 16);
        HALMA_LAUNCH_CHECK();
    }
    const int64_t n = plan.n_tst;
    if (n <= kWaveMergeTests) {
        k_soft_merge_wave<MODE><<<static_cast<unsigned>((n * kWave + 255) / 256), 256, 0, st>>>(
            w.halos, w.tst_halo, n, w.part, plan.n_part, pot, acc, wsum);
    } else {
        k_soft_merge<MODE><<<static_cast<unsigned>((n + 255) / 256), 256, 0, st>>>(
            w.halos, w.tst_halo, n, w.part, plan.n_part, pot, acc, wsum);
    }
    HALMA_LAUNCH_CHECK();
    return 0;
}

// The body of every entry: sources as separate arrays (packed here into the workspace)
// or already packed by the caller as (x, y, z, m) float4; tests FP64 points, or the
// sources themselves (self).
int soft_run(int64_t n_halo, const int64_t* src_off, const float* sm, const float* sx,
             const float* sy, const float* sz, const float* src_xyzm, const int64_t* tst_off,
             const double* tx, const double* ty, const double* tz, bool self, double eps,
             double* pot, double* acc, double* wsum, void* ws, size_t ws_bytes, void* stream) {
    if (n_halo < 0 || !src_off || !tst_off)
        return set_error(HALMA_E_ARG, "halma_pot_soft: bad halo count or offsets");
    if (!(eps >= 0.0) || !isfinite(eps)) return set_error(HALMA_E_ARG, "eps must be >= 0");
    if (n_halo == 0) return 0;
    int rc = check_offsets(n_halo, src_off);
    if (!rc) rc = check_offsets(n_halo, tst_off);
    if (rc) return rc;
    const int64_t n_src = src_off[n_halo] - src_off[0];
    const int64_t n_tst = tst_off[n_halo] - tst_off[0];
    if (n_tst == 0) return 0;
    if ((!self && (!tx || !ty || !tz)) ||
        (n_src > 0 && !src_xyzm && (!sm || !sx || !sy || !sz)))
        return set_error(HALMA_E_ARG, "halma_pot_soft: null pointer");
    if (src_xyzm && reinterpret_cast<uintptr_t>(src_xyzm) % alignof(float4) != 0)
        return set_error(HALMA_E_ARG, "packed sources must be 16-byte aligned");
    if (!pot && !acc && !wsum) return 0;
    hipStream_t st = static_cast<hipStream_t>(stream);
    SoftPlan plan;
    rc = plan_soft(n_halo, src_off, tst_off, plan);
    if (rc) return rc;
    SoftWs w = carve_soft(ws, plan, !src_xyzm);
    if (ws_bytes < w.bytes)
        return set_error(HALMA_E_WORKSPACE, "workspace %zu < required %zu", ws_bytes, w.bytes);
    const int64_t s0 = src_off[0], t0 = tst_off[0];
    const float4* src4 = w.src4;
    if (src_xyzm) {
        src4 = reinterpret_cast<const float4*>(src_xyzm) + s0;
    } else if (n_src > 0) {
        k_soft_pack<<<static_cast<unsigned>((n_src + 255) / 256), 256, 0, st>>>(
            sm + s0, sx + s0, sy + s0, sz + s0, n_src, w.src4);
        HALMA_LAUNCH_CHECK();
    }
    rc = upload_async(w.halos, plan.halos.data(), plan.halos.size() * sizeof(SoftDesc), st);
    if (!rc && !plan.tile_items.empty())
        rc = upload_async(w.tile_items, plan.tile_items.data(),
                          plan.tile_items.size() * sizeof(int4), st);
    if (!rc && !plan.few_items.empty())
        rc = upload_async(w.few_items, plan.few_items.data(),
                          plan.few_items.size() * sizeof(int4), st);
    if (rc) return rc;
    k_soft_tst_halo<<<static_cast<unsigned>(n_halo), 256, 0, st>>>(
        w.halos, static_cast<int>(n_halo), w.tst_halo);
    HALMA_LAUNCH_CHECK();
    const float e2 = static_cast<float>(eps * eps);
    const int cus = std::max(device_cu_count(), 1);
    const int grid = std::max(1, cus * g_tuning.waves_per_cu / 4);
    const double* txo = self ? nullptr : tx + t0;
    const double* tyo = self ? nullptr : ty + t0;
    const double* tzo = self ? nullptr : tz + t0;
    double* pot_o = pot ? pot + t0 : nullptr;
    double* acc_o = acc ? acc + 3 * t0 : nullptr;
    double* wsum_o = wsum ? wsum + t0 : nullptr;
    const bool want_pot = pot != nullptr, want_acc = acc || wsum;
    if (want_pot && want_acc)
        return self ? launch_mode<kModeBoth, true>(plan, w, src4, txo, tyo, tzo, e2, grid, pot_o,
                                                   acc_o, wsum_o, st)
                    : launch_mode<kModeBoth, false>(plan, w, src4, txo, tyo, tzo, e2, grid,
                                                    pot_o, acc_o, wsum_o, st);
    if (want_pot)
        return self ? launch_mode<kModePot, true>(plan, w, src4, txo, tyo, tzo, e2, grid, pot_o,
                                                  nullptr, nullptr, st)
                    : launch_mode<kModePot, false>(plan, w, src4, txo, tyo, tzo, e2, grid, pot_o,
                                                   nullptr, nullptr, st);
    return self ? launch_mode<kModeAcc, true>(plan, w, src4, txo, tyo, tzo, e2, grid, nullptr,
                                              acc_o, wsum_o, st)
                : launch_mode<kModeAcc, false>(plan, w, src4, txo, tyo, tzo, e2, grid, nullptr,
                                               acc_o, wsum_o, st);
}

size_t soft_ws_bytes(int64_t n_halo, const int64_t* src_off, const int64_t* tst_off, bool pack) {
    if (n_halo < 0 || !src_off || !tst_off) return 0;
    if (check_offsets(n_halo, src_off) || check_offsets(n_halo, tst_off)) return 0;
    SoftPlan plan;
    if (plan_soft(n_halo, src_off, tst_off, plan)) return 0;
    return carve_soft(nullptr, plan, pack).bytes;
}

}  // namespace

}  // namespace halma

using namespace halma;

extern "C" {

size_t halma_soft_workspace_size(int64_t n_halo, const int64_t* src_off,
                                 const int64_t* tst_off) {
    return soft_ws_bytes(n_halo, src_off, tst_off, true);
}

int halma_pot_soft_batched(int64_t n_halo, const int64_t* src_off, const float* sm,
                           const float* sx, const float* sy, const float* sz,
                           const int64_t* tst_off, const double* tx, const double* ty,
                           const double* tz, double eps, double* pot, double* acc,
                           double* wsum, void* ws, size_t ws_bytes, void* stream) {
    return soft_run(n_halo, src_off, sm, sx, sy, sz, nullptr, tst_off, tx, ty, tz, false, eps,
                    pot, acc, wsum, ws, ws_bytes, stream);
}

int halma_pot_soft_batched_packed(int64_t n_halo, const int64_t* src_off,
                                  const float* src_xyzm, const int64_t* tst_off,
                                  const double* tx, const double* ty, const double* tz,
                                  double eps, double* pot, double* acc, double* wsum, void* ws,
                                  size_t ws_bytes, void* stream) {
    if (!src_xyzm && n_halo > 0 && src_off && src_off[n_halo] > src_off[0])
        return set_error(HALMA_E_ARG, "halma_pot_soft_batched_packed: null sources");
    return soft_run(n_halo, src_off, nullptr, nullptr, nullptr, nullptr, src_xyzm, tst_off, tx,
                    ty, tz, false, eps, pot, acc, wsum, ws, ws_bytes, stream);
}

size_t halma_soft_self_workspace_size(int64_t n_halo, const int64_t* offsets) {
    return soft_ws_bytes(n_halo, offsets, offsets, false);
}

int halma_pot_soft_self_batched(int64_t n_halo, const int64_t* offsets, const float* src_xyzm,
                                double eps, double* pot, double* acc, double* wsum, void* ws,
                                size_t ws_bytes, void* stream) {
    if (!src_xyzm && n_halo > 0 && offsets && offsets[n_halo] > offsets[0])
        return set_error(HALMA_E_ARG, "halma_pot_soft_self_batched: null sources");
    return soft_run(n_halo, offsets, nullptr, nullptr, nullptr, nullptr, src_xyzm, offsets,
                    nullptr, nullptr, nullptr, true, eps, pot, acc, wsum, ws, ws_bytes, stream);
}

}  // extern "C"
