// Internal interface of the pair-potential kernels (shared by potential.hip and the
// batched unbinding driver).
#pragma once
#include <vector>

#include "halma_common.h"

namespace halma {

// Per-halo descriptor of one batched pair-sum launch.  Sources of halo h are
// src4[src_off, src_off + src_len), tests tst4[tst_off, tst_off + tst_len); chunk c of
// the sources covers [c*chunk, (c+1)*chunk); partial sums of test i for chunk c live at
// part[part_off + c*stride + i].  With a dynamic length array (self-potential of a
// shrinking bound set) src_len = tst_len = dyn_len[h] at run time, never above the
// planned lengths.
struct __attribute__((aligned(8))) HaloDesc {
    int64_t src_off, tst_off, part_off;
    int32_t src_len, tst_len, chunk, stride;
};

struct PotPlan {
    std::vector<HaloDesc> halos;
    std::vector<int4> items;   // {halo, test tile, source chunk, 0}, largest first
    std::vector<int32_t> order;  // the haloes in item order (largest item first)
    int64_t n_items = 0;         // item count (items is filled only when asked for)
    int64_t n_src = 0, n_tst = 0, n_part = 0;
    int tests_per_lane = 2, subtile = 8;
    int grid = 0;              // persistent workgroups (256 threads)
};

// The halo order of a plan: indices sorted by key, largest first, ties by index (a
// stable descending order).  Keys are non-negative and below 2^31 * 2^31.
std::vector<int32_t> order_desc(const std::vector<int64_t>& key);

// Plans a batched launch from HOST offset arrays (H+1 entries each).  The item list is
// built only with with_items (callers that generate items on the device need the count).
// tests_per_lane > 0 overrides the tuned tile height (kernels with a fixed one).
int plan_pot(int64_t n_halo, const int64_t* src_off, const int64_t* tst_off, PotPlan& plan,
             bool with_items = true, int tests_per_lane = 0);

// Live set sizes for a plan launched repeatedly over shrinking sets (all pointers may
// be null for a static launch).  When set, halo h has tst_len[h] tests and src_len[h]
// sources (never above the planned lengths), and its sources are read from src_alt
// instead of the main source array when src_sel[h] != 0 (same offsets).
// When n_items is set, the item list was written on the device and holds *n_items
// entries (never more than the planned count, which bounds it).
struct DynSets {
    const int* tst_len = nullptr;
    const int* src_len = nullptr;
    const int* src_sel = nullptr;
    const float4* src_alt = nullptr;
    const int64_t* n_items = nullptr;
    const int* chunk = nullptr;  // per-halo source chunk of this launch (else the plan's)
};

// Batched FP32 compensated pair sums over H haloes (HOST CSR offsets), FP64 (and/or
// FP32) result per test; scratch from pot_batched_workspace_bytes.  Stream-ordered.
size_t pot_batched_workspace_bytes(int64_t H, const int64_t* src_off, const int64_t* tst_off);
// ev_begin / ev_end (may be null): events recorded on the stream right before and after
// the pair kernel's launch.
int pot_batched(int64_t H, const int64_t* src_off, const float* sm, const float* sx,
                const float* sy, const float* sz, const int64_t* tst_off, const float* tx,
                const float* ty, const float* tz, double* out_f64, float* out_f32, void* ws,
                size_t ws_bytes, hipStream_t stream, hipEvent_t ev_begin = nullptr,
                hipEvent_t ev_end = nullptr);

// Launches the persistent FP32 kernel for a plan whose tables are already on the
// device (items taken round-robin by the waves).
int launch_pot_f32(const PotPlan& plan, const float4* src4, const float4* tst4,
                   const HaloDesc* d_halos, DynSets dyn, const int4* d_items, float2* part,
                   hipStream_t stream);

// Packs SoA FP32 (m, x, y, z) into AoS float4 (x, y, z, m); m may be null (-> 0).
int launch_pack4(const float* m, const float* x, const float* y, const float* z,
                 int64_t n, float4* out, hipStream_t stream);

// Source accessors for the ordered FP64 sum.
struct SrcAoS {
    const float4* __restrict__ p;
    __device__ __forceinline__ float4 operator()(int64_t j) const { return p[j]; }
};
struct SrcSoA {
    const float* __restrict__ m;
    const float* __restrict__ x;
    const float* __restrict__ y;
    const float* __restrict__ z;
    __device__ __forceinline__ float4 operator()(int64_t j) const {
        return make_float4(x[j], y[j], z[j], m[j]);
    }
};

// One FP32 pair term m_j / r added to acc: the three coordinate differences, r^2 by
// FMAs, v_rsq_f32, and the reference's rule (a pair counts only when all three
// differences are non-zero: the smallest |difference| compared with zero).
__device__ __forceinline__ float pair_acc(const float4 s, float xi, float yi, float zi,
                                          float acc) {
    const float dx = s.x - xi;
    const float dy = s.y - yi;
    const float dz = s.z - zi;
    const float d2 = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
    const float mn = fminf(fminf(fabsf(dx), fabsf(dy)), fabsf(dz));
    const float inv = mn != 0.0f ? __builtin_amdgcn_rsqf(d2) : 0.0f;
    return __builtin_fmaf(s.w, inv, acc);
}

// Kahan-compensated addition: c holds the low part lost so far (s + c is the sum; the
// sequence of s is that of the textbook form, whose compensation is -c)
__device__ __forceinline__ void kahan_add(float& s, float& c, float v) {
    const float y = v + c;
    const float t = s + y;
    c = y - (t - s);
    s = t;
}


// FP64 sum in the oracle's fixed order for one test position against ns sources:
// kOrdWidth strided partials L[k] = sum over j = k, k + kOrdWidth, ... ascending (each
// from 0.0), then the pairwise tree L[k] = L[k] + L[k + s] for s = kOrdWidth/2, ..., 1
// (k < s); the result is L[0].  Used by the band re-test and the FP64 checker.  Must be
// called by a whole block of kOrdWidth threads (sh: kOrdWidth doubles of LDS); every
// thread returns the same value.
constexpr int kOrdWidth = 256;

__device__ __forceinline__ double ord_term(double acc, float4 q, float tx, float ty, float tz,
                                           double xd, double yd, double zd) {
    if (q.x != tx && q.y != ty && q.z != tz) {
        const double dx = static_cast<double>(q.x) - xd;
        const double dy = static_cast<double>(q.y) - yd;
        const double dz = static_cast<double>(q.z) - zd;
        const double d2 = (dx * dx + dy * dy) + dz * dz;
        acc = acc + static_cast<double>(q.w) / __builtin_sqrt(d2);
    }
    return acc;
}

template <class Src>
__device__ double block_pot_f64_ordered(Src src, int64_t ns, float tx, float ty, float tz,
                                        double* sh) {
    const int t = threadIdx.x;
    const double xd = tx, yd = ty, zd = tz;
    double acc = 0.0;
    int64_t j = t;
    // four loads in flight per step; the additions stay in ascending j
    for (; j + 3 * kOrdWidth < ns; j += 4 * kOrdWidth) {
        float4 q[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) q[u] = src(j + u * kOrdWidth);
#pragma unroll
        for (int u = 0; u < 4; ++u) acc = ord_term(acc, q[u], tx, ty, tz, xd, yd, zd);
    }
    for (; j < ns; j += kOrdWidth) acc = ord_term(acc, src(j), tx, ty, tz, xd, yd, zd);
    sh[t] = acc;
    __syncthreads();
    for (int s = kOrdWidth / 2; s >= kWave; s >>= 1) {
        if (t < s) sh[t] = sh[t] + sh[t + s];
        __syncthreads();
    }
    if (t < kWave) {
        double r = sh[t];
        for (int s = kWave / 2; s >= 1; s >>= 1) r = r + __shfl_down(r, s, kWave);
        if (t == 0) sh[0] = r;
    }
    __syncthreads();
    const double out = sh[0];
    __syncthreads();
    return out;
}

// The same sum and the same bits from one wavefront: lane l keeps the partials l, l + 64,
// l + 128 and l + 192 (each over its sources in ascending j), folds the tree's first two
// levels in its registers (L[l] + L[l + 128], L[l + 64] + L[l + 192], then their sum, as
// the levels s = 128 and s = 64 add them) and the remaining six by shuffles.  Every lane
// returns the same value.
template <class Src>
__device__ double wave_pot_f64_ordered(Src src, int64_t ns, float tx, float ty, float tz) {
    static_assert(kOrdWidth == 4 * kWave, "four partials per lane");
    const int l = __lane_id();
    const double xd = tx, yd = ty, zd = tz;
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
    int64_t j = l;
    for (; j + 3 * kWave < ns; j += kOrdWidth) {
        const float4 q0 = src(j), q1 = src(j + kWave), q2 = src(j + 2 * kWave),
                     q3 = src(j + 3 * kWave);
        a0 = ord_term(a0, q0, tx, ty, tz, xd, yd, zd);
        a1 = ord_term(a1, q1, tx, ty, tz, xd, yd, zd);
        a2 = ord_term(a2, q2, tx, ty, tz, xd, yd, zd);
        a3 = ord_term(a3, q3, tx, ty, tz, xd, yd, zd);
    }
    if (j < ns) a0 = ord_term(a0, src(j), tx, ty, tz, xd, yd, zd);
    if (j + kWave < ns) a1 = ord_term(a1, src(j + kWave), tx, ty, tz, xd, yd, zd);
    if (j + 2 * kWave < ns) a2 = ord_term(a2, src(j + 2 * kWave), tx, ty, tz, xd, yd, zd);
    a0 = a0 + a2;
    a1 = a1 + a3;
    double r = a0 + a1;
    for (int s = kWave / 2; s >= 1; s >>= 1) r = r + __shfl_down(r, s, kWave);
    return __shfl(r, 0, kWave);
}

}  // namespace halma
