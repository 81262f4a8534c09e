// Half-mass radius of many groups at once (halo_properties.py:366-392): the distance to
// the centre of the last member, members taken by increasing distance, before the
// running mass first exceeds M / 2 (the smallest distance if the first member already
// exceeds it, the largest if none does).
//
// Distances are FP64, sqrt((dx^2 + dy^2) + dz^2) as the reference writes them.  Members
// are put in (group, distance, position) order by three stable radix passes over 32-bit
// keys (distance low word, distance high word, group: a non-negative double orders like
// its bit pattern), then one 1024-thread block per group runs a block-wide FP64 prefix
// sum of the masses in that order, with a carry between chunks, and takes the first
// position where the running mass passes M / 2.  The sums are deterministic (a fixed
// tree); they may round differently from the reference's sequential loop, which only
// matters for a member whose running mass lies within rounding of M / 2.
#include <math.h>

#include <algorithm>

#include "halma_common.h"
#include "radix.h"

namespace halma {

namespace {

constexpr int kThreads = 256;
constexpr int kBlock = 1024;
constexpr int kRedStride = 16;  // fields per group of the reduction output

inline unsigned blocks_for(int64_t n) { return static_cast<unsigned>((n + kThreads - 1) / kThreads); }

struct Pos {
    const void* x;
    const void* y;
    const void* z;
    int pos_t;
    const void* m;
    int mass_t;
};

__global__ void k_hmr_group(const int64_t* __restrict__ off, int64_t G, int32_t* __restrict__ grp) {
    for (int64_t g = blockIdx.x; g < G; g += gridDim.x)
        for (int64_t k = off[g] + threadIdx.x; k < off[g + 1]; k += blockDim.x)
            grp[k] = static_cast<int32_t>(g);
}

__global__ void k_hmr_dist(int64_t nm, const int32_t* __restrict__ mem,
                           const int32_t* __restrict__ grp, Pos s,
                           const double* __restrict__ red, double* __restrict__ dist,
                           uint32_t* __restrict__ klo, uint32_t* __restrict__ khi) {
    const int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= nm) return;
    const int64_t i = mem[k];
    const double* c = red + static_cast<int64_t>(grp[k]) * kRedStride + 1;
    const double dx = c[0] - ldf(s.x, s.pos_t, i);
    const double dy = c[1] - ldf(s.y, s.pos_t, i);
    const double dz = c[2] - ldf(s.z, s.pos_t, i);
    const double d = sqrt((dx * dx + dy * dy) + dz * dz);
    dist[k] = d;
    const unsigned long long b = static_cast<unsigned long long>(__double_as_longlong(d));
    klo[k] = static_cast<uint32_t>(b & 0xffffffffull);
    khi[k] = static_cast<uint32_t>(b >> 32);
}

__global__ void k_hmr_gather(const uint32_t* __restrict__ src, const int32_t* __restrict__ perm,
                             int64_t n, uint32_t* __restrict__ dst) {
    const int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k < n) dst[k] = src[perm[k]];
}

// inclusive block prefix sum (FP64): wave scans, then the wave totals
__device__ double block_incl_scan(double v, double* sh) {
    const int lane = __lane_id();
    const int w = threadIdx.x / kWave;
    double x = v;
    for (int d = 1; d < kWave; d <<= 1) {
        const double y = __shfl_up(x, d, kWave);
        if (lane >= d) x += y;
    }
    if (lane == kWave - 1) sh[w] = x;
    __syncthreads();
    if (w == 0) {
        const int nw = static_cast<int>(blockDim.x) / kWave;
        double t = lane < nw ? sh[lane] : 0.0;
        for (int d = 1; d < kWave; d <<= 1) {
            const double y = __shfl_up(t, d, kWave);
            if (lane >= d) t += y;
        }
        if (lane < nw) sh[kWave + lane] = t;  // inclusive wave totals
    }
    __syncthreads();
    const double before = w > 0 ? sh[kWave + w - 1] : 0.0;
    const double r = before + x;
    __syncthreads();
    return r;
}

__global__ __launch_bounds__(kBlock) void k_hmr_pick(const int64_t* __restrict__ off, int64_t G,
                                                     const int32_t* __restrict__ order,
                                                     const int32_t* __restrict__ mem, Pos s,
                                                     const double* __restrict__ dist,
                                                     const double* __restrict__ red,
                                                     double* __restrict__ out) {
    __shared__ double sh[2 * kWave];
    __shared__ double run[kBlock];
    __shared__ int first;
    for (int64_t g = blockIdx.x; g < G; g += gridDim.x) {
        const int64_t b = off[g], e = off[g + 1];
        if (e <= b) {
            if (threadIdx.x == 0) out[g] = 0.0;
            continue;  // block-uniform
        }
        const double half = 0.5 * red[g * kRedStride];
        if (threadIdx.x == 0) first = kBlock;  // none yet
        __syncthreads();
        double carry = 0.0;
        int64_t found = -1;
        for (int64_t c0 = b; c0 < e; c0 += kBlock) {
            const int64_t k = c0 + threadIdx.x;
            const double v = k < e ? ldf(s.m, s.mass_t, mem[order[k]]) : 0.0;
            const double incl = carry + block_incl_scan(v, sh);
            run[threadIdx.x] = incl;
            __syncthreads();
            const double prev = threadIdx.x == 0 ? carry : run[threadIdx.x - 1];
            if (k < e && incl > half && !(prev > half)) atomicMin(&first, static_cast<int>(threadIdx.x));
            __syncthreads();
            const int f = first;
            carry = run[kBlock - 1];
            __syncthreads();
            if (f < kBlock) {
                found = c0 + f - b;
                break;
            }
        }
        if (threadIdx.x == 0) {
            int64_t at;
            if (found < 0)
                at = e - 1;          // the running mass never passes M / 2
            else if (found == 0)
                at = b;              // the first member already does: the smallest distance
            else
                at = b + found - 1;  // the last member before the crossing
            out[g] = dist[order[at]];
        }
        __syncthreads();
    }
}

struct HmrWs {
    int32_t* grp;
    double* dist;
    uint32_t* k0;
    uint32_t* k1;
    uint32_t* k2;
    int32_t* p0;
    int32_t* p1;
    void* rws;
    size_t bytes;
};

HmrWs carve_hmr(void* base, int64_t n) {
    Carver c(base);
    const int64_t n1 = n > 0 ? n : 1;
    HmrWs w;
    w.grp = c.take<int32_t>(n1);
    w.dist = c.take<double>(n1);
    w.k0 = c.take<uint32_t>(n1);
    w.k1 = c.take<uint32_t>(n1);
    w.k2 = c.take<uint32_t>(n1);
    w.p0 = c.take<int32_t>(n1);
    w.p1 = c.take<int32_t>(n1);
    w.rws = c.take<char>(radix_workspace_bytes(n1));
    w.bytes = c.used;
    return w;
}

}  // namespace

}  // namespace halma

using namespace halma;

extern "C" {

size_t halma_half_mass_radius_workspace_size(int64_t n_members) {
    if (n_members < 0) return 0;
    return carve_hmr(nullptr, n_members).bytes;
}

int halma_half_mass_radius(int64_t n_groups, int64_t n_members, const int64_t* offsets,
                           const int32_t* members, const void* x, const void* y, const void* z,
                           int pos_dtype, const void* m, int mass_dtype, const double* reduced,
                           double* out, void* ws, size_t ws_bytes, void* stream) {
    if (n_groups < 0 || n_members < 0) return set_error(HALMA_E_ARG, "negative count");
    if (n_groups == 0) return 0;
    if (!offsets || !x || !y || !z || !m || !reduced || !out || (n_members > 0 && !members))
        return set_error(HALMA_E_ARG, "null pointer");
    if ((pos_dtype != HALMA_F32 && pos_dtype != HALMA_F64) ||
        (mass_dtype != HALMA_F32 && mass_dtype != HALMA_F64))
        return set_error(HALMA_E_ARG, "bad dtype");
    if (n_members > INT32_MAX || n_groups > INT32_MAX)
        return set_error(HALMA_E_LIMIT, "at most 2^31-1 members and groups");
    HmrWs w = carve_hmr(ws, n_members);
    if (ws_bytes < w.bytes)
        return set_error(HALMA_E_WORKSPACE, "workspace %zu < required %zu", ws_bytes, w.bytes);
    hipStream_t st = static_cast<hipStream_t>(stream);
    Pos s{x, y, z, pos_dtype, m, mass_dtype};
    const unsigned ggrid = static_cast<unsigned>(std::min<int64_t>(n_groups, 65536));
    if (n_members > 0) {
        k_hmr_group<<<ggrid, kThreads, 0, st>>>(offsets, n_groups, w.grp);
        HALMA_LAUNCH_CHECK();
        k_hmr_dist<<<blocks_for(n_members), kThreads, 0, st>>>(n_members, members, w.grp, s,
                                                               reduced, w.dist, w.k0, w.k1);
        HALMA_LAUNCH_CHECK();
        // (group, distance, position) order: low word, high word, group; stable passes
        int rc = radix_sort_pairs(w.k0, nullptr, w.k2, w.p0, n_members, 32, w.rws, st);
        if (rc) return rc;
        k_hmr_gather<<<blocks_for(n_members), kThreads, 0, st>>>(w.k1, w.p0, n_members, w.k0);
        HALMA_LAUNCH_CHECK();
        rc = radix_sort_pairs(w.k0, w.p0, w.k2, w.p1, n_members, 32, w.rws, st);
        if (rc) return rc;
        k_hmr_gather<<<blocks_for(n_members), kThreads, 0, st>>>(
            reinterpret_cast<const uint32_t*>(w.grp), w.p1, n_members, w.k0);
        HALMA_LAUNCH_CHECK();
        int gbits = 1;
        while (gbits < 32 && (int64_t{1} << gbits) < n_groups) ++gbits;
        rc = radix_sort_pairs(w.k0, w.p1, w.k2, w.p0, n_members, gbits, w.rws, st);
        if (rc) return rc;
    }
    k_hmr_pick<<<ggrid, kBlock, 0, st>>>(offsets, n_groups, w.p0, members, s, w.dist, reduced,
                                         out);
    HALMA_LAUNCH_CHECK();
    return 0;
}

}  // extern "C"
