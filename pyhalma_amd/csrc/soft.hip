// Softened pair potential and its gradient (the potential-minimum centring).
//
//   S_i   = sum_j m_j (r_ij^2 + eps^2)^(-1/2)
//   a_i   = dS_i / dt_i = sum_j m_j (x_j - t_i) (r_ij^2 + eps^2)^(-3/2)
//   w_i   = sum_j m_j (r_ij^2 + eps^2)^(-3/2)
// over the sources j of test i's halo whose three coordinates all differ from the
// test's (the rule of particle_subroutines.f90:499-501, which also drops the self pair).
// Sources are FP32 (x, y, z, m); tests are FP64 and enter the pair terms as an FP32
// (hi, lo) pair, dx = (x_j - hi) - lo, so a centre can move by less than an FP32 ulp
// of its box coordinate.  The same persistent work queue as the unsoftened kernel
// (halo, 64-test tile, source chunk), FP32 sub-tiles of 8 terms Kahan-added, chunk
// partials merged in FP64 (in chunk order by a thread per test; for calls with few tests
// by a wavefront per test, lanes over strided chunks then a fixed butterfly: the same
// result on every run).  Nothing of the N^2 pair terms is stored: the gradient kernel
// recomputes every term.
#include <math.h>

#include <algorithm>
#include <vector>

#include "halma_common.h"
#include "potential.h"
#include "staging.h"

namespace halma {

namespace {

constexpr int kSub = 8;

struct SoftTest {
    float hx, hy, hz, lx, ly, lz;
};

__device__ __forceinline__ SoftTest split_test(double x, double y, double z) {
    SoftTest t;
    t.hx = static_cast<float>(x);
    t.hy = static_cast<float>(y);
    t.hz = static_cast<float>(z);
    t.lx = static_cast<float>(x - static_cast<double>(t.hx));
    t.ly = static_cast<float>(y - static_cast<double>(t.hy));
    t.lz = static_cast<float>(z - static_cast<double>(t.hz));
    return t;
}

__device__ __forceinline__ void kahan(float& s, float& c, float v) {
    const float y = v - c;
    const float t = s + y;
    c = (t - s) - y;
    s = t;
}

// ACC = false: S; ACC = true: a (3 components) and w.  part holds 1 (S) or 4 (ax, ay,
// az, w) planes of n_part float2 {sum, compensation}.
template <bool ACC>
__global__ __launch_bounds__(256) void k_soft(const float4* __restrict__ src,
                                              const double* __restrict__ tx,
                                              const double* __restrict__ ty,
                                              const double* __restrict__ tz,
                                              const HaloDesc* __restrict__ halos,
                                              const int4* __restrict__ items, int n_items,
                                              float e2, float2* __restrict__ part,
                                              int64_t n_part, int* __restrict__ head) {
    const int lane = __lane_id();
    const int n_waves = static_cast<int>(gridDim.x * (blockDim.x / kWave));
    int k = __builtin_amdgcn_readfirstlane(
        static_cast<int>(blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave));
    constexpr int NQ = ACC ? 4 : 1;
    for (bool first = true;; first = false) {
        if (!first) {
            int t = 0;
            if (lane == 0) t = atomicAdd(head, 1);
            k = n_waves + __builtin_amdgcn_readfirstlane(t);
        }
        if (k >= n_items) break;
        const int4 it = uniform_load(items + k);
        const HaloDesc d = uniform_load(halos + it.x);
        const int nt = d.tst_len, ns = d.src_len;
        const int i0 = it.y * kWave;
        const int j0 = it.z * d.chunk;
        const int j1 = min(ns, j0 + d.chunk);
        if (i0 >= nt || j0 >= j1) continue;
        const int i = min(i0 + lane, nt - 1);
        const int64_t ti = d.tst_off + i;
        const SoftTest t = split_test(tx[ti], ty[ti], tz[ti]);
        float as[NQ], ac[NQ];
#pragma unroll
        for (int q = 0; q < NQ; ++q) as[q] = ac[q] = 0.0f;
        const float4* sp = src + d.src_off;
        for (int j = j0; j < j1; j += kSub) {
            float ps[NQ];
#pragma unroll
            for (int q = 0; q < NQ; ++q) ps[q] = 0.0f;
            const int jn = min(kSub, j1 - j);
            for (int s = 0; s < jn; ++s) {
                const float4 p = uniform_load(sp + j + s);
                const float dx = (p.x - t.hx) - t.lx;
                const float dy = (p.y - t.hy) - t.ly;
                const float dz = (p.z - t.hz) - t.lz;
                const float mn = fminf(fminf(fabsf(dx), fabsf(dy)), fabsf(dz));
                const float d2 = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx)) + e2;
                const float inv = mn != 0.0f ? __builtin_amdgcn_rsqf(d2) : 0.0f;
                if (ACC) {
                    const float w = p.w * (inv * inv * inv);
                    ps[0] = __builtin_fmaf(w, dx, ps[0]);
                    ps[1] = __builtin_fmaf(w, dy, ps[1]);
                    ps[2] = __builtin_fmaf(w, dz, ps[2]);
                    ps[NQ - 1] += w;
                } else {
                    ps[0] = __builtin_fmaf(p.w, inv, ps[0]);
                }
            }
#pragma unroll
            for (int q = 0; q < NQ; ++q) kahan(as[q], ac[q], ps[q]);
        }
        if (i0 + lane < nt) {
            const int64_t o = d.part_off + static_cast<int64_t>(it.z) * d.stride + i;
#pragma unroll
            for (int q = 0; q < NQ; ++q) part[q * n_part + o] = make_float2(as[q], ac[q]);
        }
    }
}

__global__ void k_soft_tst_halo(const HaloDesc* __restrict__ halos, int n_halo,
                                int* __restrict__ tst_halo) {
    const int h = blockIdx.x;
    if (h >= n_halo) return;
    const HaloDesc d = halos[h];
    for (int i = threadIdx.x; i < d.tst_len; i += blockDim.x) tst_halo[d.tst_off + i] = h;
}

__global__ void k_soft_clear(int* __restrict__ p) {
    if (threadIdx.x == 0) *p = 0;
}

// The same merge with one wavefront per test, for calls with few tests (a centre per
// halo against hundreds of source chunks): lane l adds chunks l, l + 64, ... in order,
// then a butterfly over the lanes (lanes a and a ^ m add the same two values, so every
// lane ends with the same total, run after run).
template <bool ACC>
__global__ __launch_bounds__(256) void k_soft_merge_wave(
    const HaloDesc* __restrict__ halos, const int* __restrict__ tst_halo, int64_t n_tst,
    const float2* __restrict__ part, int64_t n_part, double* __restrict__ pot,
    double* __restrict__ acc, double* __restrict__ wsum) {
    const int64_t g = (static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x) / kWave;
    const int lane = __lane_id();
    if (g >= n_tst) return;  // wave-uniform
    const HaloDesc d = halos[tst_halo[g]];
    const int64_t p = g - d.tst_off;
    const int nch = d.src_len > 0 ? (d.src_len + d.chunk - 1) / d.chunk : 0;
    constexpr int NQ = ACC ? 4 : 1;
    double s[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) s[q] = 0.0;
    for (int c = lane; c < nch; c += kWave) {
        const int64_t o = d.part_off + static_cast<int64_t>(c) * d.stride + p;
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            const float2 v = part[q * n_part + o];
            s[q] += static_cast<double>(v.x) + static_cast<double>(v.y);
        }
    }
#pragma unroll
    for (int m = kWave / 2; m >= 1; m >>= 1)
#pragma unroll
        for (int q = 0; q < NQ; ++q) s[q] += __shfl_xor(s[q], m, kWave);
    if (lane != 0) return;
    if (ACC) {
        if (acc) {
            acc[3 * g + 0] = s[0];
            acc[3 * g + 1] = s[1];
            acc[3 * g + 2] = s[2];
        }
        if (wsum) wsum[g] = s[NQ - 1];
    } else if (pot) {
        pot[g] = s[0];
    }
}

// tests per call up to which the merge takes a wavefront per test
constexpr int64_t kWaveMergeTests = 16384;

// FP64 merge of the chunk partials of every test (chunk order); out planes as in k_soft
template <bool ACC>
__global__ void k_soft_merge(const HaloDesc* __restrict__ halos,
                             const int* __restrict__ tst_halo, int64_t n_tst,
                             const float2* __restrict__ part, int64_t n_part,
                             double* __restrict__ pot, double* __restrict__ acc,
                             double* __restrict__ wsum) {
    const int64_t g = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (g >= n_tst) return;
    const HaloDesc d = halos[tst_halo[g]];
    const int64_t p = g - d.tst_off;
    const int nch = d.src_len > 0 ? (d.src_len + d.chunk - 1) / d.chunk : 0;
    constexpr int NQ = ACC ? 4 : 1;
    double s[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) s[q] = 0.0;
    for (int c = 0; c < nch; ++c) {
        const int64_t o = d.part_off + static_cast<int64_t>(c) * d.stride + p;
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            const float2 v = part[q * n_part + o];
            s[q] += static_cast<double>(v.x) + static_cast<double>(v.y);
        }
    }
    if (ACC) {
        if (acc) {
            acc[3 * g + 0] = s[0];
            acc[3 * g + 1] = s[1];
            acc[3 * g + 2] = s[2];
        }
        if (wsum) wsum[g] = s[NQ - 1];
    } else if (pot) {
        pot[g] = s[0];
    }
}

__global__ void k_soft_pack(const float* __restrict__ m, const float* __restrict__ x,
                            const float* __restrict__ y, const float* __restrict__ z, int64_t n,
                            float4* __restrict__ out) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) out[i] = make_float4(x[i], y[i], z[i], m[i]);
}

struct SoftWs {
    float4* src4;
    HaloDesc* halos;
    int4* items;
    float2* part;
    int* tst_halo;
    int* head;
    size_t bytes;
};

SoftWs carve_soft(void* base, const PotPlan& p) {
    Carver c(base);
    SoftWs w;
    w.src4 = c.take<float4>(std::max<int64_t>(p.n_src, 1));
    w.halos = c.take<HaloDesc>(std::max<size_t>(p.halos.size(), 1));
    w.items = c.take<int4>(std::max<int64_t>(p.n_items, 1));
    w.part = c.take<float2>(4 * std::max<int64_t>(p.n_part, 1));
    w.tst_halo = c.take<int>(std::max<int64_t>(p.n_tst, 1));
    w.head = c.take<int>(64);
    w.bytes = c.used;
    return w;
}

int check_offsets(int64_t H, const int64_t* off) {
    for (int64_t h = 0; h < H; ++h)
        if (off[h + 1] < off[h]) return set_error(HALMA_E_ARG, "offsets must not decrease");
    return 0;
}

}  // namespace

}  // namespace halma

using namespace halma;

extern "C" {

size_t halma_soft_workspace_size(int64_t n_halo, const int64_t* src_off,
                                 const int64_t* tst_off) {
    if (n_halo < 0 || !src_off || !tst_off) return 0;
    PotPlan plan;
    if (plan_pot(n_halo, src_off, tst_off, plan, false, 1)) return 0;
    return carve_soft(nullptr, plan).bytes;
}

}  // extern "C"

namespace {

// The body of both entries: sources as separate arrays (packed here into the
// workspace) or, with src_xyzm, already packed by the caller as (x, y, z, m) float4.
int soft_batched(int64_t n_halo, const int64_t* src_off, const float* sm, const float* sx,
                 const float* sy, const float* sz, const float* src_xyzm,
                 const int64_t* tst_off, const double* tx, const double* ty, const double* tz,
                 double eps, double* pot, double* acc, double* wsum, void* ws,
                 size_t ws_bytes, void* stream) {
    if (n_halo < 0 || !src_off || !tst_off)
        return set_error(HALMA_E_ARG, "halma_pot_soft_batched: bad halo count or offsets");
    if (!(eps >= 0.0) || !isfinite(eps)) return set_error(HALMA_E_ARG, "eps must be >= 0");
    if (n_halo == 0) return 0;
    int rc = check_offsets(n_halo, src_off);
    if (!rc) rc = check_offsets(n_halo, tst_off);
    if (rc) return rc;
    const int64_t n_src = src_off[n_halo] - src_off[0];
    const int64_t n_tst = tst_off[n_halo] - tst_off[0];
    if (n_tst == 0) return 0;
    if (!tx || !ty || !tz ||
        (n_src > 0 && !src_xyzm && (!sm || !sx || !sy || !sz)))
        return set_error(HALMA_E_ARG, "halma_pot_soft_batched: null pointer");
    if (src_xyzm && reinterpret_cast<uintptr_t>(src_xyzm) % alignof(float4) != 0)
        return set_error(HALMA_E_ARG, "packed sources must be 16-byte aligned");
    hipStream_t st = static_cast<hipStream_t>(stream);
    PotPlan plan;
    rc = plan_pot(n_halo, src_off, tst_off, plan, true, 1);
    if (rc) return rc;
    SoftWs w = carve_soft(ws, plan);
    if (ws_bytes < w.bytes)
        return set_error(HALMA_E_WORKSPACE, "workspace %zu < required %zu", ws_bytes, w.bytes);
    const int64_t s0 = src_off[0], t0 = tst_off[0];
    for (auto& d : plan.halos) {  // the descriptors index the packed copies from zero
        d.src_off -= s0;
        d.tst_off -= t0;
    }
    const float4* src4 = w.src4;
    if (src_xyzm) {
        src4 = reinterpret_cast<const float4*>(src_xyzm) + s0;
    } else if (n_src > 0) {
        k_soft_pack<<<static_cast<unsigned>((n_src + 255) / 256), 256, 0, st>>>(
            sm + s0, sx + s0, sy + s0, sz + s0, n_src, w.src4);
        HALMA_LAUNCH_CHECK();
    }
    rc = upload_async(w.halos, plan.halos.data(), plan.halos.size() * sizeof(HaloDesc), st);
    if (!rc && plan.n_items > 0)
        rc = upload_async(w.items, plan.items.data(), plan.items.size() * sizeof(int4), st);
    if (rc) return rc;
    k_soft_tst_halo<<<static_cast<unsigned>(n_halo), 256, 0, st>>>(
        w.halos, static_cast<int>(n_halo), w.tst_halo);
    HALMA_LAUNCH_CHECK();
    const float e2 = static_cast<float>(eps * eps);
    const int n_items = static_cast<int>(plan.n_items);
    const int grid = std::max(1, std::min(plan.grid, (n_items + 3) / 4));
    const unsigned mblocks = static_cast<unsigned>((n_tst + 255) / 256);
    const double* txo = tx + t0;
    const double* tyo = ty + t0;
    const double* tzo = tz + t0;
    for (int pass = 0; pass < 2; ++pass) {
        const bool want_acc = pass == 1;
        if (want_acc ? (!acc && !wsum) : !pot) continue;
        k_soft_clear<<<1, 64, 0, st>>>(w.head);
        HALMA_LAUNCH_CHECK();
        if (n_items > 0) {
            if (want_acc)
                k_soft<true><<<grid, 256, 0, st>>>(src4, txo, tyo, tzo, w.halos, w.items,
                                                   n_items, e2, w.part, plan.n_part, w.head);
            else
                k_soft<false><<<grid, 256, 0, st>>>(src4, txo, tyo, tzo, w.halos, w.items,
                                                    n_items, e2, w.part, plan.n_part, w.head);
            HALMA_LAUNCH_CHECK();
        }
        double* acc_o = acc ? acc + 3 * t0 : nullptr;
        double* wsum_o = wsum ? wsum + t0 : nullptr;
        if (n_tst <= kWaveMergeTests) {
            const unsigned wblocks = static_cast<unsigned>((n_tst * kWave + 255) / 256);
            if (want_acc)
                k_soft_merge_wave<true><<<wblocks, 256, 0, st>>>(
                    w.halos, w.tst_halo, n_tst, w.part, plan.n_part, nullptr, acc_o, wsum_o);
            else
                k_soft_merge_wave<false><<<wblocks, 256, 0, st>>>(
                    w.halos, w.tst_halo, n_tst, w.part, plan.n_part, pot + t0, nullptr, nullptr);
        } else if (want_acc) {
            k_soft_merge<true><<<mblocks, 256, 0, st>>>(w.halos, w.tst_halo, n_tst, w.part,
                                                        plan.n_part, nullptr, acc_o, wsum_o);
        } else {
            k_soft_merge<false><<<mblocks, 256, 0, st>>>(w.halos, w.tst_halo, n_tst, w.part,
                                                         plan.n_part, pot + t0, nullptr,
                                                         nullptr);
        }
        HALMA_LAUNCH_CHECK();
    }
    return 0;
}

}  // namespace

extern "C" {

int halma_pot_soft_batched(int64_t n_halo, const int64_t* src_off, const float* sm,
                           const float* sx, const float* sy, const float* sz,
                           const int64_t* tst_off, const double* tx, const double* ty,
                           const double* tz, double eps, double* pot, double* acc,
                           double* wsum, void* ws, size_t ws_bytes, void* stream) {
    return soft_batched(n_halo, src_off, sm, sx, sy, sz, nullptr, tst_off, tx, ty, tz, eps,
                        pot, acc, wsum, ws, ws_bytes, stream);
}

int halma_pot_soft_batched_packed(int64_t n_halo, const int64_t* src_off,
                                  const float* src_xyzm, const int64_t* tst_off,
                                  const double* tx, const double* ty, const double* tz,
                                  double eps, double* pot, double* acc, double* wsum, void* ws,
                                  size_t ws_bytes, void* stream) {
    if (!src_xyzm && n_halo > 0 && src_off && src_off[n_halo] > src_off[0])
        return set_error(HALMA_E_ARG, "halma_pot_soft_batched_packed: null sources");
    return soft_batched(n_halo, src_off, nullptr, nullptr, nullptr, nullptr, src_xyzm, tst_off,
                        tx, ty, tz, eps, pot, acc, wsum, ws, ws_bytes, stream);
}

}  // extern "C"
