// Self-potential of many haloes with the pair symmetry used: S_i = sum_{j != i} m_j / r_ij
// over the halo's own members (the sum of particle_subroutines.f90:466-514 with tests =
// sources, as the unbinding's full passes need), each pair distance computed once and
// its two terms, m_j / r to S_i and m_i / r to S_j, accumulated together.
//
// A halo of n members is cut into tiles of kTile = 128 members and the tiles into blocks
// of kSymK = 4 tiles.  A work item is a pair of blocks (A, B), A <= B.  Tests are held in
// registers (a tile of A: 2 per lane), the members of B's tiles streamed as wave-uniform
// sources; accumulated are
//   - test side: FP32 sums of 8 terms, Kahan-added, per test;
//   - source side: per source, the wave-wide sum of its terms over a tile's tests (a
//     fixed reduction tree across the lanes), Kahan-added in LDS over the tiles of A in
//     order.
// The test side keeps one Kahan sum per source tile; the tiles' sums are added in FP64
// in tile order.  A launch with many items gives each item to one wavefront (all tiles of
// B in turn); a launch with fewer (a lone large halo, the last passes of an unbinding)
// gives each to a workgroup whose wavefront w takes tile w of B, so that an item lasts
// four tile pairs instead of sixteen, and one with fewer than 1024 items to a 1024-thread
// workgroup with one tile pair per wavefront (k_pot_sym_fine).  Both forms do the same arithmetic in
// the same order: a halo's sums do not depend on the other haloes of the call.  Within a diagonal item
// (A == B) only pairs of tiles a <= b are taken, and a tile against itself is summed on
// the test side alone.  Every member then has kSymK-block partial sums in nb + 1 slots
// (nb = blocks of its halo), each written by exactly one item: slot s < I (its block) from the source side of item (s, I), slot I from the test
// side of item (I, I), slots s > I from the test side of item (I, s), and slot nb from
// the source side of item (I, I).  The slots are merged in FP64 in slot order.
#include <math.h>

#include <algorithm>
#include <vector>

#include "halma_common.h"
#include "potential.h"
#include "staging.h"
#include "sym.h"
#include "tuning.h"

namespace halma {

namespace {

constexpr int kTile = 2 * kWave;  // 128 members: two tests per lane

__device__ __forceinline__ float inv_r(float4 q, float x, float y, float z) {
    const float dx = q.x - x;
    const float dy = q.y - y;
    const float dz = q.z - z;
    const float d2 = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
    const float mn = fminf(fminf(fabsf(dx), fabsf(dy)), fabsf(dz));
    return mn != 0.0f ? __builtin_amdgcn_rsqf(d2) : 0.0f;  // the rule of pair_acc
}

// Kahan-compensated addition: c holds the low part lost so far (s + c is the sum; the
// sequence of s is that of the textbook form, whose compensation is -c)
__device__ __forceinline__ void kahan(float& s, float& c, float v) {
    const float y = v + c;
    const float t = s + y;
    c = y - (t - s);
    s = t;
}

template <int CTRL>
__device__ __forceinline__ float dpp_move(float v) {
    // lanes whose source lies outside their row of 16 read 0 (never used below)
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xf, 0xf,
                                                      true));
}

// v[0..7] per lane -> the sum over all 64 lanes of v[k], for k = source index of the lane:
// k = 4*bit5 + 2*bit4 + bit3 of the lane id, valid in the lanes with lane % 8 == 0.
// A fixed binary tree of depth 6 per total: recursive halving over lane bits 5 and 4 with
// the CDNA4 half-wave and row swaps (v_permlane32_swap / v_permlane16_swap: each lane
// ends with its half's pair sum of v[k] or v[k + 4], then of w[k] or w[k + 2]), over
// bit 3 with a row rotation by 8, then lane + (lane + 4), + (lane + 2), + (lane + 1)
// within the row (DPP row shifts).
__device__ __forceinline__ float reduce8(const float v[8], int lane) {
    float w[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const auto r = __builtin_amdgcn_permlane32_swap(__float_as_uint(v[k]),
                                                        __float_as_uint(v[k + 4]), false, false);
        w[k] = __uint_as_float(r[0]) + __uint_as_float(r[1]);
    }
    float u[2];
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const auto r = __builtin_amdgcn_permlane16_swap(__float_as_uint(w[k]),
                                                        __float_as_uint(w[k + 2]), false, false);
        u[k] = __uint_as_float(r[0]) + __uint_as_float(r[1]);
    }
    const bool h3 = (lane & 8) != 0;
    float t = (h3 ? u[1] : u[0]) + dpp_move<0x128>(h3 ? u[0] : u[1]);  // row_ror:8
    t += dpp_move<0x104>(t);                                           // row_shl:4
    t += dpp_move<0x102>(t);                                           // row_shl:2
    t += dpp_move<0x101>(t);                                           // row_shl:1
    return t;
}

// Sources q[g .. g+7] of a tile (cnt of them exist) against the lane's two tests: the
// test-side partial sums ps[r] (8 terms, source order) and the per-source pair terms
// pj[s] = m_i0 / r + m_i1 / r.  A full group (g + 8 <= cnt, wave-uniform) needs no
// padding select; the tail group clamps its loads and gives padding sources no mass.
template <bool kFull>
__device__ __forceinline__ void pair_group(const float4* __restrict__ q0, int g, int cnt,
                                           const float xi[2], const float yi[2],
                                           const float zi[2], const float mi[2], float ps[2],
                                           float pj[8]) {
#pragma unroll
    for (int s = 0; s < 8; ++s) {
        const int j = g + s;
        const float4 q = uniform_load(q0 + (kFull ? j : min(j, cnt - 1)));
        const float qw = (kFull || j < cnt) ? q.w : 0.0f;
        const float v0 = inv_r(q, xi[0], yi[0], zi[0]);
        const float v1 = inv_r(q, xi[1], yi[1], zi[1]);
        ps[0] = __builtin_fmaf(qw, v0, ps[0]);
        ps[1] = __builtin_fmaf(qw, v1, ps[1]);
        pj[s] = __builtin_fmaf(mi[1], v1, mi[0] * v0);
    }
}

// Below this many items in a launch an item is taken by a workgroup (four wavefronts,
// one source tile each), else by one wavefront: a wavefront per item fills the device
// once there are a few items per SIMD and keeps diagonal items balanced; with fewer
// items the workgroup form cuts an item's latency fourfold.
constexpr int kGroupItems = 2048;

// Below this many items an item is taken by a 1024-thread workgroup whose sixteen
// wavefronts take its sixteen tile pairs (k_pot_sym_fine): an item then lasts one tile
// pair, for the last passes of an unbinding where a few small haloes take full passes.
constexpr int kFineItems = 1024;

// One item per wavefront: tiles of A in turn against all tiles of B; test side a Kahan
// sum per tile of B, the tiles' sums added in FP64 in tile order; source side in the
// wavefront's LDS (ls, lc: kSymK tiles each).  Same arithmetic, in the same order, as
// the workgroup form below, so both give the same bits.
__device__ __forceinline__ void sym_items_wave(const float4* __restrict__ src,
                                               const SymDesc* __restrict__ halos,
                                               const int* __restrict__ live_n,
                                               const int4* __restrict__ items, int n_items,
                                               float2* __restrict__ part, int* __restrict__ head,
                                               float* ls, float* lc) {
    const int lane = __lane_id();
    const int wv = threadIdx.x / kWave;
    const int n_waves = static_cast<int>(gridDim.x * (blockDim.x / kWave));
    int k = __builtin_amdgcn_readfirstlane(static_cast<int>(blockIdx.x * (blockDim.x / kWave) + wv));
    for (bool first = true;; first = false) {
        if (!first) {
            int t = 0;
            if (lane == 0) t = atomicAdd(head, 1);
            k = n_waves + __builtin_amdgcn_readfirstlane(t);
        }
        if (k >= n_items) break;
        const int4 it = uniform_load(items + k);
        const SymDesc d = uniform_load(halos + it.x);
        const int n = live_n ? __builtin_amdgcn_readfirstlane(uniform_load(live_n + it.x)) : d.n;
        const int nt = (n + kTile - 1) / kTile;
        const int nb = (nt + kSymK - 1) / kSymK;
        const int A = it.y, B = it.z;
        if (B >= nb || A > B) continue;
        const float4* sp = src + d.off;
        float2* pp = part + d.part_off;
        const int64_t stride = d.n;  // slots are laid out for the planned size
#pragma unroll
        for (int e = 0; e < kSymK * kTile / kWave; ++e) {
            ls[e * kWave + lane] = 0.0f;
            lc[e * kWave + lane] = 0.0f;
        }
        for (int ta = 0; ta < kSymK; ++ta) {
            const int a = A * kSymK + ta;
            if (a >= nt) break;
            float xi[2], yi[2], zi[2], mi[2];
            double S[2] = {0.0, 0.0};
            int ii[2];
#pragma unroll
            for (int r = 0; r < 2; ++r) {
                ii[r] = a * kTile + r * kWave + lane;
                const float4 p = sp[min(ii[r], n - 1)];
                xi[r] = p.x;
                yi[r] = p.y;
                zi[r] = p.z;
                mi[r] = ii[r] < n ? p.w : 0.0f;  // padding lanes add nothing to sources
            }
            for (int tb = 0; tb < kSymK; ++tb) {
                const int b = B * kSymK + tb;
                if (b >= nt) break;
                if (A == B && b < a) continue;
                const int j0 = b * kTile;
                const int cnt = min(kTile, n - j0);
                const bool diag = b == a;
                float as[2] = {0.0f, 0.0f}, ac[2] = {0.0f, 0.0f};
                for (int g = 0; g < cnt; g += 8) {
                    float ps[2] = {0.0f, 0.0f};
                    float pj[8];
                    if (g + 8 <= cnt)
                        pair_group<true>(sp + j0, g, cnt, xi, yi, zi, mi, ps, pj);
                    else
                        pair_group<false>(sp + j0, g, cnt, xi, yi, zi, mi, ps, pj);
                    kahan(as[0], ac[0], ps[0]);
                    kahan(as[1], ac[1], ps[1]);
                    if (!diag) {
                        const float tot = reduce8(pj, lane);
                        const int s = (lane >> 3) & 7;
                        if ((lane & 7) == 0 && g + s < cnt) {
                            const int e = tb * kTile + g + s;
                            float sv = ls[e], cv = lc[e];
                            kahan(sv, cv, tot);
                            ls[e] = sv;
                            lc[e] = cv;
                        }
                    }
                }
#pragma unroll
                for (int r = 0; r < 2; ++r)
                    S[r] += static_cast<double>(as[r]) + static_cast<double>(ac[r]);
            }
            // test side of tile a from this item: slot B (mod the band width), the source
            // tiles' sums added in FP64 in tile order (the same bits as a workgroup's)
            const int tslot = B % d.wb;
#pragma unroll
            for (int r = 0; r < 2; ++r)
                if (ii[r] < n) {
                    const float hi = static_cast<float>(S[r]);
                    pp[static_cast<int64_t>(tslot) * stride + ii[r]] =
                        make_float2(hi, static_cast<float>(S[r] - static_cast<double>(hi)));
                }
        }
        // source side of B's tiles from this item: slot A (mod the band width), or the
        // diagonal slot min(wb, nb)
        const int slot = A < B ? A % d.wb : min(d.wb, nb);
        for (int tb = 0; tb < kSymK; ++tb) {
            const int b = B * kSymK + tb;
            if (b >= nt) break;
            const int j0 = b * kTile;
            const int cnt = min(kTile, n - j0);
#pragma unroll
            for (int r = 0; r < 2; ++r) {
                const int e = r * kWave + lane;
                if (e < cnt)
                    pp[static_cast<int64_t>(slot) * stride + j0 + e] =
                        make_float2(ls[tb * kTile + e], lc[tb * kTile + e]);
            }
        }
    }
}

// One item per workgroup: wavefront w streams tile w of B against each tile of A in turn;
// its source side stays in its LDS (ls, lc: one tile), its test-side sums per tile of A go
// to tpart[ta][w] and are added in FP64 in tile order once the four are done.
__device__ __forceinline__ void sym_items_group(const float4* __restrict__ src,
                                                const SymDesc* __restrict__ halos,
                                                const int* __restrict__ live_n,
                                                const int4* __restrict__ items, int n_items,
                                                float2* __restrict__ part,
                                                int* __restrict__ head, float* ls, float* lc,
                                                float2* tpart, int* next_item) {
    const int lane = __lane_id();
    const int wv = threadIdx.x / kWave;
    int k = static_cast<int>(blockIdx.x);
    for (bool first = true;; first = false) {
        if (!first) {
            __syncthreads();  // the previous item's LDS has been read
            if (threadIdx.x == 0) *next_item = atomicAdd(head, 1);
            __syncthreads();
            k = static_cast<int>(gridDim.x) + *next_item;
        }
        if (k >= n_items) break;
        const int4 it = uniform_load(items + k);
        const SymDesc d = uniform_load(halos + it.x);
        const int n = live_n ? __builtin_amdgcn_readfirstlane(uniform_load(live_n + it.x)) : d.n;
        const int nt = (n + kTile - 1) / kTile;
        const int nb = (nt + kSymK - 1) / kSymK;
        const int A = it.y, B = it.z;
        if (B >= nb || A > B) continue;
        const float4* sp = src + d.off;
        float2* pp = part + d.part_off;
        const int64_t stride = d.n;  // slots are laid out for the planned size
        const int b = B * kSymK + wv;  // this wavefront's source tile
        const int j0 = b * kTile;
        const int cnt = b < nt ? min(kTile, n - j0) : 0;
#pragma unroll
        for (int e = 0; e < kTile / kWave; ++e) {
            ls[e * kWave + lane] = 0.0f;
            lc[e * kWave + lane] = 0.0f;
        }
        for (int ta = 0; ta < kSymK; ++ta) {
            const int a = A * kSymK + ta;
            if (a >= nt) break;
            float as[2] = {0.0f, 0.0f}, ac[2] = {0.0f, 0.0f};
            if (cnt > 0 && !(A == B && b < a)) {
                float xi[2], yi[2], zi[2], mi[2];
#pragma unroll
                for (int r = 0; r < 2; ++r) {
                    const int i = a * kTile + r * kWave + lane;
                    const float4 p = sp[min(i, n - 1)];
                    xi[r] = p.x;
                    yi[r] = p.y;
                    zi[r] = p.z;
                    mi[r] = i < n ? p.w : 0.0f;  // padding lanes add nothing to sources
                }
                const bool diag = b == a;
                for (int g = 0; g < cnt; g += 8) {
                    float ps[2] = {0.0f, 0.0f};
                    float pj[8];
                    if (g + 8 <= cnt)
                        pair_group<true>(sp + j0, g, cnt, xi, yi, zi, mi, ps, pj);
                    else
                        pair_group<false>(sp + j0, g, cnt, xi, yi, zi, mi, ps, pj);
                    kahan(as[0], ac[0], ps[0]);
                    kahan(as[1], ac[1], ps[1]);
                    if (!diag) {
                        const float tot = reduce8(pj, lane);
                        const int s = (lane >> 3) & 7;
                        if ((lane & 7) == 0 && g + s < cnt) {
                            float sv = ls[g + s], cv = lc[g + s];
                            kahan(sv, cv, tot);
                            ls[g + s] = sv;
                            lc[g + s] = cv;
                        }
                    }
                }
            }
#pragma unroll
            for (int r = 0; r < 2; ++r)
                tpart[(ta * kSymK + wv) * kTile + r * kWave + lane] = make_float2(as[r], ac[r]);
        }
        __syncthreads();
        // test side of A's tiles from this item: slot B (mod the band width); the four
        // source tiles' compensated sums added in FP64 in tile order, stored as a pair of
        // floats (hi, lo) whose sum keeps the FP64 value to u^2
        const int tslot = B % d.wb;
        for (int t = threadIdx.x; t < kSymK * kTile; t += blockDim.x) {
            const int ta = t / kTile, e = t % kTile;
            const int i = (A * kSymK + ta) * kTile + e;
            if (A * kSymK + ta >= nt || i >= n) continue;
            double S = 0.0;
#pragma unroll
            for (int w = 0; w < kSymK; ++w) {
                const float2 v = tpart[(ta * kSymK + w) * kTile + e];
                S += static_cast<double>(v.x) + static_cast<double>(v.y);
            }
            const float hi = static_cast<float>(S);
            pp[static_cast<int64_t>(tslot) * stride + i] =
                make_float2(hi, static_cast<float>(S - static_cast<double>(hi)));
        }
        // source side of this wavefront's tile of B: slot A (mod the band width), or the
        // diagonal slot min(wb, nb)
        const int slot = A < B ? A % d.wb : min(d.wb, nb);
#pragma unroll
        for (int r = 0; r < 2; ++r) {
            const int e = r * kWave + lane;
            if (e < cnt)
                pp[static_cast<int64_t>(slot) * stride + j0 + e] = make_float2(ls[e], lc[e]);
        }
    }
}

__global__ __launch_bounds__(256) void k_pot_sym(const float4* __restrict__ src,
                                                 const SymDesc* __restrict__ halos,
                                                 const int* __restrict__ live_n,
                                                 const int4* __restrict__ items, int n_items,
                                                 const int64_t* __restrict__ n_items_dev,
                                                 float2* __restrict__ part,
                                                 int* __restrict__ head) {
    static_assert(kSymK * kWave == 256, "one wavefront per tile of a block");
    // LDS of either form: a wavefront per item keeps (sum, comp) of kSymK source tiles
    // per wavefront (16 KiB); a workgroup per item one source tile per wavefront (4 KiB)
    // and the test-side sums of each (tile of A, wavefront) (16 KiB)
    __shared__ float2 lds[2560];
    __shared__ int next_item;
    float* f = reinterpret_cast<float*>(lds);
    const int wv = threadIdx.x / kWave;
    if (n_items_dev) {
        const int64_t nd = uniform_load(n_items_dev);
        n_items = static_cast<int>(nd < n_items ? nd : n_items);
    }
    if (n_items < kFineItems) return;  // k_pot_sym_fine takes the launch
    if (n_items >= kGroupItems)
        sym_items_wave(src, halos, live_n, items, n_items, part, head,
                       f + wv * kSymK * kTile, f + (kSymK + wv) * kSymK * kTile);
    else
        sym_items_group(src, halos, live_n, items, n_items, part, head, f + wv * kTile,
                        f + (kSymK + wv) * kTile, lds + kSymK * kTile, &next_item);
}

// One item per 1024-thread workgroup (launches of fewer than kFineItems items): wavefront
// 4 ta + w takes tile ta of A against tile w of B.  The pair's test-side compensated sum
// goes to tpart[ta][w], its source-side terms (one per source of tile w, the wave-wide sum
// over tile ta's tests) to spart[ta][w]; then the test side of each tile of A is added in
// FP64 over the tiles of B in tile order and the source side of each tile of B is
// Kahan-added over the tiles of A in order, skipping the pairs the other forms skip: the
// same operations in the same order as the other forms, so the same bits.
__global__ __launch_bounds__(1024) void k_pot_sym_fine(const float4* __restrict__ src,
                                                       const SymDesc* __restrict__ halos,
                                                       const int* __restrict__ live_n,
                                                       const int4* __restrict__ items,
                                                       int n_items,
                                                       const int64_t* __restrict__ n_items_dev,
                                                       float2* __restrict__ part,
                                                       int* __restrict__ head) {
    static_assert(kSymK * kSymK * kWave == 1024, "one wavefront per tile pair of an item");
    __shared__ float2 tpart[kSymK * kSymK * kTile];  // 16 KiB
    __shared__ float spart[kSymK * kSymK * kTile];   // 8 KiB
    __shared__ int next_item;
    if (n_items_dev) {
        const int64_t nd = uniform_load(n_items_dev);
        n_items = static_cast<int>(nd < n_items ? nd : n_items);
    }
    if (n_items >= kFineItems) return;  // k_pot_sym takes the launch
    const int lane = __lane_id();
    const int wv = threadIdx.x / kWave;
    const int ta = wv / kSymK, w = wv % kSymK;
    float2* tp = tpart + wv * kTile;
    float* spw = spart + wv * kTile;
    int k = static_cast<int>(blockIdx.x);
    for (bool first = true;; first = false) {
        if (!first) {
            __syncthreads();  // the previous item's LDS has been read
            if (threadIdx.x == 0) next_item = atomicAdd(head, 1);
            __syncthreads();
            k = static_cast<int>(gridDim.x) + next_item;
        }
        if (k >= n_items) break;
        const int4 it = uniform_load(items + k);
        const SymDesc d = uniform_load(halos + it.x);
        const int n = live_n ? __builtin_amdgcn_readfirstlane(uniform_load(live_n + it.x)) : d.n;
        const int nt = (n + kTile - 1) / kTile;
        const int nb = (nt + kSymK - 1) / kSymK;
        const int A = it.y, B = it.z;
        if (B >= nb || A > B) continue;  // the same for the whole workgroup
        const float4* sp = src + d.off;
        float2* pp = part + d.part_off;
        const int64_t stride = d.n;  // slots are laid out for the planned size
        const int a = A * kSymK + ta, b = B * kSymK + w;
        const int j0 = b * kTile;
        const int cnt = b < nt ? min(kTile, n - j0) : 0;
#pragma unroll
        for (int r = 0; r < 2; ++r) spw[r * kWave + lane] = 0.0f;
        float as[2] = {0.0f, 0.0f}, ac[2] = {0.0f, 0.0f};
        if (a < nt && cnt > 0 && !(A == B && b < a)) {
            float xi[2], yi[2], zi[2], mi[2];
#pragma unroll
            for (int r = 0; r < 2; ++r) {
                const int i = a * kTile + r * kWave + lane;
                const float4 p = sp[min(i, n - 1)];
                xi[r] = p.x;
                yi[r] = p.y;
                zi[r] = p.z;
                mi[r] = i < n ? p.w : 0.0f;  // padding lanes add nothing to sources
            }
            const bool diag = b == a;
            for (int g = 0; g < cnt; g += 8) {
                float ps[2] = {0.0f, 0.0f};
                float pj[8];
                if (g + 8 <= cnt)
                    pair_group<true>(sp + j0, g, cnt, xi, yi, zi, mi, ps, pj);
                else
                    pair_group<false>(sp + j0, g, cnt, xi, yi, zi, mi, ps, pj);
                kahan(as[0], ac[0], ps[0]);
                kahan(as[1], ac[1], ps[1]);
                if (!diag) {
                    const float tot = reduce8(pj, lane);
                    const int s = (lane >> 3) & 7;
                    if ((lane & 7) == 0 && g + s < cnt) spw[g + s] = tot;
                }
            }
        }
#pragma unroll
        for (int r = 0; r < 2; ++r) tp[r * kWave + lane] = make_float2(as[r], ac[r]);
        __syncthreads();
        // test side of A's tiles: slot B (mod the band width), the four source tiles'
        // compensated sums added in FP64 in tile order, kept as a float pair (hi, lo)
        const int tslot = B % d.wb;
        for (int t = threadIdx.x; t < kSymK * kTile; t += blockDim.x) {
            const int t_a = t / kTile, e = t % kTile;
            const int i = (A * kSymK + t_a) * kTile + e;
            if (A * kSymK + t_a >= nt || i >= n) continue;
            double S = 0.0;
#pragma unroll
            for (int t_b = 0; t_b < kSymK; ++t_b) {
                const float2 v = tpart[(t_a * kSymK + t_b) * kTile + e];
                S += static_cast<double>(v.x) + static_cast<double>(v.y);
            }
            const float hi = static_cast<float>(S);
            pp[static_cast<int64_t>(tslot) * stride + i] =
                make_float2(hi, static_cast<float>(S - static_cast<double>(hi)));
        }
        // source side of B's tiles: slot A (mod the band width), or the diagonal slot
        // min(wb, nb); the terms from the tiles of A Kahan-added in tile order
        const int slot = A < B ? A % d.wb : min(d.wb, nb);
        for (int t = threadIdx.x; t < kSymK * kTile; t += blockDim.x) {
            const int t_b = t / kTile, e = t % kTile;
            const int bb = B * kSymK + t_b;
            if (bb >= nt) continue;
            const int jb = bb * kTile;
            if (e >= min(kTile, n - jb)) continue;
            float sv = 0.0f, cv = 0.0f;
            for (int t_a = 0; t_a < kSymK; ++t_a) {
                const int aa = A * kSymK + t_a;
                if (aa >= nt) break;
                if ((A == B && bb < aa) || bb == aa) continue;
                kahan(sv, cv, spart[(t_a * kSymK + t_b) * kTile + e]);
            }
            pp[static_cast<int64_t>(slot) * stride + jb + e] = make_float2(sv, cv);
        }
    }
}

// The slots of one tile launch added into each affected member's FP64 sum (out, cleared
// before the first launch): segment s = {halo, first member, count, role << 24 | local
// slots}; the member adds its local slots 0 .. nloc - 1 in order, then, on a diagonal
// tile, the diagonal slot.  With one band a member's slots 0..nb are added in order.
__global__ __launch_bounds__(256) void k_sym_merge_tile(const SymDesc* __restrict__ halos,
                                                        const int4* __restrict__ segs, int nseg,
                                                        const int64_t* __restrict__ member_base,
                                                        const float2* __restrict__ part,
                                                        double* __restrict__ out) {
    for (int s = blockIdx.y; s < nseg; s += gridDim.y) {
        const int4 sg = segs[s];
        const SymDesc d = halos[sg.x];
        const int nloc = sg.w & 0xffffff;
        const bool diag = (sg.w >> 24) == 2;
        for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < sg.z; k += gridDim.x * blockDim.x) {
            const int64_t i = sg.y + k;
            const float2* pp = part + d.part_off + i;
            const int64_t g = member_base[sg.x] + i;
            double acc = out[g];
            for (int c = 0; c < nloc; c += 8) {
                float2 q[8];
#pragma unroll
                for (int u = 0; u < 8; ++u)
                    q[u] = c + u < nloc ? pp[static_cast<int64_t>(c + u) * d.n]
                                        : make_float2(0.0f, 0.0f);
#pragma unroll
                for (int u = 0; u < 8; ++u)
                    if (c + u < nloc) acc += static_cast<double>(q[u].x) + static_cast<double>(q[u].y);
            }
            if (diag) {
                const float2 q = pp[static_cast<int64_t>(min(d.wb, d.nb)) * d.n];
                acc += static_cast<double>(q.x) + static_cast<double>(q.y);
            }
            out[g] = acc;
        }
    }
}

__global__ void k_sym_zero(double* __restrict__ p, int64_t n) {
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += static_cast<int64_t>(gridDim.x) * blockDim.x)
        p[i] = 0.0;
}

__global__ void k_sym_pack(const float* __restrict__ m, const float* __restrict__ x,
                           const float* __restrict__ y, const float* __restrict__ z, int64_t n,
                           float4* __restrict__ out) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) out[i] = make_float4(x[i], y[i], z[i], m[i]);
}

// One block per tile (P, Q) of a halo: its items (A, B), B-major, A <= B, from the
// tile's first item index on.
__global__ __launch_bounds__(256) void k_sym_expand(const int4* __restrict__ tiles,
                                                    const SymDesc* __restrict__ halos,
                                                    int4* __restrict__ items) {
    const int4 t = tiles[blockIdx.x];
    const SymDesc d = halos[t.x];
    const int wb = d.wb, nb = d.nb;
    const int a0 = t.y * wb, a1 = min(nb, (t.y + 1) * wb);
    const int b0 = t.z * wb, b1 = min(nb, (t.z + 1) * wb);
    int64_t at = t.w;
    for (int B = b0; B < b1; ++B) {
        const int ae = t.y == t.z ? B + 1 : a1;  // diagonal tile: A <= B
        for (int A = a0 + static_cast<int>(threadIdx.x); A < ae; A += blockDim.x)
            items[at + (A - a0)] = make_int4(t.x, A, B, 0);
        at += ae - a0;
    }
}

__global__ void k_sym_clear(int* __restrict__ p) {
    if (threadIdx.x == 0) *p = 0;
}

struct SymWs {
    float4* src4;
    SymDesc* halos;
    int4* items;
    float2* part;
    int4* segs;
    int4* tiles;
    int64_t* base;
    int* head;
    size_t bytes;
};

// n_segs: the merge segments of a banded plan (at most two per halo and tile)
SymWs carve_sym(void* b, const SymPlan& p, int64_t H, int64_t n_segs) {
    Carver c(b);
    SymWs w;
    w.src4 = c.take<float4>(std::max<int64_t>(p.n_members, 1));
    w.halos = c.take<SymDesc>(std::max<int64_t>(H, 1));
    w.items = c.take<int4>(std::max<int64_t>(p.n_items, 1));
    w.part = c.take<float2>(std::max<int64_t>(p.n_part, 1));
    w.segs = c.take<int4>(std::max<int64_t>(n_segs, 1));
    w.tiles = c.take<int4>(std::max<size_t>(p.tiles.size(), 1));
    w.base = c.take<int64_t>(H + 1);
    w.head = c.take<int>(64);
    w.bytes = c.used;
    return w;
}

}  // namespace

int plan_sym(int64_t H, const int64_t* off, SymPlan& plan, bool with_items, int max_members,
             int band) {
    plan = SymPlan();
    plan.halos.resize(H);
    int64_t part = 0;
    std::vector<int64_t> work(H);
    for (int64_t h = 0; h < H; ++h) {
        const int64_t n = off[h + 1] - off[h];
        if (n < 0 || n > INT32_MAX) return set_error(HALMA_E_ARG, "halo %lld: bad offsets",
                                                      static_cast<long long>(h));
        SymDesc& d = plan.halos[h];
        d.off = off[h] - off[0];
        d.part_off = part;
        d.n = static_cast<int32_t>(n);
        d.nb = (max_members != 0 && d.n > max_members) ? 0 : sym_blocks(d.n);
        d.wb = (band > 0 && d.nb > band) ? band : d.nb + 1;
        d.pad = 0;
        part += static_cast<int64_t>(std::min(d.wb, d.nb) + 1) * n;
        work[h] = n;
    }
    plan.order = order_desc(work);
    int64_t n_items = 0;
    for (int64_t h = 0; h < H; ++h) {
        const int64_t nb = plan.halos[h].nb;
        n_items += nb * (nb + 1) / 2;
    }
    if (n_items > INT32_MAX) return set_error(HALMA_E_LIMIT, "too many work items");
    plan.n_items = n_items;
    if (with_items) {
        int64_t tiles = 1;
        for (int64_t h = 0; h < H; ++h) {
            const SymDesc& d = plan.halos[h];
            const int64_t K = d.nb > 0 ? (d.nb + d.wb - 1) / d.wb : 0;
            tiles = std::max<int64_t>(tiles, K * (K + 1) / 2);
        }
        plan.launch_off.push_back(0);
        plan.seg_off.push_back(0);
        int64_t at = 0;  // items listed so far
        for (int64_t t = 0; t < tiles; ++t) {
            for (int32_t h : plan.order) {
                const SymDesc& d = plan.halos[h];
                const int nb = d.nb, wb = d.wb;
                const int64_t K = nb > 0 ? (nb + wb - 1) / wb : 0;
                if (t >= K * (K + 1) / 2) continue;
                int Q = 0;
                while (static_cast<int64_t>(Q + 1) * (Q + 2) / 2 <= t) ++Q;
                const int P = static_cast<int>(t - static_cast<int64_t>(Q) * (Q + 1) / 2);
                const int b0 = Q * wb, b1 = std::min(nb, (Q + 1) * wb);
                const int a0 = P * wb, a1 = std::min(nb, (P + 1) * wb);
                plan.tiles.push_back(make_int4(h, P, Q, static_cast<int>(at)));
                at += P == Q ? static_cast<int64_t>(b1 - b0) * (b1 - b0 + 1) / 2
                             : static_cast<int64_t>(b1 - b0) * (a1 - a0);
                // the members of bands P and Q take this tile's slots into their sums
                auto seg = [&](int band, int role, int nloc) {
                    const int64_t m0 = static_cast<int64_t>(band) * wb * 4 * 128;
                    const int64_t m1 = std::min<int64_t>(d.n, m0 + static_cast<int64_t>(wb) * 512);
                    plan.segs.push_back(make_int4(h, static_cast<int>(m0),
                                                  static_cast<int>(m1 - m0), (role << 24) | nloc));
                };
                if (P == Q) {
                    seg(P, 2, b1 - b0);
                } else {
                    seg(P, 0, b1 - b0);
                    seg(Q, 1, a1 - a0);
                }
            }
            plan.launch_off.push_back(at);
            plan.seg_off.push_back(static_cast<int64_t>(plan.segs.size()));
        }
        if (at != n_items) return set_error(HALMA_E_STATE, "symmetric plan: item count");
    }
    plan.n_members = off[H] - off[0];
    plan.n_part = part;
    return 0;
}

int expand_sym_items(const int4* d_tiles, int64_t n_tiles, const SymDesc* d_halos, int4* items,
                     hipStream_t st) {
    if (n_tiles <= 0) return 0;
    if (n_tiles > INT32_MAX) return set_error(HALMA_E_LIMIT, "too many symmetric tiles");
    k_sym_expand<<<static_cast<unsigned>(n_tiles), 256, 0, st>>>(d_tiles, d_halos, items);
    HALMA_LAUNCH_CHECK();
    return 0;
}

int launch_pot_sym(const SymPlan& plan, const float4* src4, const SymDesc* d_halos,
                   const int* live_n, const int4* d_items, const int64_t* n_items_dev,
                   float2* part, int* head, hipStream_t st, bool reset_head) {
    if (reset_head) {
        k_sym_clear<<<1, 64, 0, st>>>(head);
        HALMA_LAUNCH_CHECK();
    }
    const int n_items = static_cast<int>(plan.n_items);
    if (n_items == 0) return 0;
    int cus = device_cu_count();
    if (cus <= 0) cus = 256;
    // both forms are queued when the item count is known only on the device and the plan
    // allows either (the one not chosen returns at once); the device count never exceeds
    // the planned n_items
    if (n_items >= kFineItems) {
        const int grid = std::max(1, std::min(cus * 8, n_items));  // a workgroup or more per item
        k_pot_sym<<<grid, 256, 0, st>>>(src4, d_halos, live_n, d_items, n_items, n_items_dev,
                                        part, head);
        HALMA_LAUNCH_CHECK();
    }
    if (n_items < kFineItems || n_items_dev) {
        const int grid = std::max(1, std::min(cus, n_items));
        k_pot_sym_fine<<<grid, 1024, 0, st>>>(src4, d_halos, live_n, d_items, n_items,
                                              n_items_dev, part, head);
        HALMA_LAUNCH_CHECK();
    }
    return 0;
}

}  // namespace halma

using namespace halma;

extern "C" {

double halma_pot_sym_error_bound(void) { return pot_sym_error_bound(); }

size_t halma_pot_sum_self_workspace_size(int64_t n_halo, const int64_t* offsets) {
    if (n_halo < 0 || !offsets) return 0;
    for (int64_t h = 0; h < n_halo; ++h)
        if (offsets[h + 1] < offsets[h]) return 0;
    SymPlan plan;
    if (plan_sym(n_halo, offsets, plan, true, 0, kSymBand)) return 0;
    return carve_sym(nullptr, plan, n_halo, static_cast<int64_t>(plan.segs.size())).bytes;
}

int halma_pot_sum_self_batched(int64_t n_halo, const int64_t* offsets, const float* m,
                               const float* x, const float* y, const float* z, double* out,
                               void* ws, size_t ws_bytes, void* stream) {
    if (n_halo < 0 || !offsets) return set_error(HALMA_E_ARG, "bad halo count or offsets");
    if (n_halo == 0) return 0;
    for (int64_t h = 0; h < n_halo; ++h)
        if (offsets[h + 1] < offsets[h]) return set_error(HALMA_E_ARG, "offsets must not decrease");
    const int64_t n = offsets[n_halo] - offsets[0];
    if (n == 0) return 0;
    if (!m || !x || !y || !z || !out) return set_error(HALMA_E_ARG, "null pointer");
    hipStream_t st = static_cast<hipStream_t>(stream);
    // haloes of more than kSymBand blocks run tile by tile (slots O(n kSymBand) per member)
    SymPlan plan;
    int rc = plan_sym(n_halo, offsets, plan, true, 0, kSymBand);
    if (rc) return rc;
    SymWs w = carve_sym(ws, plan, n_halo, static_cast<int64_t>(plan.segs.size()));
    if (ws_bytes < w.bytes)
        return set_error(HALMA_E_WORKSPACE, "workspace %zu < required %zu", ws_bytes, w.bytes);
    const int64_t o0 = offsets[0];
    k_sym_pack<<<static_cast<unsigned>((n + 255) / 256), 256, 0, st>>>(m + o0, x + o0, y + o0,
                                                                       z + o0, n, w.src4);
    HALMA_LAUNCH_CHECK();
    std::vector<int64_t> base(n_halo + 1);
    for (int64_t h = 0; h <= n_halo; ++h) base[h] = offsets[h] - o0;
    rc = upload_async(w.halos, plan.halos.data(), plan.halos.size() * sizeof(SymDesc), st);
    if (!rc) rc = upload_async(w.base, base.data(), base.size() * sizeof(int64_t), st);
    if (!rc && !plan.tiles.empty())
        rc = upload_async(w.tiles, plan.tiles.data(), plan.tiles.size() * sizeof(int4), st);
    if (!rc && !plan.segs.empty())
        rc = upload_async(w.segs, plan.segs.data(), plan.segs.size() * sizeof(int4), st);
    if (!rc)
        rc = expand_sym_items(w.tiles, static_cast<int64_t>(plan.tiles.size()), w.halos,
                              w.items, st);
    if (rc) return rc;
    double* o = out + o0;
    k_sym_zero<<<256, 256, 0, st>>>(o, n);
    HALMA_LAUNCH_CHECK();
    int cus = device_cu_count();
    if (cus <= 0) cus = 256;
    const int64_t T = static_cast<int64_t>(plan.launch_off.size()) - 1;
    for (int64_t t = 0; t < T; ++t) {
        const int64_t i0 = plan.launch_off[t], ni = plan.launch_off[t + 1] - i0;
        const int64_t s0 = plan.seg_off[t], ns = plan.seg_off[t + 1] - s0;
        if (ni > 0) {
            k_sym_clear<<<1, 64, 0, st>>>(w.head);
            HALMA_LAUNCH_CHECK();
            if (ni >= kFineItems) {
                const int grid = static_cast<int>(std::min<int64_t>(cus * 8, ni));
                k_pot_sym<<<grid, 256, 0, st>>>(w.src4, w.halos, nullptr, w.items + i0,
                                                static_cast<int>(ni), nullptr, w.part, w.head);
            } else {
                const int grid = static_cast<int>(std::min<int64_t>(cus, ni));
                k_pot_sym_fine<<<grid, 1024, 0, st>>>(w.src4, w.halos, nullptr, w.items + i0,
                                                      static_cast<int>(ni), nullptr, w.part,
                                                      w.head);
            }
            HALMA_LAUNCH_CHECK();
        }
        if (ns > 0) {
            const dim3 mg(32, static_cast<unsigned>(std::min<int64_t>(ns, 4096)));
            k_sym_merge_tile<<<mg, 256, 0, st>>>(w.halos, w.segs + s0, static_cast<int>(ns),
                                                 w.base, w.part, o);
            HALMA_LAUNCH_CHECK();
        }
    }
    return 0;
}

}  // extern "C"
