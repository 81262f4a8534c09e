// Batched iterative unbinding.
//
// The energy test of halo_properties.py:342-359 / halo_gas.py:456-476,
//   E_i = 1/2 |v_i - v_cm|^2 - factor * G * S_i,  bound iff E_i <= 0,
// with S_i the pair sum of particle_subroutines.f90:466-514 over the halo's bound
// members (positions and masses cast to FP32 as halo_gas.py:172-178 does), iterated per
// halo until a pass removes nobody: each pass recomputes v_cm (mass-weighted over the
// bound set, halo_properties.py:45-60) and S over the bound set and removes members with
// E > 0.  (The reference evaluates the test once; iterating it is this path's design.)
//
// One pass on the GPU, for all haloes at once:
//   1. compaction    k_chunk_sums / k_halo_sets, then k_items_write, over fixed chunks
//                    of 1024 member positions: stable compaction of each active halo's
//                    bound members into the front of its range (AoS float4), of the
//                    members removed by the previous pass into a second list, the
//                    double-double v_cm over the bound set, and the pass mode: FULL
//                    (first pass, or the previous pass removed more than half as many
//                    members as it kept) sums over the bound set afresh; INCREMENTAL
//                    subtracts the removed members' contributions from each bound
//                    member's stored sum, S_new = S_old - sum_{j removed} m_j / r_ij,
//                    which costs n_bound x n_removed pairs instead of n_bound^2.
//                    k_halo_sets also sets the live test/source counts and each halo's
//                    number of pair-kernel items; k_items_write writes both pair
//                    kernels' item lists (haloes in the plan's largest-first order) and,
//                    in its other blocks, the compacted lists;
//   2. pair kernels  full passes on the symmetric kernel (k_pot_sym, or k_pot_sym_fine
//                    for fewer than 1024 block pairs), incremental passes (and full
//                    passes of haloes planned without slots) on the chunked FP32 kernel
//                    k_pot_f32;
//   3. k_energy      merges the slots or chunk partials (FP64), updates S and its error
//                    bound (the bounds of every FP32 sum that went into S add up),
//                    evaluates E in FP64 and decides every member whose |E| exceeds
//                    factor*G*bound + 8 ulp(kin + factor*G*S);
//   4. k_retest      members inside the band: S in FP64 with the oracle's fixed summation
//                    order over the bound set (a block or a wavefront per member), so E
//                    and the decision are the oracle's bit for bit (S and its bound are
//                    reset to that value); stale members: a fresh FP32 S;
//   5. k_finish      per-halo pass counts, convergence flags, the removed total (a thread
//                    per halo).
// The parts are plain launches (or, on request, hipGraphs captured once and replayed);
// the host reads one counter per pass to stop.
#include <math.h>

#include <algorithm>
#include <mutex>
#include <vector>

#include "halma_common.h"
#include "potential.h"
#include "scan.h"
#include "sym.h"
#include "staging.h"
#include "tuning.h"

namespace halma {

namespace {

constexpr int kThreads = 256;
constexpr int kIncChunk = 256;      // smallest source chunk of an incremental pass
constexpr int kRetestBlock = 2048;  // live sets from this size are re-summed by a block

// member states
constexpr uint8_t kGone = 0;      // removed by an earlier pass
constexpr uint8_t kBound = 1;     // currently bound
constexpr uint8_t kJustGone = 2;  // removed by the last pass (its contribution is pending)

inline unsigned blocks_for(int64_t n, int t = kThreads) {
    return static_cast<unsigned>((n + t - 1) / t);
}

struct Inputs {
    const void *x, *y, *z;
    int pos_t;
    const void *vx, *vy, *vz;
    int vel_t;
    const void* m;
    int mass_t;
};

// member k -> FP32 position/mass (the potential's inputs), FP64 velocity and mass
template <int DT>
__global__ void k_gather(Inputs in, const int32_t* __restrict__ mem, int64_t nm,
                         float4* __restrict__ p4, double* __restrict__ vel,
                         double* __restrict__ m64, uint8_t* __restrict__ state) {
    const int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= nm) return;
    const int64_t i = mem[k];
    const double m = ldt<DT>(in.m, in.mass_t, i);
    p4[k] = make_float4(static_cast<float>(ldt<DT>(in.x, in.pos_t, i)),
                        static_cast<float>(ldt<DT>(in.y, in.pos_t, i)),
                        static_cast<float>(ldt<DT>(in.z, in.pos_t, i)), static_cast<float>(m));
    vel[3 * k + 0] = ldt<DT>(in.vx, in.vel_t, i);
    vel[3 * k + 1] = ldt<DT>(in.vy, in.vel_t, i);
    vel[3 * k + 2] = ldt<DT>(in.vz, in.vel_t, i);
    m64[k] = m;
    state[k] = kBound;
}

__global__ void k_member_halo(const int64_t* __restrict__ off, int H, int32_t* __restrict__ mh) {
    const int h = blockIdx.x;
    if (h >= H) return;
    for (int64_t k = off[h] + threadIdx.x; k < off[h + 1]; k += blockDim.x) mh[k] = h;
}

__global__ void k_init_halo(const int64_t* __restrict__ off, int H, int32_t* __restrict__ active,
                            int32_t* __restrict__ iters, int32_t* __restrict__ removed,
                            int32_t* __restrict__ stale, int32_t* __restrict__ need_full,
                            int32_t* __restrict__ refresh) {
    const int h = blockIdx.x * blockDim.x + threadIdx.x;
    if (h >= H) return;
    active[h] = off[h + 1] > off[h] ? 1 : 0;
    iters[h] = 0;
    removed[h] = 0;
    stale[h] = 0;
    need_full[h] = 0;
    refresh[h] = 0;
}

// A halo's pass is full (tests = sources = the live bound set, symmetric kernel) on
// the first pass, whenever the previous pass removed more than half as many members as
// it kept (an incremental pass costs kept x removed pair terms, a full one kept^2 / 2
// with the pair symmetry), and after a pass that left too many stale members (below);
// otherwise incremental (the removed members' terms are subtracted).
__device__ __forceinline__ bool pass_is_full(int pass_no, int run, int rrun, int need_full) {
    return pass_no == 0 || need_full != 0 || 2 * static_cast<int64_t>(rrun) > run;
}

// Largest relative error bound a member's S may keep.  A kept member whose accumulated
// bound (the full pass's, plus that of every incremental subtraction since) goes above it
// is stale.  Its decision is exact (it lies outside the decision band), so it only needs
// a fresh S: when its halo has few stale members (at most 1/kStaleDiv of the live set)
// the re-test kernel re-sums them in FP32 with every term Kahan-added (bound 8u =
// 4.8e-7); otherwise the
// halo's next pass is a full one (the symmetric kernel recomputes every S), and a halo
// that removed nobody takes that full pass only to refresh its S (not counted as an
// iteration: its decisions are exact and do not change).  Both full-pass kernels stay
// below the limit (halma_pot_error_bound at sub-tiles of 4 or 8, halma_pot_sym_error_bound).
constexpr double kPotRelMax = 1e-6;

// A wavefront re-sums one stale member, a latency-bound chain of about live/64 terms per
// lane, while a full pass adds a halo's block pairs to a launch that runs anyway: beyond
// 1/32 of the live set the full pass costs less (measured on clump and bridged-pair
// catalogues).
constexpr int kStaleDiv = 32;

__device__ __forceinline__ bool too_stale(int stale, int live) {
    return kStaleDiv * static_cast<int64_t>(stale) > live;
}

// band-list entries of stale members carry this bit (members of the decision band do not)
constexpr uint32_t kStaleBit = 0x80000000u;

__global__ void k_zero_counters(unsigned long long* __restrict__ c, int n) {
    if (static_cast<int>(threadIdx.x) < n) c[threadIdx.x] = 0;
}

// block-wide exclusive scan of one int per thread; returns the block total
__device__ int block_excl_scan(int v, int& excl, int* sh) {
    const int lane = __lane_id();
    const int w = threadIdx.x / kWave;
    int x = v;
    for (int s = 1; s < kWave; s <<= 1) {
        const int y = __shfl_up(x, s, kWave);
        if (lane >= s) x += y;
    }
    if (lane == kWave - 1) sh[w] = x;
    __syncthreads();
    int base = 0, tot = 0;
    for (int k = 0; k < kThreads / kWave; ++k) {
        if (k < w) base += sh[k];
        tot += sh[k];
    }
    excl = base + x - v;
    __syncthreads();
    return tot;
}

// Compaction of the bound sets, in fixed chunks of kChunk member positions (a halo of n
// members has ceil(n / kChunk) chunks, listed halo by halo): k_chunk_sums counts each
// chunk's kept and just-removed members and sums the kept members' mass and momentum in
// double-double; k_halo_sets turns the counts into per-chunk output offsets and merges
// the sums in chunk order into v_cm; k_chunk_write writes the compacted lists.  Large
// haloes thus spread over many wavefronts and blocks instead of one.
constexpr int kChunk = 1024;

__device__ __forceinline__ void chunk_range(const int64_t* off, int2 ch, int64_t& b,
                                            int64_t& e) {
    b = off[ch.x] + static_cast<int64_t>(ch.y) * kChunk;
    e = off[ch.x + 1] < b + kChunk ? off[ch.x + 1] : b + kChunk;
}

// One wavefront per chunk: lanes stride over the chunk's members, then the counts and the
// double-double sums are added across the lanes by a fixed butterfly (no barriers: a
// pass's chunks, many of them in haloes already converged, cost a wavefront each).
__global__ __launch_bounds__(kThreads) void k_chunk_sums(
    const int2* __restrict__ chunks, int n_chunks, const int64_t* __restrict__ off,
    const int32_t* __restrict__ active, const uint8_t* __restrict__ state,
    const double* __restrict__ vel, const double* __restrict__ m64,
    int32_t* __restrict__ ccnt, double* __restrict__ csum) {
    const int lane = __lane_id();
    const int64_t nw = static_cast<int64_t>(gridDim.x) * (kThreads / kWave);
    for (int64_t ci = (static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x) / kWave;
         ci < n_chunks; ci += nw) {
        const int2 ch = chunks[ci];
        if (!active[ch.x]) continue;  // wave-uniform
        int64_t b, e;
        chunk_range(off, ch, b, e);
        // every load of the chunk is issued before its first use (the states, then mass
        // and velocity of the bound members eight positions at a time): three memory
        // round trips per wavefront instead of two per position; the lane's terms are
        // added in the same order as one position at a time
        constexpr int kPer = kChunk / kWave;
        uint8_t sv[kPer];
#pragma unroll
        for (int j = 0; j < kPer; ++j) {
            const int64_t k = b + lane + static_cast<int64_t>(j) * kWave;
            sv[j] = k < e ? state[k] : kGone;
        }
        int packed = 0;  // kept count | just-removed count << 16 (at most kChunk each)
        double acc[4][2] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}};
#pragma unroll
        for (int j0 = 0; j0 < kPer; j0 += 8) {
            if (b + static_cast<int64_t>(j0) * kWave >= e) break;  // wave-uniform
            double mv[8][4];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int64_t k = b + lane + static_cast<int64_t>(j0 + u) * kWave;
                const bool ld = sv[j0 + u] == kBound;  // positions past e hold kGone
                mv[u][0] = ld ? m64[k] : 0.0;
                mv[u][1] = ld ? vel[3 * k + 0] : 0.0;
                mv[u][2] = ld ? vel[3 * k + 1] : 0.0;
                mv[u][3] = ld ? vel[3 * k + 2] : 0.0;
            }
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                if (sv[j0 + u] == kBound) {
                    packed += 1;
                    const double m = mv[u][0];
                    dd_add(acc[0][0], acc[0][1], m);
                    dd_add(acc[1][0], acc[1][1], m * mv[u][1]);
                    dd_add(acc[2][0], acc[2][1], m * mv[u][2]);
                    dd_add(acc[3][0], acc[3][1], m * mv[u][3]);
                } else if (sv[j0 + u] == kJustGone) {
                    packed += 1 << 16;
                }
            }
        }
        for (int m = kWave / 2; m >= 1; m >>= 1) {
            packed += __shfl_xor(packed, m, kWave);
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const double h2 = __shfl_xor(acc[q][0], m, kWave);
                const double l2 = __shfl_xor(acc[q][1], m, kWave);
                dd_add_dd(acc[q][0], acc[q][1], h2, l2);
            }
        }
        if (lane == 0) {
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                csum[8 * ci + 2 * q] = acc[q][0];
                csum[8 * ci + 2 * q + 1] = acc[q][1];
            }
            ccnt[2 * ci] = packed & 0xffff;
            ccnt[2 * ci + 1] = packed >> 16;
        }
    }
}

// One thread per halo, in the plan's halo order (h = order[pos]): chunk output offsets,
// bound and removed counts, v_cm, the pass mode (full recompute, or incremental when the
// last pass removed at most half as many members as it kept), the live sets of the pair kernel (tests =
// the bound set; sources = the bound set in a full pass, the members removed by the
// previous pass in an incremental one; none when inactive) and the halo's number of
// pair-kernel items (test tiles x source chunks).  Also clears the removed-members
// counter for the pass.
__device__ __forceinline__ double shfl_xor_f64(double v, int m) {
    return __shfl_xor(v, m, kWave);
}

// One wavefront per halo position: the lanes take the halo's chunks (lane l: chunks l,
// l + 64, ...), turn the counts into output offsets with a wave prefix scan, and merge the
// double-double sums by a fixed butterfly across the lanes.
__global__ __launch_bounds__(kThreads) void k_halo_sets(
    int H, const int32_t* __restrict__ order, const HaloDesc* __restrict__ halos, int tile,
    const int32_t* __restrict__ cfirst, const int32_t* __restrict__ active,
    const int32_t* __restrict__ pass_no, int32_t* __restrict__ ccnt,
    const double* __restrict__ csum, int32_t* __restrict__ n_bound, int32_t* __restrict__ n_rem,
    int32_t* __restrict__ mode, double* __restrict__ vcm, int32_t* __restrict__ tst_len,
    int32_t* __restrict__ src_len, int32_t* __restrict__ src_sel, int32_t* __restrict__ pchunk,
    int64_t* __restrict__ icnt, int64_t* __restrict__ icnt_sym, const SymDesc* __restrict__ sym,
    int32_t* __restrict__ stale, int32_t* __restrict__ need_full,
    unsigned long long* __restrict__ counters) {
    const int lane = __lane_id();
    const int pos = static_cast<int>((static_cast<int64_t>(blockIdx.x) * blockDim.x +
                                      threadIdx.x) / kWave);
    if (pos == 0 && lane == 0) {
        counters[0] = 0;
        counters[5] = 0;
    }
    if (pos >= H) return;  // wave-uniform
    const int h = order[pos];
    if (!active[h]) {
        if (lane == 0) {
            n_bound[h] = 0;
            n_rem[h] = 0;
            tst_len[h] = 0;
            src_len[h] = 0;
            src_sel[h] = 0;
            icnt[pos] = 0;
            icnt_sym[pos] = 0;
        }
        return;
    }
    const int c0 = cfirst[h], c1 = cfirst[h + 1];
    int run = 0, rrun = 0;  // totals of the chunks before the current 64
    double acc[4][2] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}};
    for (int cb = c0; cb < c1; cb += kWave) {
        const int c = cb + lane;
        const bool in = c < c1;
        const int kc = in ? ccnt[2 * c] : 0, gc = in ? ccnt[2 * c + 1] : 0;
        // inclusive prefix of the packed pair (kept in the low, removed in the high half
        // of a 64-bit word: both stay below 2^31)
        unsigned long long v = static_cast<unsigned long long>(kc) |
                               (static_cast<unsigned long long>(gc) << 32);
        for (int d = 1; d < kWave; d <<= 1) {
            const unsigned long long u = __shfl_up(v, d, kWave);
            if (lane >= d) v += u;
        }
        const int incl_k = static_cast<int>(v & 0xffffffffu), incl_g = static_cast<int>(v >> 32);
        if (in) {
            ccnt[2 * c] = run + incl_k - kc;  // counts become the chunk's output offsets
            ccnt[2 * c + 1] = rrun + incl_g - gc;
            for (int q = 0; q < 4; ++q)
                dd_add_dd(acc[q][0], acc[q][1], csum[8 * static_cast<int64_t>(c) + 2 * q],
                          csum[8 * static_cast<int64_t>(c) + 2 * q + 1]);
        }
        const unsigned long long tot = __shfl(v, kWave - 1, kWave);
        run += static_cast<int>(tot & 0xffffffffu);
        rrun += static_cast<int>(tot >> 32);
    }
    for (int m = kWave / 2; m >= 1; m >>= 1)
        for (int q = 0; q < 4; ++q) {
            const double oh = shfl_xor_f64(acc[q][0], m), ol = shfl_xor_f64(acc[q][1], m);
            // the lower lane of each pair adds the upper one's value: the same tree and
            // operand order in every lane
            if (lane & m)
                dd_add_dd(acc[q][0], acc[q][1], oh, ol);
            else {
                double th = oh, tl = ol;
                dd_add_dd(th, tl, acc[q][0], acc[q][1]);
                acc[q][0] = th;
                acc[q][1] = tl;
            }
        }
    if (lane != 0) return;
    const double M = acc[0][0] + acc[0][1];
    for (int d = 0; d < 3; ++d)
        vcm[3 * h + d] = M > 0.0 ? (acc[1 + d][0] + acc[1 + d][1]) / M : 0.0;
    const int full = pass_is_full(*pass_no, run, rrun, need_full[h]) ? 1 : 0;
    need_full[h] = 0;
    stale[h] = 0;
    n_bound[h] = run;
    n_rem[h] = rrun;
    mode[h] = full;
    const int nt = run;
    const int ns = full ? run : rrun;
    tst_len[h] = nt;
    src_len[h] = ns;
    src_sel[h] = full ? 0 : 1;
    // full passes (tests = sources = the bound set) go to the symmetric kernel: block
    // pairs A <= B of its live size; incremental ones, and full ones of haloes planned
    // without slots (above kSymMaxMembers), to the chunked kernel
    const bool use_sym = full && sym[h].nb > 0;
    // source chunk of this pass: the plan's for a full pass; for an incremental one (few
    // sources: the members just removed) as many of the planned chunk slots as chunks of
    // at least kIncChunk sources fill, so that a large halo's removals still spread over
    // several items (ns never exceeds the planned source count, so the chunks fit)
    const HaloDesc hd = halos[h];
    int ch = hd.chunk;
    if (!full && ns > 0) {
        const int slots = (hd.src_len + hd.chunk - 1) / hd.chunk;
        int c = (ns + slots - 1) / slots;
        c = c < kIncChunk ? kIncChunk : c;
        c = (c + 7) / 8 * 8;
        ch = c < ch ? c : ch;
    }
    if (lane == 0) pchunk[h] = ch;
    icnt[pos] = (!use_sym && nt > 0 && ns > 0)
                    ? static_cast<int64_t>((nt + tile - 1) / tile) * ((ns + ch - 1) / ch)
                    : 0;
    const int64_t nbs = use_sym ? sym_blocks(nt) : 0;
    icnt_sym[pos] = nbs * (nbs + 1) / 2;
}

// The compaction by blocks bid, bid + nblk, ... of a launch (its own kernel, or the
// second part of k_items_write)
__device__ __forceinline__ void chunk_write_blocks(
    const int2* __restrict__ chunks, int n_chunks, const int64_t* __restrict__ off,
    const int32_t* __restrict__ active, const int32_t* __restrict__ ccnt,
    uint8_t* __restrict__ state, const float4* __restrict__ p4, float4* __restrict__ src4,
    int32_t* __restrict__ cidx, float4* __restrict__ rem4, int bid, int nblk) {
    __shared__ int shs[kThreads / kWave];
    for (int ci = bid; ci < n_chunks; ci += nblk) {
        const int2 ch = chunks[ci];
        if (!active[ch.x]) continue;  // block-uniform
        int64_t b, e;
        chunk_range(off, ch, b, e);
        const int64_t kbase = off[ch.x] + ccnt[2 * ci];
        const int64_t gbase = off[ch.x] + ccnt[2 * ci + 1];
        // the chunk's states and the positions of its live and just-removed members are
        // loaded before the first scan (one memory round trip instead of two per round)
        constexpr int kPer = kChunk / kThreads;
        uint8_t sv[kPer];
        float4 pv[kPer];
#pragma unroll
        for (int j = 0; j < kPer; ++j) {
            const int64_t k = b + static_cast<int64_t>(j) * kThreads + threadIdx.x;
            sv[j] = k < e ? state[k] : kGone;
        }
#pragma unroll
        for (int j = 0; j < kPer; ++j) {
            const int64_t k = b + static_cast<int64_t>(j) * kThreads + threadIdx.x;
            pv[j] = sv[j] != kGone ? p4[k] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        }
        int run = 0, rrun = 0;
#pragma unroll
        for (int j = 0; j < kPer; ++j) {
            if (b + static_cast<int64_t>(j) * kThreads >= e) break;  // block-uniform
            const int64_t k = b + static_cast<int64_t>(j) * kThreads + threadIdx.x;
            const uint8_t st = sv[j];
            const int v = st == kBound ? 1 : (st == kJustGone ? (1 << 16) : 0);
            int excl;
            const int tot = block_excl_scan(v, excl, shs);
            if (st == kBound) {
                const int64_t pos = kbase + run + (excl & 0xffff);
                src4[pos] = pv[j];
                cidx[pos] = static_cast<int32_t>(k);
            } else if (st == kJustGone) {
                rem4[gbase + rrun + (excl >> 16)] = pv[j];
                state[k] = kGone;
            }
            run += tot & 0xffff;
            rrun += tot >> 16;
        }
    }
}

__global__ __launch_bounds__(kThreads) void k_chunk_write(
    const int2* __restrict__ chunks, int n_chunks, const int64_t* __restrict__ off,
    const int32_t* __restrict__ active, const int32_t* __restrict__ ccnt,
    uint8_t* __restrict__ state, const float4* __restrict__ p4, float4* __restrict__ src4,
    int32_t* __restrict__ cidx, float4* __restrict__ rem4) {
    chunk_write_blocks(chunks, n_chunks, off, active, ccnt, state, p4, src4, cidx, rem4,
                       blockIdx.x, gridDim.x);
}

struct EnergyArgs {
    const HaloDesc* halos;
    const int32_t* mhalo;
    const int32_t* n_bound;
    const int32_t* n_src;
    const int32_t* pchunk;
    const int32_t* mode;
    const int32_t* cidx;
    const float2* part;
    const SymDesc* sym;      // full passes: symmetric kernel's slots
    const float2* sym_part;
    const double* vel;
    const double* vcm;
    double cg;         // factor * G
    double bound_rel;  // FP32 relative error bound B of one pair sum (incremental passes)
    double bound_sym;  // the same for the symmetric kernel (full passes)
    double bound_resum;  // the same for the FP32 re-sum of a stale member
    uint8_t* state;
    double* sacc;      // running S per member position
    double* eacc;      // absolute error bound of sacc
    int32_t* removed;
    int32_t* band;       // members to re-test: small live sets from the front, large from
    int64_t band_cap;    // the back (band_cap entries)
    int32_t* stale;      // per halo: stale members this pass
    int32_t* need_full;  // per halo: the next pass is a full one
    int32_t* refresh;    // per halo: this pass only refreshes S (not an iteration)
    unsigned long long* counters;  // [0] removed total, [1] band count, [2] band total,
                                   // [3] pairs, [4] symmetric-kernel pairs, [5] haloes
                                   // still active after the pass, [6] members re-summed
                                   // for their error bound alone, [8] the band
                                   // count of large live sets
    double* s_out;     // optional: S per member position (FP64)
    int64_t nm;
    // the end of a pass (k_finish)
    int H;
    int32_t* active;
    int32_t* iters;
    int32_t* pass_no;
};

__device__ __forceinline__ double kinetic(const double* vel, int64_t k, const double* vc) {
    const double dvx = vel[3 * k + 0] - vc[0];
    const double dvy = vel[3 * k + 1] - vc[1];
    const double dvz = vel[3 * k + 2] - vc[2];
    return 0.5 * ((dvx * dvx + dvy * dvy) + dvz * dvz);
}

__global__ __launch_bounds__(kThreads) void k_energy(EnergyArgs a) {
    const int64_t g = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (g >= a.nm) return;
    const int h = a.mhalo[g];
    const HaloDesc d = a.halos[h];
    const int64_t p = g - d.tst_off;
    const int n = a.n_bound[h];
    if (p >= n) return;
    const bool full = a.mode[h] != 0;
    const bool use_sym = full && a.sym[h].nb > 0;
    // symmetric full pass: the kernel's nb + 1 slots (nb of the live set); chunked pass
    // (incremental, or full for haloes planned without slots): the chunk partials; both
    // summed in order, loads issued eight at a time
    int nch;
    int64_t stride;
    const float2* pp;
    if (use_sym) {
        const SymDesc sd = a.sym[h];
        nch = sym_blocks(n) + 1;
        stride = sd.n;
        pp = a.sym_part + sd.part_off + p;
    } else {
        const int ns = a.n_src[h];
        const int ch = a.pchunk[h];
        nch = (ns + ch - 1) / ch;
        stride = d.stride;
        pp = a.part + d.part_off + p;
    }
    // every batch (the last one too) has its loads in flight together: a member pays
    // ceil(nch / 8) memory round trips, never one per slot of a ragged tail
    double sum = 0.0;
    for (int c = 0; c < nch; c += 8) {
        float2 q[8];
#pragma unroll
        for (int u = 0; u < 8; ++u)
            q[u] = c + u < nch ? pp[static_cast<int64_t>(c + u) * stride] : make_float2(0.0f, 0.0f);
#pragma unroll
        for (int u = 0; u < 8; ++u)
            if (c + u < nch) sum += static_cast<double>(q[u].x) + static_cast<double>(q[u].y);
    }
    const int64_t k = a.cidx[g];
    double S, err;
    if (full) {
        S = sum;
        err = (use_sym ? a.bound_sym : a.bound_rel) * sum;
    } else {
        S = a.sacc[k] - sum;
        err = a.eacc[k] + a.bound_rel * sum;
    }
    a.sacc[k] = S;
    a.eacc[k] = err;
    const double kin = kinetic(a.vel, k, a.vcm + 3 * h);
    const double pot = a.cg * S;
    const double E = kin - pot;
    // FP32 bound, FP64 rounding of the updates and of E, and the oracle's own FP64
    // rounding (far below 1e-10 relative for any halo that fits in memory)
    const double band = a.cg * err * 1.01 + 8.0 * 0x1p-53 * (kin + fabs(pot)) +
                        1e-10 * fabs(pot);
    if (a.s_out) a.s_out[k] = S;
    const bool in_band = fabs(E) <= band;
    const bool gone = !in_band && E > 0.0;
    // a kept member whose error bound went above kPotRelMax: its S is re-summed (its
    // decision, already exact, stays the same)
    const bool stale = !gone && !in_band && err > kPotRelMax * S;
    const int lane = __lane_id();
    // stale members per halo, one atomic per halo present in the wave (members of a halo
    // are contiguous, so a wave spans few haloes); the count returned tells whether the
    // halo already has too many for a re-sum (it then takes a full pass next and
    // k_retest would skip them): such members are not listed
    bool list_stale = false;
    unsigned long long pending = __ballot(stale);
    while (pending) {
        const int leader = __ffsll(static_cast<long long>(pending)) - 1;
        const int hl = __shfl(h, leader, kWave);
        const unsigned long long same = __ballot(stale && h == hl);
        int before = 0;
        if (lane == leader) before = atomicAdd(a.stale + hl, __popcll(same));
        before = __shfl(before, leader, kWave);
        if (stale && h == hl) list_stale = !too_stale(before + __popcll(same), n);
        pending &= ~same;
    }
    // members to re-test: of small live sets from the front of the band list, of large
    // ones (a block each in k_retest) from its back; one atomic per wavefront and list
    const bool push = in_band || list_stale;
    const bool big = n >= kRetestBlock;
    const int32_t val = static_cast<int32_t>(static_cast<uint32_t>(g) | (stale ? kStaleBit : 0u));
    const unsigned long long below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    const unsigned long long m_small = __ballot(push && !big), m_big = __ballot(push && big);
    if (m_small) {
        const int leader = __ffsll(static_cast<long long>(m_small)) - 1;
        unsigned long long base = 0;
        if (lane == leader) base = atomicAdd(a.counters + 1, static_cast<unsigned long long>(__popcll(m_small)));
        base = __shfl(base, leader, kWave);
        if (push && !big) a.band[base + __popcll(m_small & below)] = val;
    }
    if (m_big) {
        const int leader = __ffsll(static_cast<long long>(m_big)) - 1;
        unsigned long long base = 0;
        if (lane == leader) base = atomicAdd(a.counters + 8, static_cast<unsigned long long>(__popcll(m_big)));
        base = __shfl(base, leader, kWave);
        if (push && big) a.band[a.band_cap - 1 - static_cast<int64_t>(base + __popcll(m_big & below))] = val;
    }
    if (gone) a.state[k] = kJustGone;
    // removals per halo, one atomic per halo present in the wave
    pending = __ballot(gone);
    while (pending) {
        const int leader = __ffsll(static_cast<long long>(pending)) - 1;
        const int hl = __shfl(h, leader, kWave);
        const unsigned long long same = __ballot(gone && h == hl);
        if (lane == leader) atomicAdd(a.removed + hl, __popcll(same));
        pending &= ~same;
    }
}

__device__ __forceinline__ unsigned long long wave_sum_u64(unsigned long long v) {
    for (int o = kWave / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
    return v;
}

// End of a pass for the halo h (every thread of the block calls it): per-halo iteration
// counts, the halo stays active while it removed someone (or, once, to refresh the S of
// too many stale members), and the counters (wave-aggregated atomics): [0] removed
// total, [3] pairs of the reference's loop this pass (n_bound x n_src), [4] those of the
// symmetric kernel's passes, [5] haloes still active.
__device__ void finish_haloes(const EnergyArgs& a, int h) {
    unsigned long long rem = 0, pairs = 0, full_pairs = 0, live = 0;
    if (h < a.H && a.active[h]) {
        const int nb = a.n_bound[h];
        const int r = a.removed[h];
        const bool refreshing = a.refresh[h] != 0;
        if (nb > 0) {
            if (!(refreshing && r == 0)) a.iters[h] += 1;
            pairs = static_cast<unsigned long long>(nb) *
                    static_cast<unsigned long long>(a.n_src[h]);
            if (a.mode[h] && a.sym[h].nb > 0) full_pairs = pairs;  // symmetric kernel's
        }
        const bool stay = r > 0 && nb > 0;
        const bool redo = nb > 0 && too_stale(a.stale[h], nb);
        a.need_full[h] = redo ? 1 : 0;
        const bool refresh = !stay && redo && !refreshing;
        a.refresh[h] = refresh ? 1 : 0;
        if (!stay && !refresh) a.active[h] = 0;
        live = (stay || refresh) ? 1 : 0;
        rem = static_cast<unsigned long long>(r);
        a.removed[h] = 0;
    }
    rem = wave_sum_u64(rem);
    pairs = wave_sum_u64(pairs);
    full_pairs = wave_sum_u64(full_pairs);
    live = wave_sum_u64(live);
    if ((threadIdx.x & (kWave - 1)) == 0) {
        if (rem) atomicAdd(a.counters + 0, rem);
        if (pairs) atomicAdd(a.counters + 3, pairs);
        if (full_pairs) atomicAdd(a.counters + 4, full_pairs);
        if (live) atomicAdd(a.counters + 5, live);
    }
}

// A stale member (kept, outside the decision band: its decision is exact) only needs a
// fresh S within kPotRelMax: one FP32 sum over its halo's live set with the chunked
// kernel's term and every term Kahan-added (relative bound pot_error_bound(1), 8u), the
// threads of a wavefront or block in stride over the sources (four loads in flight),
// merged in FP64.  About five times cheaper per term than the FP64 ordered sum.
__device__ __forceinline__ double resum_f32_part(const float4* __restrict__ src, int n,
                                                 float4 t, int i0, int step) {
    float s = 0.0f, c = 0.0f;
    int j = i0;
    for (; j + 3 * step < n; j += 4 * step) {
        float4 q[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) q[u] = src[j + u * step];
#pragma unroll
        for (int u = 0; u < 4; ++u) kahan_add(s, c, pair_acc(q[u], t.x, t.y, t.z, 0.0f));
    }
    for (; j < n; j += step) kahan_add(s, c, pair_acc(src[j], t.x, t.y, t.z, 0.0f));
    return static_cast<double>(s) + static_cast<double>(c);
}

__device__ __forceinline__ double wave_resum_f32(const float4* __restrict__ src, int n, float4 t) {
    double r = resum_f32_part(src, n, t, __lane_id(), kWave);
    for (int m = kWave / 2; m >= 1; m >>= 1) r += __shfl_xor(r, m, kWave);
    return r;
}

__device__ double block_resum_f32(const float4* __restrict__ src, int n, float4 t, double* sh) {
    double r = resum_f32_part(src, n, t, threadIdx.x, blockDim.x);
    for (int m = kWave / 2; m >= 1; m >>= 1) r += __shfl_xor(r, m, kWave);
    if (__lane_id() == 0) sh[threadIdx.x / kWave] = r;
    __syncthreads();
    double tot = 0.0;
    for (int w = 0; w < static_cast<int>(blockDim.x) / kWave; ++w) tot += sh[w];
    __syncthreads();
    return tot;
}

// The FP64 sum in the oracle's order decides every listed member (a block each for live
// sets of kRetestBlock or more, a wavefront each for smaller ones); stale members get an
// FP32 re-sum.  k_finish then ends the pass.
__global__ __launch_bounds__(kOrdWidth) void k_retest(EnergyArgs a,
                                                      const float4* __restrict__ src4) {
    static_assert(kOrdWidth == kThreads, "the FP64 block sums take kThreads threads");
    __shared__ double sh[kOrdWidth];
    const int64_t nb = static_cast<int64_t>(a.counters[1]);
    // members of the decision band: the sum in the oracle's order (same bits either way),
    // a block per member of a large live set (a shorter chain per thread), a wavefront per
    // member of a small one (more members in flight); stale members: an FP32 re-sum.  The
    // re-test and re-sum counts are kept by the thread that writes and added to the
    // global counters once per thread at the end (one address, many members per pass).
    unsigned long long n_retest = 0, n_resum = 0;
    auto fresh = [&](int64_t g, double S) {
        const int64_t k = a.cidx[g];
        a.sacc[k] = S;
        a.eacc[k] = a.bound_resum * S;
        if (a.s_out) a.s_out[k] = S;
        ++n_retest;
        ++n_resum;
    };
    auto decide = [&](int64_t g, int h, double S) {
        const int64_t k = a.cidx[g];
        const double E = kinetic(a.vel, k, a.vcm + 3 * h) - a.cg * S;
        a.sacc[k] = S;
        a.eacc[k] = 1e-12 * S;
        if (a.s_out) a.s_out[k] = S;
        ++n_retest;
        return E;
    };
    const int64_t nbig = static_cast<int64_t>(a.counters[8]);
    for (int64_t w = blockIdx.x; w < nbig; w += gridDim.x) {  // large live sets: a block each
        const uint32_t entry = static_cast<uint32_t>(a.band[a.band_cap - 1 - w]);
        const int64_t g = entry & ~kStaleBit;
        const int h = a.mhalo[g];
        const int n = a.n_bound[h];
        // stale members of a halo that takes a full pass next are left to that pass
        if ((entry & kStaleBit) && too_stale(a.stale[h], n)) continue;  // uniform
        const float4 t = src4[g];
        if (entry & kStaleBit) {  // a fresh S only
            const double S = block_resum_f32(src4 + a.halos[h].src_off, n, t, sh);
            if (threadIdx.x == 0) fresh(g, S);
            continue;
        }
        SrcAoS src{src4 + a.halos[h].src_off};
        const double S = block_pot_f64_ordered(src, n, t.x, t.y, t.z, sh);
        if (threadIdx.x == 0 && decide(g, h, S) > 0.0) {
            a.state[a.cidx[g]] = kJustGone;
            atomicAdd(a.removed + h, 1);
        }
    }
    const int waves = kThreads / kWave;
    for (int64_t w = static_cast<int64_t>(blockIdx.x) * waves + threadIdx.x / kWave; w < nb;
         w += static_cast<int64_t>(gridDim.x) * waves) {
        const uint32_t entry = static_cast<uint32_t>(a.band[w]);
        const int64_t g = entry & ~kStaleBit;
        const int h = a.mhalo[g];
        const int n = a.n_bound[h];
        if ((entry & kStaleBit) && too_stale(a.stale[h], n)) continue;
        const float4 t = src4[g];
        if (entry & kStaleBit) {  // a fresh S only
            const double S = wave_resum_f32(src4 + a.halos[h].src_off, n, t);
            if (__lane_id() == 0) fresh(g, S);
            continue;
        }
        SrcAoS src{src4 + a.halos[h].src_off};
        const double S = wave_pot_f64_ordered(src, n, t.x, t.y, t.z);
        if (__lane_id() == 0 && decide(g, h, S) > 0.0) {
            a.state[a.cidx[g]] = kJustGone;
            atomicAdd(a.removed + h, 1);
        }
    }
    if (n_retest) atomicAdd(a.counters + 2, n_retest);
    if (n_resum) atomicAdd(a.counters + 6, n_resum);
}

// The end of a pass, after every re-test: per-halo counts and modes (finish_haloes), one
// thread per halo; the band lists cleared for the next pass.
__global__ __launch_bounds__(kThreads) void k_finish(EnergyArgs a) {
    const int h = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    finish_haloes(a, h);
    if (h == 0) {
        a.counters[1] = 0;  // the band lists are rebuilt by the next pass
        a.counters[8] = 0;
        *a.pass_no += 1;
    }
}

// The pass's item list: the items of the haloes with live work, in the plan's halo
// order, test tiles fastest ({halo, tile, chunk, 0}, as in the plan).  ioff is the
// exclusive scan of the per-position counts; the list never exceeds the planned one
// (live sets only shrink, no pass has more chunks than the plan), and is cut at cap all
// the same.
__global__ void k_item_fill(int H, const int32_t* __restrict__ order,
                            const int64_t* __restrict__ ioff, const int64_t* __restrict__ itotal,
                            const int32_t* __restrict__ tst_len, int tile, int64_t cap,
                            int4* __restrict__ items) {
    const int64_t total = itotal[0] < cap ? itotal[0] : cap;
    for (int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; k < total;
         k += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        int lo = 0, hi = H - 1;  // last position with ioff <= k (its count is > 0)
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (ioff[mid] <= k)
                lo = mid;
            else
                hi = mid - 1;
        }
        const int h = order[lo];
        const int64_t local = k - ioff[lo];
        const int tiles = (tst_len[h] + tile - 1) / tile;
        items[k] = make_int4(h, static_cast<int>(local % tiles), static_cast<int>(local / tiles),
                             0);
    }
}

// The pass's items of the symmetric kernel: block pairs (A, B), A <= B, of every
// full-pass halo, in the plan's halo order, B-major (local index B(B+1)/2 + A).
__global__ void k_sym_item_fill(int H, const int32_t* __restrict__ order,
                                const int64_t* __restrict__ ioff,
                                const int64_t* __restrict__ itotal, int64_t cap,
                                int4* __restrict__ items, int* __restrict__ head) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *head = 0;
    const int64_t total = itotal[0] < cap ? itotal[0] : cap;
    for (int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; k < total;
         k += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        int lo = 0, hi = H - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (ioff[mid] <= k)
                lo = mid;
            else
                hi = mid - 1;
        }
        const int64_t local = k - ioff[lo];
        int64_t B = static_cast<int64_t>((sqrt(8.0 * static_cast<double>(local) + 1.0) - 1.0) / 2.0);
        while (B * (B + 1) / 2 > local) --B;
        while ((B + 1) * (B + 2) / 2 <= local) ++B;
        const int64_t A = local - B * (B + 1) / 2;
        items[k] = make_int4(order[lo], static_cast<int>(A), static_cast<int>(B), 0);
    }
}

// Both item lists of a pass in one launch, for at most kFuseH haloes: every block scans
// the per-position counts of both lists in LDS (32-bit: the plans keep both totals below
// 2^31) and fills its share of each list as k_item_fill and k_sym_item_fill do; block 0
// publishes the totals and clears both queue heads.
constexpr int kFuseH = 4096;
constexpr int kFusePer = kFuseH / kThreads;

__device__ __forceinline__ int last_le(const int* off, int H, int64_t k) {
    int lo = 0, hi = H - 1;  // last position with off <= k (its count is > 0)
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (off[mid] <= k)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

__device__ __forceinline__ void items_fused_blocks(
    int H, const int32_t* __restrict__ order, const int64_t* __restrict__ icnt,
    const int64_t* __restrict__ icnt_sym, const int32_t* __restrict__ tst_len, int tile,
    int64_t cap, int64_t sym_cap, int4* __restrict__ items, int4* __restrict__ sym_items,
    int64_t* __restrict__ itotal, int64_t* __restrict__ sym_itotal, int* __restrict__ sym_head,
    int bid, int nblk) {
    __shared__ int off_a[kFuseH];
    __shared__ int off_b[kFuseH];
    __shared__ int shs[kThreads / kWave];
    const int base = threadIdx.x * kFusePer;
    int va[kFusePer], vb[kFusePer];
    int sa = 0, sb = 0;
#pragma unroll
    for (int k = 0; k < kFusePer; ++k) {
        const int p = base + k;
        va[k] = p < H ? static_cast<int>(icnt[p]) : 0;
        vb[k] = p < H ? static_cast<int>(icnt_sym[p]) : 0;
        sa += va[k];
        sb += vb[k];
    }
    int ea, eb;
    const int ta = block_excl_scan(sa, ea, shs);
    const int tb = block_excl_scan(sb, eb, shs);
#pragma unroll
    for (int k = 0; k < kFusePer; ++k) {
        off_a[base + k] = ea;
        off_b[base + k] = eb;
        ea += va[k];
        eb += vb[k];
    }
    __syncthreads();
    if (bid == 0 && threadIdx.x == 0) {
        *itotal = ta;
        *sym_itotal = tb;
        *sym_head = 0;
    }
    const int64_t stride = static_cast<int64_t>(nblk) * blockDim.x;
    const int64_t t0 = static_cast<int64_t>(bid) * blockDim.x + threadIdx.x;
    const int64_t na = ta < cap ? ta : cap;
    for (int64_t k = t0; k < na; k += stride) {
        const int pos = last_le(off_a, H, k);
        const int h = order[pos];
        const int64_t local = k - off_a[pos];
        const int tiles = (tst_len[h] + tile - 1) / tile;
        items[k] = make_int4(h, static_cast<int>(local % tiles), static_cast<int>(local / tiles),
                             0);
    }
    const int64_t nb = tb < sym_cap ? tb : sym_cap;
    for (int64_t k = t0; k < nb; k += stride) {
        const int pos = last_le(off_b, H, k);
        const int64_t local = k - off_b[pos];
        int64_t B = static_cast<int64_t>((sqrt(8.0 * static_cast<double>(local) + 1.0) - 1.0) / 2.0);
        while (B * (B + 1) / 2 > local) --B;
        while ((B + 1) * (B + 2) / 2 <= local) ++B;
        const int64_t A = local - B * (B + 1) / 2;
        sym_items[k] = make_int4(order[pos], static_cast<int>(A), static_cast<int>(B), 0);
    }
}

// The pass's item lists (blocks 0 .. items_blocks - 1) and the bound-set compaction (the
// other blocks) in one launch: the two only need k_halo_sets' offsets.
__global__ __launch_bounds__(kThreads) void k_items_write(
    int H, const int32_t* __restrict__ order, const int64_t* __restrict__ icnt,
    const int64_t* __restrict__ icnt_sym, const int32_t* __restrict__ tst_len, int tile,
    int64_t cap, int64_t sym_cap, int4* __restrict__ items, int4* __restrict__ sym_items,
    int64_t* __restrict__ itotal, int64_t* __restrict__ sym_itotal, int* __restrict__ sym_head,
    int items_blocks, const int2* __restrict__ chunks, int n_chunks,
    const int64_t* __restrict__ off, const int32_t* __restrict__ active,
    const int32_t* __restrict__ ccnt, uint8_t* __restrict__ state, const float4* __restrict__ p4,
    float4* __restrict__ src4, int32_t* __restrict__ cidx, float4* __restrict__ rem4) {
    if (static_cast<int>(blockIdx.x) < items_blocks)
        items_fused_blocks(H, order, icnt, icnt_sym, tst_len, tile, cap, sym_cap, items,
                           sym_items, itotal, sym_itotal, sym_head, blockIdx.x, items_blocks);
    else
        chunk_write_blocks(chunks, n_chunks, off, active, ccnt, state, p4, src4, cidx, rem4,
                           blockIdx.x - items_blocks, gridDim.x - items_blocks);
}

__global__ void k_final_mask(int64_t nm, const uint8_t* __restrict__ state,
                             uint8_t* __restrict__ bound) {
    const int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k < nm) bound[k] = state[k] == kBound ? 1 : 0;
}

struct UnbWs {
    float4* p4;
    double* vel;
    double* m64;
    uint8_t* state;
    float4* src4;
    float4* rem4;
    int32_t* cidx;
    int32_t* mhalo;
    double* sacc;
    double* eacc;
    int32_t* n_bound;
    int32_t* n_rem;
    int32_t* mode;
    int32_t* tst_len;
    int32_t* src_len;
    int32_t* src_sel;
    int32_t* pchunk;  // per-halo source chunk of the current pass
    int32_t* active;
    int32_t* iters;
    int32_t* removed;
    int32_t* stale;
    int32_t* need_full;
    int32_t* refresh;
    int32_t* pass_no;
    double* vcm;
    int32_t* band;
    unsigned long long* counters;
    HaloDesc* halos;
    int4* items;
    int32_t* order;
    int64_t* icnt;
    int64_t* ioff;
    int64_t* itotal;
    void* iscan_ws;
    float2* part;
    // the symmetric kernel of the full passes
    SymDesc* sym;     // per halo (in the tables region)
    int4* sym_items;
    int64_t* sym_icnt;
    int64_t* sym_ioff;
    int64_t* sym_itotal;
    void* sym_scan_ws;
    float2* sym_part;
    int* sym_head;
    int64_t sym_cap;
    int64_t* off;
    int2* chunks;     // compaction chunks {halo, chunk number}, halo by halo
    int32_t* cfirst;  // first chunk of each halo (H + 1)
    int32_t* ccnt;    // per chunk: kept / removed counts, then their output offsets
    double* csum;     // per chunk: double-double mass and momentum sums (8 doubles)
    int64_t n_chunks;
    // host tables (off, halos, order, chunks, cfirst) sit in one region, uploaded at once
    char* tables;
    size_t t_off, t_halos, t_order, t_chunks, t_cfirst, t_sym, t_bytes;
    size_t bytes;
};

int64_t count_chunks(int64_t H, const int64_t* h_off) {
    int64_t n = 0;
    for (int64_t h = 0; h < H; ++h) n += (h_off[h + 1] - h_off[h] + kChunk - 1) / kChunk;
    return n;
}

UnbWs carve_unb(void* base, int64_t H, int64_t nm, const PotPlan& plan, const SymPlan& splan,
                int64_t n_chunks) {
    Carver c(base);
    const int64_t n1 = nm > 0 ? nm : 1;
    const int64_t h1 = H > 0 ? H : 1;
    const int64_t c1 = n_chunks > 0 ? n_chunks : 1;
    UnbWs w;
    Carver t(nullptr);
    w.t_off = t.used;
    t.take<int64_t>(h1 + 1);
    w.t_halos = t.used;
    t.take<HaloDesc>(h1);
    w.t_order = t.used;
    t.take<int32_t>(h1);
    w.t_chunks = t.used;
    t.take<int2>(c1);
    w.t_cfirst = t.used;
    t.take<int32_t>(h1 + 1);
    w.t_sym = t.used;
    t.take<SymDesc>(h1);
    w.t_bytes = t.used;
    w.tables = c.take<char>(w.t_bytes);
    w.off = reinterpret_cast<int64_t*>(w.tables + w.t_off);
    w.halos = reinterpret_cast<HaloDesc*>(w.tables + w.t_halos);
    w.order = reinterpret_cast<int32_t*>(w.tables + w.t_order);
    w.chunks = reinterpret_cast<int2*>(w.tables + w.t_chunks);
    w.cfirst = reinterpret_cast<int32_t*>(w.tables + w.t_cfirst);
    w.sym = reinterpret_cast<SymDesc*>(w.tables + w.t_sym);
    w.n_chunks = n_chunks;
    w.ccnt = c.take<int32_t>(2 * c1);
    w.csum = c.take<double>(8 * c1);
    w.p4 = c.take<float4>(n1);
    w.vel = c.take<double>(3 * n1);
    w.m64 = c.take<double>(n1);
    w.state = c.take<uint8_t>(n1);
    w.src4 = c.take<float4>(n1);
    w.rem4 = c.take<float4>(n1);
    w.cidx = c.take<int32_t>(n1);
    w.mhalo = c.take<int32_t>(n1);
    w.sacc = c.take<double>(n1);
    w.eacc = c.take<double>(n1);
    w.n_bound = c.take<int32_t>(h1);
    w.n_rem = c.take<int32_t>(h1);
    w.mode = c.take<int32_t>(h1);
    w.tst_len = c.take<int32_t>(h1);
    w.src_len = c.take<int32_t>(h1);
    w.src_sel = c.take<int32_t>(h1);
    w.pchunk = c.take<int32_t>(h1);
    w.active = c.take<int32_t>(h1);
    w.iters = c.take<int32_t>(h1);
    w.removed = c.take<int32_t>(h1);
    w.stale = c.take<int32_t>(h1);
    w.need_full = c.take<int32_t>(h1);
    w.refresh = c.take<int32_t>(h1);
    w.pass_no = c.take<int32_t>(64);
    w.vcm = c.take<double>(3 * h1);
    w.band = c.take<int32_t>(n1);
    w.counters = c.take<unsigned long long>(16);
    w.items = c.take<int4>(std::max<int64_t>(plan.n_items, 1));
    w.icnt = c.take<int64_t>(h1);
    w.ioff = c.take<int64_t>(h1);
    w.itotal = c.take<int64_t>(4);
    w.iscan_ws = c.take<char>(scan_workspace_bytes(h1));
    w.part = c.take<float2>(std::max<int64_t>(plan.n_part, 1));
    w.sym_cap = std::max<int64_t>(splan.n_items, 1);
    w.sym_items = c.take<int4>(w.sym_cap);
    w.sym_icnt = c.take<int64_t>(h1);
    w.sym_ioff = c.take<int64_t>(h1);
    w.sym_itotal = c.take<int64_t>(4);
    w.sym_scan_ws = c.take<char>(scan_workspace_bytes(h1));
    w.sym_part = c.take<float2>(std::max<int64_t>(splan.n_part, 1));
    w.sym_head = c.take<int>(64);
    w.bytes = c.used;
    return w;
}

// library-owned capture stream per device (default priority)
hipStream_t capture_stream() {
    static std::mutex mu;
    static std::vector<std::pair<int, hipStream_t>> streams;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return nullptr;
    std::lock_guard<std::mutex> lock(mu);
    for (auto& p : streams)
        if (p.first == dev) return p.second;
    hipStream_t s = nullptr;
    if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) return nullptr;
    streams.emplace_back(dev, s);
    return s;
}

struct PassLaunch {
    int64_t H, nm;
    int grid_chunks, grid_chunk_waves, grid_retest, grid_items;
    const PotPlan* plan;
    const SymPlan* splan;
    UnbWs w;
    uint8_t* bound;
    EnergyArgs ea;
};

// A pass in four parts (each captured into its own graph): prepare (compaction, live
// lengths, item lists), the symmetric kernel (full passes), the chunked kernel
// (incremental passes), and the energy decisions.  Timing events, when requested, are
// recorded on the stream between the parts, so that they bracket the pair kernels alone.
int launch_pre(const PassLaunch& L, hipStream_t st) {
    const UnbWs& w = L.w;
    const int nc = static_cast<int>(w.n_chunks);
    const int tile = L.plan->tests_per_lane * kWave;
    if (nc > 0) {
        k_chunk_sums<<<L.grid_chunk_waves, kThreads, 0, st>>>(w.chunks, nc, w.off, w.active,
                                                              w.state, w.vel, w.m64, w.ccnt,
                                                              w.csum);
        HALMA_LAUNCH_CHECK();
    }
    k_halo_sets<<<static_cast<unsigned>((L.H + kThreads / kWave - 1) / (kThreads / kWave)),
                  kThreads, 0, st>>>(
        static_cast<int>(L.H), w.order, w.halos, tile, w.cfirst, w.active, w.pass_no, w.ccnt,
        w.csum, w.n_bound, w.n_rem, w.mode, w.vcm, w.tst_len, w.src_len, w.src_sel, w.pchunk,
        w.icnt, w.sym_icnt, w.sym, w.stale, w.need_full, w.counters);
    HALMA_LAUNCH_CHECK();
    if (L.H <= kFuseH) {
        const int wblocks = nc > 0 ? L.grid_chunks : 0;
        k_items_write<<<L.grid_items + wblocks, kThreads, 0, st>>>(
            static_cast<int>(L.H), w.order, w.icnt, w.sym_icnt, w.tst_len, tile, L.plan->n_items,
            w.sym_cap, w.items, w.sym_items, w.itotal, w.sym_itotal, w.sym_head, L.grid_items,
            w.chunks, nc, w.off, w.active, w.ccnt, w.state, w.p4, w.src4, w.cidx, w.rem4);
        HALMA_LAUNCH_CHECK();
        return 0;
    } else {
        {
            const int rc = exclusive_scan_i64(w.icnt, w.ioff, L.H, w.itotal, w.iscan_ws, st);
            if (rc) return rc;
            const int64_t cap = L.plan->n_items;
            const unsigned fill_grid = static_cast<unsigned>(
                std::max<int64_t>(1, std::min<int64_t>(1024, (cap + 255) / 256)));
            k_item_fill<<<fill_grid, kThreads, 0, st>>>(static_cast<int>(L.H), w.order, w.ioff,
                                                         w.itotal, w.tst_len, tile, cap, w.items);
            HALMA_LAUNCH_CHECK();
        }
        {
            const int rc = exclusive_scan_i64(w.sym_icnt, w.sym_ioff, L.H, w.sym_itotal,
                                              w.sym_scan_ws, st);
            if (rc) return rc;
            const unsigned sym_grid = static_cast<unsigned>(
                std::max<int64_t>(1, std::min<int64_t>(1024, (w.sym_cap + 255) / 256)));
            k_sym_item_fill<<<sym_grid, kThreads, 0, st>>>(static_cast<int>(L.H), w.order,
                                                           w.sym_ioff, w.sym_itotal, w.sym_cap,
                                                           w.sym_items, w.sym_head);
            HALMA_LAUNCH_CHECK();
        }
    }
    if (nc > 0) {
        k_chunk_write<<<L.grid_chunks, kThreads, 0, st>>>(w.chunks, nc, w.off, w.active, w.ccnt,
                                                          w.state, w.p4, w.src4, w.cidx, w.rem4);
        HALMA_LAUNCH_CHECK();
    }
    return 0;
}

// The queue heads are cleared by the item fills in the prepare part, so that the two
// pair-kernel parts hold one kernel each (their timing events bracket nothing else).
int launch_pot_sym_part(const PassLaunch& L, hipStream_t st) {
    const UnbWs& w = L.w;
    return launch_pot_sym(*L.splan, w.src4, w.sym, w.n_bound, w.sym_items, w.sym_itotal,
                          w.sym_part, w.sym_head, st, false);
}

int launch_pot_f32_part(const PassLaunch& L, hipStream_t st) {
    const UnbWs& w = L.w;
    DynSets dyn;
    dyn.tst_len = w.tst_len;
    dyn.src_len = w.src_len;
    dyn.src_sel = w.src_sel;
    dyn.src_alt = w.rem4;
    dyn.n_items = w.itotal;
    dyn.chunk = w.pchunk;
    return launch_pot_f32(*L.plan, w.src4, w.src4, w.halos, dyn, w.items, w.part, st);
}

int launch_post(const PassLaunch& L, hipStream_t st) {
    const UnbWs& w = L.w;
    if (L.nm > 0) {
        k_energy<<<blocks_for(L.nm), kThreads, 0, st>>>(L.ea);
        HALMA_LAUNCH_CHECK();
    }
    // re-tests, then the end of the pass
    k_retest<<<L.grid_retest, kOrdWidth, 0, st>>>(L.ea, w.src4);
    HALMA_LAUNCH_CHECK();
    k_finish<<<blocks_for(L.H), kThreads, 0, st>>>(L.ea);
    HALMA_LAUNCH_CHECK();
    return 0;
}

typedef int (*PartFn)(const PassLaunch&, hipStream_t);

int capture_part(PartFn fn, const PassLaunch& L, hipStream_t cs, hipGraph_t* g,
                 hipGraphExec_t* e) {
    HALMA_HIP(hipStreamBeginCapture(cs, hipStreamCaptureModeRelaxed));
    const int rc = fn(L, cs);
    hipGraph_t g2 = nullptr;
    const hipError_t ec = hipStreamEndCapture(cs, &g2);
    if (rc) {
        if (g2) (void)hipGraphDestroy(g2);
        return rc;
    }
    HALMA_HIP(ec);
    *g = g2;
    HALMA_HIP(hipGraphInstantiate(e, g2, nullptr, nullptr, 0));
    return 0;
}

// The plans of the last haloes planned on this thread: the workspace query and the run
// of one unbinding plan the same haloes, and so do repeated runs over one catalogue.
// Keyed by the relative offsets, the tuning and the device's CU count.
struct UnbPlans {
    std::vector<int64_t> rel;
    Tuning tun;
    int cus = -1;
    bool valid = false;
    PotPlan plan;
    SymPlan splan;
};

int unb_plans(int64_t H, const int64_t* h_off, const UnbPlans** out) {
    thread_local UnbPlans c;
    const int cus = device_cu_count();
    const Tuning t = g_tuning;
    bool hit = c.valid && c.cus == cus && static_cast<int64_t>(c.rel.size()) == H + 1 &&
               c.tun.tests_per_lane == t.tests_per_lane && c.tun.subtile == t.subtile &&
               c.tun.min_chunk == t.min_chunk && c.tun.items_per_wave == t.items_per_wave &&
               c.tun.waves_per_cu == t.waves_per_cu && c.tun.sym_max == t.sym_max;
    for (int64_t h = 0; hit && h <= H; ++h) hit = c.rel[h] == h_off[h] - h_off[0];
    if (!hit) {
        c.valid = false;
        c.rel.resize(H + 1);
        for (int64_t h = 0; h <= H; ++h) c.rel[h] = h_off[h] - h_off[0];
        int rc = plan_pot(H, c.rel.data(), c.rel.data(), c.plan, false);
        if (rc) return rc;
        // haloes above the limit plan no slots (their full passes take the chunked kernel);
        // a limit of 0 plans none at all
        rc = plan_sym(H, c.rel.data(), c.splan, false, t.sym_max > 0 ? t.sym_max : -1);
        if (rc) return rc;
        c.tun = t;
        c.cus = cus;
        c.valid = true;
    }
    *out = &c;
    return 0;
}

int unbind_run(int64_t H, const int64_t* h_off, const int32_t* members, Inputs in,
               double factor, int max_iter, int flags, uint8_t* bound_out, int32_t* iters_out,
               double* vcm_out, double* s_out, int64_t* stats, void* ws, size_t ws_bytes,
               hipStream_t st) {
    const int64_t nm = h_off[H] - h_off[0];
    // every halo's tests are its own (live) members
    const UnbPlans* P = nullptr;
    int rc = unb_plans(H, h_off, &P);
    if (rc) return rc;
    const std::vector<int64_t>& rel = P->rel;
    const PotPlan& plan = P->plan;
    const SymPlan& splan = P->splan;
    const int64_t n_chunks = count_chunks(H, h_off);
    if (n_chunks > INT32_MAX) return set_error(HALMA_E_LIMIT, "too many members");
    UnbWs w = carve_unb(ws, H, nm, plan, splan, n_chunks);
    if (ws_bytes < w.bytes)
        return set_error(HALMA_E_WORKSPACE, "workspace %zu < required %zu", ws_bytes, w.bytes);
    std::vector<int2> chunks(static_cast<size_t>(n_chunks));
    std::vector<int32_t> cfirst(H + 1);
    {
        int32_t q = 0;
        for (int64_t h = 0; h < H; ++h) {
            cfirst[h] = q;
            const int64_t nch = (rel[h + 1] - rel[h] + kChunk - 1) / kChunk;
            for (int64_t c = 0; c < nch; ++c)
                chunks[q++] = make_int2(static_cast<int>(h), static_cast<int>(c));
        }
        cfirst[H] = q;
    }
    const UploadPiece pieces[6] = {
        {w.t_off, rel.data(), (H + 1) * sizeof(int64_t)},
        {w.t_halos, plan.halos.data(), plan.halos.size() * sizeof(HaloDesc)},
        {w.t_order, plan.order.data(), plan.order.size() * sizeof(int32_t)},
        {w.t_chunks, chunks.data(), chunks.size() * sizeof(int2)},
        {w.t_cfirst, cfirst.data(), cfirst.size() * sizeof(int32_t)},
        {w.t_sym, splan.halos.data(), splan.halos.size() * sizeof(SymDesc)}};
    rc = upload_pieces(w.tables, w.t_bytes, pieces, 6, st);
    if (rc) return rc;
    const int32_t* mem = members + h_off[0];
    if (nm > 0) {
        const bool one_type = in.pos_t == in.vel_t && in.pos_t == in.mass_t;
        auto gather = one_type ? (in.pos_t == HALMA_F32 ? k_gather<1> : k_gather<2>) : k_gather<0>;
        gather<<<blocks_for(nm), kThreads, 0, st>>>(in, mem, nm, w.p4, w.vel, w.m64, w.state);
        HALMA_LAUNCH_CHECK();
        k_member_halo<<<static_cast<unsigned>(H), kThreads, 0, st>>>(w.off, static_cast<int>(H),
                                                                     w.mhalo);
        HALMA_LAUNCH_CHECK();
    }
    k_init_halo<<<blocks_for(H), kThreads, 0, st>>>(w.off, static_cast<int>(H), w.active, w.iters,
                                                    w.removed, w.stale, w.need_full, w.refresh);
    HALMA_LAUNCH_CHECK();
    k_zero_counters<<<1, 64, 0, st>>>(w.counters, 16);
    HALMA_LAUNCH_CHECK();
    k_zero_counters<<<1, 64, 0, st>>>(reinterpret_cast<unsigned long long*>(w.pass_no), 1);
    HALMA_LAUNCH_CHECK();

    double G = 4.3 * 1e-3;
    G *= 1e-6;
    PassLaunch L;
    L.H = H;
    L.nm = nm;
    const int cus = std::max(device_cu_count(), 1);
    L.grid_chunks = static_cast<int>(std::max<int64_t>(1, std::min<int64_t>(n_chunks, cus * 8)));
    // k_chunk_sums: a wavefront per chunk, four per block
    L.grid_chunk_waves =
        static_cast<int>(std::max<int64_t>(1, std::min<int64_t>((n_chunks + 3) / 4, cus * 8)));
    // item lists: every block scans the per-halo counts itself, so one block per compute
    // unit (more would repeat that scan in a second round)
    const int64_t most_items = std::max<int64_t>(plan.n_items, w.sym_cap);
    L.grid_items = static_cast<int>(
        std::max<int64_t>(1, std::min<int64_t>(cus, (most_items + 255) / 256)));
    // re-tests: blocks (large live sets) and wavefronts (small ones) over the band lists
    L.grid_retest = cus * 4;
    L.plan = &plan;
    L.splan = &splan;
    L.w = w;
    L.bound = bound_out;
    L.ea.halos = w.halos;
    L.ea.mhalo = w.mhalo;
    L.ea.n_bound = w.n_bound;
    L.ea.n_src = w.src_len;
    L.ea.pchunk = w.pchunk;
    L.ea.mode = w.mode;
    L.ea.cidx = w.cidx;
    L.ea.part = w.part;
    L.ea.sym = w.sym;
    L.ea.sym_part = w.sym_part;
    L.ea.vel = w.vel;
    L.ea.vcm = w.vcm;
    L.ea.cg = factor * G;
    L.ea.bound_rel = pot_error_bound(plan.subtile);
    L.ea.bound_sym = pot_sym_error_bound();
    L.ea.bound_resum = pot_error_bound(1);
    L.ea.state = w.state;
    L.ea.sacc = w.sacc;
    L.ea.eacc = w.eacc;
    L.ea.removed = w.removed;
    L.ea.band = w.band;
    L.ea.band_cap = nm > 0 ? nm : 1;  // the entries of w.band
    L.ea.stale = w.stale;
    L.ea.need_full = w.need_full;
    L.ea.refresh = w.refresh;
    L.ea.counters = w.counters;
    L.ea.s_out = s_out;
    L.ea.nm = nm;
    L.ea.H = static_cast<int>(H);
    L.ea.active = w.active;
    L.ea.iters = w.iters;
    L.ea.pass_no = w.pass_no;

    const bool use_graph = (flags & HALMA_UNBIND_GRAPHS) != 0;
    const bool timing = (flags & HALMA_UNBIND_TIME_POT) != 0;
    // Passes are launched one ahead of the host: pass p + 1 is queued before pass p's
    // counters are read, so the GPU never waits for the host between passes.  Pass p's
    // counters go to slot p % 2 of a pinned buffer, with slot events; when pass p removes
    // nobody, the queued pass p + 1 finds no active halo and does nothing.
    hipEvent_t ev[2][3] = {{nullptr, nullptr, nullptr}, {nullptr, nullptr, nullptr}};
    hipEvent_t done[2] = {nullptr, nullptr};
    for (auto& d : done) HALMA_HIP(hipEventCreateWithFlags(&d, hipEventDisableTiming));
    if (timing)
        for (auto& q : ev)
            for (auto& e : q) HALMA_HIP(hipEventCreate(&e));
    double pot_ms = 0.0, sym_ms = 0.0;
    long long pot_launches = 0, sym_launches = 0, f32_launches = 0;
    const PartFn parts[4] = {launch_pre, launch_pot_sym_part, launch_pot_f32_part, launch_post};
    hipGraph_t graphs[4] = {nullptr, nullptr, nullptr, nullptr};
    hipGraphExec_t execs[4] = {nullptr, nullptr, nullptr, nullptr};
    auto cleanup = [&]() {
        for (int q = 0; q < 4; ++q) {
            if (execs[q]) (void)hipGraphExecDestroy(execs[q]);
            if (graphs[q]) (void)hipGraphDestroy(graphs[q]);
        }
        for (auto& q : ev)
            for (auto& e : q)
                if (e) (void)hipEventDestroy(e);
        for (auto& d : done)
            if (d) (void)hipEventDestroy(d);
    };
    // every pass launches both pair kernels (full passes of haloes above the symmetric
    // limit run on the chunked kernel); the one without items returns at once
    if (use_graph) {
        hipStream_t cs = capture_stream();
        if (!cs) {
            cleanup();
            return set_error(HALMA_E_STATE, "cannot create the capture stream");
        }
        for (int q = 0; q < 4; ++q) {
            rc = capture_part(parts[q], L, cs, &graphs[q], &execs[q]);
            if (rc) {
                cleanup();
                return rc;
            }
        }
    }
    auto run_part = [&](int q) -> int {
        if (use_graph) {
            const hipError_t e = hipGraphLaunch(execs[q], st);
            return e == hipSuccess ? 0 : hip_fail(e, "hipGraphLaunch", __FILE__, __LINE__);
        }
        return parts[q](L, st);
    };
    long long* host = pinned_words(16);
    if (!host) {
        cleanup();
        return set_error(HALMA_E_STATE, "pinned host buffer unavailable");
    }
    auto record = [&](hipEvent_t e) -> int {
        return hipEventRecord(e, st) == hipSuccess
                   ? 0
                   : set_error(HALMA_E_STATE, "event record failed");
    };
    auto launch_pass = [&](int p) -> int {
        const int q = p & 1;
        int r = run_part(0);
        if (!r && timing) r = record(ev[q][0]);
        if (!r) r = run_part(1);
        if (!r && timing) r = record(ev[q][1]);
        if (!r) r = run_part(2);
        if (!r && timing) r = record(ev[q][2]);
        if (!r) r = run_part(3);
        if (r) return r;
        const hipError_t e = hipMemcpyAsync(host + 8 * q, w.counters, 7 * sizeof(long long),
                                            hipMemcpyDeviceToHost, st);
        if (e != hipSuccess) return hip_fail(e, "counter read-back", __FILE__, __LINE__);
        return record(done[q]);
    };
    // waits for pass p and adds its pair-kernel times
    auto collect = [&](int p) -> int {
        const int q = p & 1;
        const hipError_t e = hipEventSynchronize(done[q]);
        if (e != hipSuccess) return hip_fail(e, "pass completion", __FILE__, __LINE__);
        sym_launches += 1;
        f32_launches += 1;
        if (timing) {
            float ms_all = 0.0f, ms_sym = 0.0f;
            hipError_t e4 = hipEventElapsedTime(&ms_all, ev[q][0], ev[q][2]);
            if (e4 == hipSuccess) e4 = hipEventElapsedTime(&ms_sym, ev[q][0], ev[q][1]);
            if (e4 != hipSuccess) return hip_fail(e4, "pair-kernel timing", __FILE__, __LINE__);
            pot_ms += ms_all;
            sym_ms += ms_sym;
            ++pot_launches;
        }
        return 0;
    };
    int passes = 0, launched = 0;
    long long band_total = 0;
    const long long* last = host;
    int status = max_iter > 0 ? launch_pass(0) : 0;
    if (max_iter > 0) launched = 1;
    while (!status && passes < launched) {
        const int p = passes;
        if (launched < max_iter) {
            status = launch_pass(launched);
            if (status) break;
            ++launched;
        }
        status = collect(p);
        if (status) break;
        last = host + 8 * (p & 1);
        band_total = last[2];
        passes = p + 1;
        if (last[5] == 0) break;  // converged: a queued pass, if any, does nothing
    }
    // the queued pass after convergence: its (empty) kernels are counted like the others
    if (!status && launched > passes) status = collect(passes);
    cleanup();
    if (status) return status;
    if (nm > 0) {
        k_final_mask<<<blocks_for(nm), kThreads, 0, st>>>(nm, w.state, bound_out);
        HALMA_LAUNCH_CHECK();
    }
    if (iters_out) HALMA_HIP(hipMemcpyAsync(iters_out, w.iters, H * sizeof(int32_t),
                                            hipMemcpyDeviceToDevice, st));
    if (vcm_out) HALMA_HIP(hipMemcpyAsync(vcm_out, w.vcm, 3 * H * sizeof(double),
                                          hipMemcpyDeviceToDevice, st));
    if (stats) {
        stats[0] = passes;
        stats[1] = band_total;
        stats[2] = splan.n_items;
        stats[3] = plan.tests_per_lane;
        stats[4] = plan.subtile;
        stats[5] = static_cast<int64_t>(pot_ms * 1e6);  // ns
        stats[6] = pot_launches;
        stats[7] = last[3];
        stats[8] = static_cast<int64_t>(sym_ms * 1e6);  // ns
        stats[9] = sym_launches;
        stats[10] = last[4];
        stats[11] = f32_launches;
        stats[12] = last[6];
    }
    return 0;
}

// ------------------------------------------------------------ plain FP64 path ----
// One halo at a time, one launch per step, FP64 storage and arithmetic, sequential
// per-thread sums: the comparator the batched path is measured against.
__global__ void k_plain_gather(Inputs in, const int32_t* __restrict__ mem, int64_t n,
                               double* __restrict__ px, double* __restrict__ py,
                               double* __restrict__ pz, double* __restrict__ pm,
                               double* __restrict__ vel, double* __restrict__ m64,
                               uint8_t* __restrict__ bound) {
    const int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int64_t i = mem[k];
    // positions and masses as the FP32 values the pair sum uses, held in FP64
    px[k] = static_cast<float>(ldf(in.x, in.pos_t, i));
    py[k] = static_cast<float>(ldf(in.y, in.pos_t, i));
    pz[k] = static_cast<float>(ldf(in.z, in.pos_t, i));
    pm[k] = static_cast<float>(ldf(in.m, in.mass_t, i));
    vel[3 * k + 0] = ldf(in.vx, in.vel_t, i);
    vel[3 * k + 1] = ldf(in.vy, in.vel_t, i);
    vel[3 * k + 2] = ldf(in.vz, in.vel_t, i);
    m64[k] = ldf(in.m, in.mass_t, i);
    bound[k] = 1;
}

// single block: compaction of the bound set of one halo and its v_cm (plain FP64 sums)
__global__ __launch_bounds__(kThreads) void k_plain_compact(
    int64_t n, const uint8_t* __restrict__ bound, const double* __restrict__ px,
    const double* __restrict__ py, const double* __restrict__ pz, const double* __restrict__ pm,
    const double* __restrict__ vel, const double* __restrict__ m64, double* __restrict__ cx,
    double* __restrict__ cy, double* __restrict__ cz, double* __restrict__ cm,
    int32_t* __restrict__ cidx, int32_t* __restrict__ cnt, double* __restrict__ vcm) {
    __shared__ int shs[kThreads / kWave];
    __shared__ double red[4][kThreads];
    int run = 0;
    double s[4] = {0, 0, 0, 0};
    for (int64_t k0 = 0; k0 < n; k0 += kThreads) {
        const int64_t k = k0 + threadIdx.x;
        const int keep = (k < n && bound[k]) ? 1 : 0;
        int excl;
        const int tot = block_excl_scan(keep, excl, shs);
        if (keep) {
            const int64_t pos = run + excl;
            cx[pos] = px[k];
            cy[pos] = py[k];
            cz[pos] = pz[k];
            cm[pos] = pm[k];
            cidx[pos] = static_cast<int32_t>(k);
            s[0] += m64[k];
            for (int d = 0; d < 3; ++d) s[1 + d] += m64[k] * vel[3 * k + d];
        }
        run += tot;
    }
    for (int q = 0; q < 4; ++q) red[q][threadIdx.x] = s[q];
    __syncthreads();
    if (threadIdx.x == 0) {
        double t[4] = {0, 0, 0, 0};
        for (int k = 0; k < kThreads; ++k)
            for (int q = 0; q < 4; ++q) t[q] += red[q][k];
        for (int d = 0; d < 3; ++d) vcm[d] = t[0] > 0.0 ? t[1 + d] / t[0] : 0.0;
        *cnt = run;
    }
}

__global__ void k_plain_energy(const int32_t* __restrict__ cnt, const int32_t* __restrict__ cidx,
                               const double* __restrict__ S, const double* __restrict__ vel,
                               const double* __restrict__ vcm, double cg,
                               uint8_t* __restrict__ bound, int32_t* __restrict__ removed) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= *cnt) return;
    const int k = cidx[p];
    const double E = kinetic(vel, k, vcm) - cg * S[p];
    if (E > 0.0) {
        bound[k] = 0;
        atomicAdd(removed, 1);
    }
}

}  // namespace

size_t unbind_workspace_bytes(int64_t H, const int64_t* h_off) {
    const UnbPlans* P = nullptr;
    if (unb_plans(H, h_off, &P)) return 0;
    return carve_unb(nullptr, H, h_off[H] - h_off[0], P->plan, P->splan, count_chunks(H, h_off))
        .bytes;
}

int unbind_plain_run(int64_t H, const int64_t* h_off, const int32_t* members, Inputs in,
                     double factor, int max_iter, uint8_t* bound_out, int32_t* iters_host,
                     void* ws, size_t ws_bytes, hipStream_t st) {
    int64_t nmax = 0;
    for (int64_t h = 0; h < H; ++h) nmax = std::max(nmax, h_off[h + 1] - h_off[h]);
    Carver c(ws);
    const int64_t n1 = nmax > 0 ? nmax : 1;
    double* px = c.take<double>(n1);
    double* py = c.take<double>(n1);
    double* pz = c.take<double>(n1);
    double* pm = c.take<double>(n1);
    double* vel = c.take<double>(3 * n1);
    double* m64 = c.take<double>(n1);
    double* cx = c.take<double>(n1);
    double* cy = c.take<double>(n1);
    double* cz = c.take<double>(n1);
    double* cm = c.take<double>(n1);
    int32_t* cidx = c.take<int32_t>(n1);
    double* S = c.take<double>(n1);
    double* vcm = c.take<double>(4);
    int32_t* cnt = c.take<int32_t>(64);
    int32_t* removed = c.take<int32_t>(64);
    if (ws_bytes < c.used) return set_error(HALMA_E_WORKSPACE, "workspace too small");
    double G = 4.3 * 1e-3;
    G *= 1e-6;
    const double cg = factor * G;
    long long* host = pinned_words(2);
    if (!host) return set_error(HALMA_E_STATE, "pinned host buffer unavailable");
    int* hcnt = reinterpret_cast<int*>(host);
    for (int64_t h = 0; h < H; ++h) {
        const int64_t n = h_off[h + 1] - h_off[h];
        if (iters_host) iters_host[h] = 0;
        if (n == 0) continue;
        uint8_t* bh = bound_out + (h_off[h] - h_off[0]);
        k_plain_gather<<<blocks_for(n), kThreads, 0, st>>>(in, members + h_off[h], n, px, py, pz,
                                                           pm, vel, m64, bh);
        HALMA_LAUNCH_CHECK();
        for (int it = 0; it < max_iter; ++it) {
            k_plain_compact<<<1, kThreads, 0, st>>>(n, bh, px, py, pz, pm, vel, m64, cx, cy, cz,
                                                    cm, cidx, cnt, vcm);
            HALMA_LAUNCH_CHECK();
            HALMA_HIP(hipMemcpyAsync(hcnt, cnt, sizeof(int), hipMemcpyDeviceToHost, st));
            HALMA_HIP(hipStreamSynchronize(st));
            const int live = hcnt[0];
            if (live == 0) break;
            if (iters_host) iters_host[h] = it + 1;
            int rc = halma_pot_sum_f64_plain(cm, cx, cy, cz, live, cx, cy, cz, live, S, st);
            if (rc) return rc;
            k_zero_counters<<<1, 64, 0, st>>>(reinterpret_cast<unsigned long long*>(removed), 1);
            HALMA_LAUNCH_CHECK();
            k_plain_energy<<<blocks_for(live), kThreads, 0, st>>>(cnt, cidx, S, vel, vcm, cg, bh,
                                                                  removed);
            HALMA_LAUNCH_CHECK();
            HALMA_HIP(hipMemcpyAsync(hcnt, removed, sizeof(int), hipMemcpyDeviceToHost, st));
            HALMA_HIP(hipStreamSynchronize(st));
            if (hcnt[0] == 0) break;
        }
    }
    return 0;
}

size_t unbind_plain_workspace_bytes(int64_t H, const int64_t* h_off) {
    int64_t nmax = 0;
    for (int64_t h = 0; h < H; ++h) nmax = std::max(nmax, h_off[h + 1] - h_off[h]);
    const int64_t n1 = nmax > 0 ? nmax : 1;
    Carver c(nullptr);
    for (int k = 0; k < 4; ++k) c.take<double>(n1);
    c.take<double>(3 * n1);
    c.take<double>(n1);
    for (int k = 0; k < 4; ++k) c.take<double>(n1);
    c.take<int32_t>(n1);
    c.take<double>(n1);
    c.take<double>(4);
    c.take<int32_t>(64);
    c.take<int32_t>(64);
    return c.used;
}

}  // namespace halma

using namespace halma;

namespace {
int check_unbind_args(int64_t H, const int64_t* h_off, const int32_t* members, const void* x,
                      const void* y, const void* z, int pos_t, const void* vx, const void* vy,
                      const void* vz, int vel_t, const void* m, int mass_t, double factor,
                      int max_iter, uint8_t* bound) {
    if (H < 0 || !h_off) return set_error(HALMA_E_ARG, "bad halo offsets");
    for (int64_t h = 0; h < H; ++h)
        if (h_off[h + 1] < h_off[h]) return set_error(HALMA_E_ARG, "offsets must not decrease");
    if (H > INT32_MAX) return set_error(HALMA_E_LIMIT, "too many haloes");
    if (h_off[H] - h_off[0] > INT32_MAX) return set_error(HALMA_E_LIMIT, "too many members");
    for (int64_t h = 0; h < H; ++h)
        if (h_off[h + 1] - h_off[h] > INT32_MAX) return set_error(HALMA_E_LIMIT, "halo too large");
    if (h_off[H] > h_off[0] &&
        (!members || !x || !y || !z || !vx || !vy || !vz || !m || !bound))
        return set_error(HALMA_E_ARG, "null pointer");
    auto bad = [](int t) { return t != HALMA_F32 && t != HALMA_F64; };
    if (bad(pos_t) || bad(vel_t) || bad(mass_t)) return set_error(HALMA_E_ARG, "bad dtype");
    if (!(factor >= 0.0) || !isfinite(factor))
        return set_error(HALMA_E_ARG, "factor must be finite and non-negative");
    if (max_iter < 1) return set_error(HALMA_E_ARG, "max_iter must be >= 1");
    return 0;
}

Inputs make_inputs(const void* x, const void* y, const void* z, int pos_t, const void* vx,
                   const void* vy, const void* vz, int vel_t, const void* m, int mass_t) {
    Inputs in;
    in.x = x;
    in.y = y;
    in.z = z;
    in.pos_t = pos_t;
    in.vx = vx;
    in.vy = vy;
    in.vz = vz;
    in.vel_t = vel_t;
    in.m = m;
    in.mass_t = mass_t;
    return in;
}
}  // namespace

extern "C" {

size_t halma_unbind_workspace_size(int64_t n_halo, const int64_t* offsets) {
    if (n_halo < 0 || !offsets) return 0;
    return unbind_workspace_bytes(n_halo, offsets);
}

int halma_unbind_batched(int64_t n_halo, const int64_t* offsets, const int32_t* members,
                         const void* x, const void* y, const void* z, int pos_dtype,
                         const void* vx, const void* vy, const void* vz, int vel_dtype,
                         const void* m, int mass_dtype, double factor, int max_iter, int flags,
                         uint8_t* bound, int32_t* iters, double* vcm, double* pot,
                         int64_t* stats, void* ws, size_t ws_bytes, void* stream) {
    int rc = check_unbind_args(n_halo, offsets, members, x, y, z, pos_dtype, vx, vy, vz,
                               vel_dtype, m, mass_dtype, factor, max_iter, bound);
    if (rc) return rc;
    if (n_halo == 0) return 0;
    for (int64_t h = 0; h < n_halo; ++h)
        if (offsets[h + 1] - offsets[h] < 0) return set_error(HALMA_E_ARG, "bad offsets");
    Inputs in = make_inputs(x, y, z, pos_dtype, vx, vy, vz, vel_dtype, m, mass_dtype);
    return unbind_run(n_halo, offsets, members, in, factor, max_iter, flags, bound, iters, vcm,
                      pot, stats, ws, ws_bytes, static_cast<hipStream_t>(stream));
}

size_t halma_unbind_plain_workspace_size(int64_t n_halo, const int64_t* offsets) {
    if (n_halo < 0 || !offsets) return 0;
    return unbind_plain_workspace_bytes(n_halo, offsets);
}

int halma_unbind_plain_f64(int64_t n_halo, const int64_t* offsets, const int32_t* members,
                           const void* x, const void* y, const void* z, int pos_dtype,
                           const void* vx, const void* vy, const void* vz, int vel_dtype,
                           const void* m, int mass_dtype, double factor, int max_iter,
                           uint8_t* bound, int32_t* iters_host, void* ws, size_t ws_bytes,
                           void* stream) {
    int rc = check_unbind_args(n_halo, offsets, members, x, y, z, pos_dtype, vx, vy, vz,
                               vel_dtype, m, mass_dtype, factor, max_iter, bound);
    if (rc) return rc;
    Inputs in = make_inputs(x, y, z, pos_dtype, vx, vy, vz, vel_dtype, m, mass_dtype);
    return unbind_plain_run(n_halo, offsets, members, in, factor, max_iter, bound, iters_host,
                            ws, ws_bytes, static_cast<hipStream_t>(stream));
}

}  // extern "C"
