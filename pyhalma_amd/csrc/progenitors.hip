// Progenitor matching (pyHALMA.py:1211-1266) for every new halo at once, as a hash join
// on particle IDs instead of one np.in1d per (new, old) halo pair.
//
// 1. IDs become unsigned keys (sign bit flipped) relative to the smallest ID of both
//    catalogues.  Every old member is entered in an ID table: indexed by the key itself
//    when the keys span a small range (IDs that are particle numbers), else an
//    open-addressing hash (linear probing, Fibonacci hash).  Members sharing an ID (an ID
//    listed in several old haloes) are chained from the ID's entry.
// 2. Every new member looks its ID up and keeps the chain's head.
// 3. One block per new halo walks its members' chains into a hash set of old haloes in
//    LDS (haloes with more than 2048 distinct old haloes go, one after another, through
//    a global-memory set of 2 slots per old halo): the number of old haloes that share an ID with it (n_mergers) and the two
//    most massive (equal masses: the higher old index first, as np.flip of a stable
//    argsort), then the reference's classification: -1 when 1.2 * mass < the main
//    progenitor's mass; with two or more, 2 when 1/20 < second / first < 1/3 and 3
//    otherwise (the type 1 of second / first > 1/3 is overwritten by the following
//    if / else, as there).
// Chains and sets are filled with atomics, so their order varies from run to run; what is
// read from them (a set, its size, its two largest under a strict order) does not.
#include <algorithm>

#include "halma_common.h"

namespace halma {

namespace {

constexpr int kThreads = 256;
constexpr int kRangeBlocks = 256;
constexpr unsigned long long kEmpty = ~0ull;

inline unsigned blocks_for(int64_t n) {
    return static_cast<unsigned>((n + kThreads - 1) / kThreads);
}

__device__ __forceinline__ unsigned long long id_key(int64_t v) {
    return static_cast<unsigned long long>(v) ^ 0x8000000000000000ull;
}

__device__ __forceinline__ unsigned long long fib_hash(unsigned long long k, int shift) {
    return (k * 0x9E3779B97F4A7C15ull) >> shift;
}

// smallest and largest key of ids[0..n): one (lo, hi) partial per block ...
__global__ __launch_bounds__(kThreads) void k_range(const int64_t* __restrict__ ids, int64_t n,
                                                    unsigned long long* __restrict__ part) {
    __shared__ unsigned long long sh[2 * (kThreads / kWave)];
    unsigned long long lo = ~0ull, hi = 0ull;
    for (int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; k < n;
         k += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const unsigned long long v = id_key(ids[k]);
        lo = v < lo ? v : lo;
        hi = v > hi ? v : hi;
    }
    for (int s = kWave / 2; s >= 1; s >>= 1) {
        const unsigned long long a = __shfl_down(lo, s, kWave);
        const unsigned long long b = __shfl_down(hi, s, kWave);
        lo = a < lo ? a : lo;
        hi = b > hi ? b : hi;
    }
    const int w = threadIdx.x / kWave;
    if (__lane_id() == 0) {
        sh[2 * w] = lo;
        sh[2 * w + 1] = hi;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int q = 1; q < kThreads / kWave; ++q) {
            lo = sh[2 * q] < lo ? sh[2 * q] : lo;
            hi = sh[2 * q + 1] > hi ? sh[2 * q + 1] : hi;
        }
        part[2 * blockIdx.x] = lo;
        part[2 * blockIdx.x + 1] = hi;
    }
}

// ... then over the np partials (one wavefront)
__global__ void k_range_final(const unsigned long long* __restrict__ part, int np,
                              unsigned long long* __restrict__ r) {
    unsigned long long lo = ~0ull, hi = 0ull;
    for (int q = __lane_id(); q < np; q += kWave) {
        lo = part[2 * q] < lo ? part[2 * q] : lo;
        hi = part[2 * q + 1] > hi ? part[2 * q + 1] : hi;
    }
    for (int s = kWave / 2; s >= 1; s >>= 1) {
        const unsigned long long a = __shfl_down(lo, s, kWave);
        const unsigned long long b = __shfl_down(hi, s, kWave);
        lo = a < lo ? a : lo;
        hi = b > hi ? b : hi;
    }
    if (__lane_id() == 0) {
        r[0] = lo;
        r[1] = hi;
    }
}

__global__ void k_fill_u64(unsigned long long* __restrict__ a, int64_t n, unsigned long long v) {
    const int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k < n) a[k] = v;
}

__global__ void k_fill_i32(int32_t* __restrict__ a, int64_t n, int32_t v) {
    const int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k < n) a[k] = v;
}

struct Table {
    unsigned long long* key;  // hash: cap slots, kEmpty when free (unused when direct)
    int32_t* head;            // per slot (hash) or per key (direct): last old member entered
    int32_t* next;            // per old member: the previous member with the same ID, or -1
    int32_t* halo;            // per old member: its old halo
    int shift;                // 64 - log2(cap)
    unsigned long long mask;  // cap - 1
    int direct;               // head indexed by the key itself
};

// the table entry of a key: its slot, or -1 when absent (lookups) / claimed (inserts)
__device__ __forceinline__ int64_t table_slot(const Table& t, unsigned long long key,
                                              bool insert) {
    if (t.direct) return static_cast<int64_t>(key);
    unsigned long long s = fib_hash(key, t.shift);
    for (;;) {  // the table is at most half full: a free or matching slot is always found
        if (insert) {
            const unsigned long long prev = atomicCAS(t.key + s, kEmpty, key);
            if (prev == kEmpty || prev == key) return static_cast<int64_t>(s);
        } else {
            const unsigned long long v = t.key[s];
            if (v == kEmpty) return -1;
            if (v == key) return static_cast<int64_t>(s);
        }
        s = (s + 1) & t.mask;
    }
}

// the old halo of every old member (one block per halo, grid-stride)
__global__ void k_old_halo(const int64_t* __restrict__ off, int64_t G, int32_t* __restrict__ halo) {
    for (int64_t g = blockIdx.x; g < G; g += gridDim.x)
        for (int64_t k = off[g] + threadIdx.x; k < off[g + 1]; k += blockDim.x)
            halo[k] = static_cast<int32_t>(g);
}

// direct table, unique IDs without atomics: every old member stores itself at its key
// (one of several members sharing a key wins) ...
__global__ void k_direct_store(int64_t n, const int64_t* __restrict__ ids,
                               const unsigned long long* __restrict__ r, int32_t* __restrict__ head) {
    const int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= n) return;
    head[id_key(ids[k]) - r[0]] = static_cast<int32_t>(k);
}

// ... and the others chain themselves behind the winner (next[] starts at -1)
__global__ void k_direct_chain(int64_t n, const int64_t* __restrict__ ids,
                               const unsigned long long* __restrict__ r,
                               const int32_t* __restrict__ head, int32_t* __restrict__ next) {
    const int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int32_t w = head[id_key(ids[k]) - r[0]];
    if (w != static_cast<int32_t>(k)) next[k] = atomicExch(next + w, static_cast<int32_t>(k));
}

// hash table: every old member into the ID table
__global__ void k_table_build(int64_t n, const int64_t* __restrict__ ids,
                              const unsigned long long* __restrict__ r, Table t) {
    const int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const unsigned long long key = id_key(ids[k]) - r[0];
    const int64_t s = table_slot(t, key, true);
    t.next[k] = atomicExch(t.head + s, static_cast<int32_t>(k));
}

// every new member: the old halo of the head of its ID's chain (-1: none) and the rest of
// the chain (-1: the ID is in one old member only)
__global__ void k_member_lookup(int64_t n, const int64_t* __restrict__ ids,
                                const unsigned long long* __restrict__ r, Table t,
                                int32_t* __restrict__ mhalo, int32_t* __restrict__ mrest) {
    const int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const unsigned long long key = id_key(ids[k]) - r[0];
    const int64_t s = key <= r[1] - r[0] ? table_slot(t, key, false) : -1;
    const int32_t j = s >= 0 ? t.head[s] : -1;
    mhalo[k] = j >= 0 ? t.halo[j] : -1;
    mrest[k] = j >= 0 ? t.next[j] : -1;
}

// a ranks before b: larger mass, or equal mass and larger old index
__device__ __forceinline__ bool ranks_before(double ma, int32_t a, double mb, int32_t b) {
    return b < 0 || (a >= 0 && (ma > mb || (ma == mb && a > b)));
}

struct Top2 {
    int32_t a, b;  // best, second (-1: none)
    double ma, mb;
};

__device__ __forceinline__ void top2_add(Top2& t, int32_t oh, double m) {
    if (ranks_before(m, oh, t.ma, t.a)) {
        t.b = t.a;
        t.mb = t.ma;
        t.a = oh;
        t.ma = m;
    } else if (ranks_before(m, oh, t.mb, t.b)) {
        t.b = oh;
        t.mb = m;
    }
}

constexpr int kRankThreads = 512;
constexpr int kLdsSet = 4096;  // LDS set slots; a halo filling more than half -> fallback

// insert oh into a set of cap slots (-1 free): 1 inserted, 0 already there, -1 no free slot
// (a plain read first: most inserts find their old halo already there)
__device__ __forceinline__ int set_insert(int32_t* set, int64_t cap, int32_t oh) {
    int64_t q = static_cast<int64_t>(fib_hash(static_cast<unsigned long long>(oh), 0) %
                                     static_cast<unsigned long long>(cap));
    for (int64_t probe = 0; probe < cap; ++probe) {
        if (set[q] == oh) return 0;
        const int32_t prev = atomicCAS(set + q, -1, oh);
        if (prev == -1) return 1;
        if (prev == oh) return 0;
        q = q + 1 == cap ? 0 : q + 1;
    }
    return -1;
}

// the block's set -> count, two largest, classification (thread 0 writes the outputs)
__device__ void rank_set(const int32_t* set, int64_t cap, int64_t ih,
                         const double* __restrict__ new_mass, const double* __restrict__ old_mass,
                         Top2* sh, int* count, int32_t* __restrict__ pro1,
                         int32_t* __restrict__ pro2, int32_t* __restrict__ nmergers,
                         int32_t* __restrict__ mtype) {
    Top2 tp{-1, -1, 0.0, 0.0};
    int n = 0;
    for (int64_t q = threadIdx.x; q < cap; q += blockDim.x) {
        const int32_t oh = set[q];
        if (oh >= 0) {
            ++n;
            top2_add(tp, oh, old_mass[oh]);
        }
    }
    // merge the lanes' top twos (shuffle tree), then the waves' in LDS
    for (int m = kWave / 2; m >= 1; m >>= 1) {
        const int32_t oa = __shfl_xor(tp.a, m, kWave), ob = __shfl_xor(tp.b, m, kWave);
        const double oma = __shfl_xor(tp.ma, m, kWave), omb = __shfl_xor(tp.mb, m, kWave);
        if (oa >= 0) top2_add(tp, oa, oma);
        if (ob >= 0) top2_add(tp, ob, omb);
        n += __shfl_xor(n, m, kWave);
    }
    if (__lane_id() == 0) sh[threadIdx.x / kWave] = tp;
    if (threadIdx.x == 0) *count = 0;
    __syncthreads();
    if (__lane_id() == 0 && n) atomicAdd(count, n);
    __syncthreads();
    if (threadIdx.x == 0) {
        const int total = *count;
        Top2 r = sh[0];
        for (int i = 1; i < static_cast<int>(blockDim.x) / kWave; ++i) {
            const Top2 o = sh[i];
            if (o.a >= 0) top2_add(r, o.a, o.ma);
            if (o.b >= 0) top2_add(r, o.b, o.mb);
        }
        int32_t p1 = 0, p2 = 0, ty = 0;
        if (total > 0) {
            p1 = r.a + 1;
            if (1.2 * new_mass[ih] < r.ma) {
                ty = -1;  // the old halo breaking apart
            } else if (total > 1) {
                p2 = r.b + 1;
                const double frac = r.mb / r.ma;
                ty = (1.0 / 20.0 < frac && frac < 1.0 / 3.0) ? 2 : 3;
            }
        }
        pro1[ih] = p1;
        pro2[ih] = p2;
        nmergers[ih] = total;
        mtype[ih] = ty;
    }
    __syncthreads();
}

// one block per new halo: its members' chains into an LDS set of old haloes, then
// rank_set; a halo with more than kLdsSet / 2 distinct old haloes is flagged for
// k_rank_fallback
__global__ __launch_bounds__(kRankThreads) void k_halo_rank(
    int64_t H, const int64_t* __restrict__ noff, const int32_t* __restrict__ mhalo,
    const int32_t* __restrict__ mrest, Table t, const double* __restrict__ new_mass,
    const double* __restrict__ old_mass, int32_t* __restrict__ flag, int32_t* __restrict__ pro1, int32_t* __restrict__ pro2,
    int32_t* __restrict__ nmergers, int32_t* __restrict__ mtype) {
    __shared__ int32_t lset[kLdsSet];
    __shared__ Top2 sh[kRankThreads / kWave];
    __shared__ int count;
    __shared__ int full;
    for (int64_t ih = blockIdx.x; ih < H; ih += gridDim.x) {
        const int64_t b = noff[ih], e = noff[ih + 1];
        for (int q = threadIdx.x; q < kLdsSet; q += blockDim.x) lset[q] = -1;
        if (threadIdx.x == 0) {
            count = 0;
            full = 0;
        }
        __syncthreads();
        for (int64_t k = b + threadIdx.x; k < e; k += blockDim.x) {
            int32_t oh = mhalo[k];
            int32_t j = mrest[k];
            while (oh >= 0 && !full) {
                const int rc = set_insert(lset, kLdsSet, oh);
                if (rc < 0 || (rc > 0 && atomicAdd(&count, 1) + 1 > kLdsSet / 2)) full = 1;
                oh = j >= 0 ? t.halo[j] : -1;
                j = j >= 0 ? t.next[j] : -1;
            }
            if (full) break;
        }
        __syncthreads();
        const bool over = full != 0;  // block-uniform
        if (threadIdx.x == 0) flag[ih] = over ? 1 : 0;
        if (!over)
            rank_set(lset, kLdsSet, ih, new_mass, old_mass, sh, &count, pro1, pro2, nmergers,
                     mtype);
        __syncthreads();
    }
}

// the flagged haloes, one after another, with a global set of 2 slots per old halo (never
// more than half full)
__global__ __launch_bounds__(kRankThreads) void k_rank_fallback(
    int64_t H, int64_t G, const int64_t* __restrict__ noff, const int32_t* __restrict__ mhalo,
    const int32_t* __restrict__ mrest, Table t, const int32_t* __restrict__ flag,
    int32_t* __restrict__ gset,
    const double* __restrict__ new_mass, const double* __restrict__ old_mass,
    int32_t* __restrict__ pro1, int32_t* __restrict__ pro2, int32_t* __restrict__ nmergers,
    int32_t* __restrict__ mtype) {
    __shared__ Top2 sh[kRankThreads / kWave];
    __shared__ int count;
    __shared__ int any;
    const int64_t cap = 2 * G;
    for (int64_t c0 = 0; c0 < H; c0 += blockDim.x) {  // flags read a block-width at a time
        __syncthreads();  // every thread has read the previous chunk's any
        if (threadIdx.x == 0) any = 0;
        __syncthreads();
        if (c0 + threadIdx.x < H && flag[c0 + threadIdx.x]) any = 1;
        __syncthreads();
        if (!any) continue;  // block-uniform
        const int64_t c1 = c0 + blockDim.x < H ? c0 + blockDim.x : H;
        for (int64_t ih = c0; ih < c1; ++ih) {
            if (!flag[ih]) continue;  // block-uniform
            for (int64_t q = threadIdx.x; q < cap; q += blockDim.x) gset[q] = -1;
            __syncthreads();
            for (int64_t k = noff[ih] + threadIdx.x; k < noff[ih + 1]; k += blockDim.x) {
                int32_t oh = mhalo[k];
                int32_t j = mrest[k];
                while (oh >= 0) {
                    set_insert(gset, cap, oh);
                    oh = j >= 0 ? t.halo[j] : -1;
                    j = j >= 0 ? t.next[j] : -1;
                }
            }
            __syncthreads();
            rank_set(gset, cap, ih, new_mass, old_mass, sh, &count, pro1, pro2, nmergers, mtype);
        }
    }
}

struct ProgWs {
    unsigned long long* part;
    unsigned long long* range;
    unsigned long long* key;
    int32_t* head;
    int32_t* next;
    int32_t* halo;
    int32_t* mhalo;
    int32_t* mrest;
    int32_t* flag;
    int32_t* gset;
    int64_t cap;    // hash slots
    int64_t dcap;   // direct entries (keys 0 .. dcap - 1)
    int shift;
    size_t bytes;
};

ProgWs carve_prog(void* base, int64_t n_new_haloes, int64_t n_new, int64_t n_old_haloes,
                  int64_t n_old) {
    Carver c(base);
    ProgWs w;
    const int64_t no = std::max<int64_t>(n_old, 1), nn = std::max<int64_t>(n_new, 1);
    int lg = 4;
    while ((int64_t{1} << lg) < 2 * no) ++lg;
    w.cap = int64_t{1} << lg;
    w.shift = 64 - lg;
    w.dcap = 4 * (no + nn) + (int64_t{1} << 16);
    w.part = c.take<unsigned long long>(4 * kRangeBlocks);
    w.range = c.take<unsigned long long>(2);
    w.key = c.take<unsigned long long>(w.cap);
    w.head = c.take<int32_t>(std::max(w.cap, w.dcap));
    w.next = c.take<int32_t>(no);
    w.halo = c.take<int32_t>(no);
    w.mhalo = c.take<int32_t>(nn);
    w.mrest = c.take<int32_t>(nn);
    w.flag = c.take<int32_t>(std::max<int64_t>(n_new_haloes, 1));
    w.gset = c.take<int32_t>(2 * std::max<int64_t>(n_old_haloes, 1));
    w.bytes = c.used;
    return w;
}

}  // namespace

}  // namespace halma

using namespace halma;

extern "C" {

size_t halma_progenitors_workspace_size(int64_t n_new, int64_t n_new_members, int64_t n_old,
                                        int64_t n_old_members) {
    if (n_new < 0 || n_new_members < 0 || n_old < 0 || n_old_members < 0) return 0;
    return carve_prog(nullptr, n_new, n_new_members, n_old, n_old_members).bytes;
}

int halma_progenitors(int64_t n_new, int64_t n_new_members, const int64_t* new_offsets,
                      const int64_t* new_ids, const double* new_mass, int64_t n_old,
                      int64_t n_old_members, const int64_t* old_offsets, const int64_t* old_ids,
                      const double* old_mass, int32_t* pro1, int32_t* pro2, int32_t* n_mergers,
                      int32_t* merger_type, void* ws, size_t ws_bytes, void* stream) {
    if (n_new < 0 || n_new_members < 0 || n_old < 0 || n_old_members < 0)
        return set_error(HALMA_E_ARG, "bad argument");
    if (n_new == 0) return 0;
    if (!new_offsets || !new_mass || !pro1 || !pro2 || !n_mergers || !merger_type ||
        (n_new_members > 0 && !new_ids) || (n_old > 0 && (!old_offsets || !old_mass)) ||
        (n_old_members > 0 && !old_ids))
        return set_error(HALMA_E_ARG, "null pointer");
    if (n_new > INT32_MAX || n_old > INT32_MAX || n_new_members > INT32_MAX / 2 ||
        n_old_members > INT32_MAX / 2)
        return set_error(HALMA_E_LIMIT, "at most 2^31-1 haloes and 2^30-1 members");
    ProgWs w = carve_prog(ws, n_new, n_new_members, n_old, n_old_members);
    if (ws_bytes < w.bytes)
        return set_error(HALMA_E_WORKSPACE, "workspace %zu < required %zu", ws_bytes, w.bytes);
    hipStream_t st = static_cast<hipStream_t>(stream);
    const bool none = n_new_members == 0 || n_old_members == 0 || n_old == 0;
    if (!none) {
        // 1. key range, ID table of the old members
        k_range<<<kRangeBlocks, kThreads, 0, st>>>(old_ids, n_old_members, w.part);
        HALMA_LAUNCH_CHECK();
        k_range<<<kRangeBlocks, kThreads, 0, st>>>(new_ids, n_new_members,
                                                   w.part + 2 * kRangeBlocks);
        HALMA_LAUNCH_CHECK();
        k_range_final<<<1, kWave, 0, st>>>(w.part, 2 * kRangeBlocks, w.range);
        HALMA_LAUNCH_CHECK();
        unsigned long long r[2];
        HALMA_HIP(hipMemcpyAsync(r, w.range, sizeof(r), hipMemcpyDeviceToHost, st));
        HALMA_HIP(hipStreamSynchronize(st));
        if (r[1] - r[0] == kEmpty)
            return set_error(HALMA_E_LIMIT, "IDs span the whole int64 range");
        const unsigned long long span = r[1] - r[0];
        const int direct = span < static_cast<unsigned long long>(w.dcap);
        const int64_t entries = direct ? static_cast<int64_t>(span) + 1 : w.cap;
        if (!direct) {
            k_fill_u64<<<blocks_for(w.cap), kThreads, 0, st>>>(w.key, w.cap, kEmpty);
            HALMA_LAUNCH_CHECK();
        }
        k_fill_i32<<<blocks_for(entries), kThreads, 0, st>>>(w.head, entries, -1);
        HALMA_LAUNCH_CHECK();
        Table t{w.key, w.head, w.next, w.halo, w.shift,
                static_cast<unsigned long long>(w.cap - 1), direct};
        k_old_halo<<<static_cast<unsigned>(std::min<int64_t>(n_old, 65536)), kThreads, 0, st>>>(
            old_offsets, n_old, w.halo);
        HALMA_LAUNCH_CHECK();
        if (direct) {
            k_fill_i32<<<blocks_for(n_old_members), kThreads, 0, st>>>(w.next, n_old_members, -1);
            HALMA_LAUNCH_CHECK();
            k_direct_store<<<blocks_for(n_old_members), kThreads, 0, st>>>(n_old_members,
                                                                           old_ids, w.range,
                                                                           w.head);
            HALMA_LAUNCH_CHECK();
            k_direct_chain<<<blocks_for(n_old_members), kThreads, 0, st>>>(
                n_old_members, old_ids, w.range, w.head, w.next);
            HALMA_LAUNCH_CHECK();
        } else {
            k_table_build<<<blocks_for(n_old_members), kThreads, 0, st>>>(n_old_members,
                                                                          old_ids, w.range, t);
            HALMA_LAUNCH_CHECK();
        }
        // 2. every new member's chain of old members
        k_member_lookup<<<blocks_for(n_new_members), kThreads, 0, st>>>(
            n_new_members, new_ids, w.range, t, w.mhalo, w.mrest);
        HALMA_LAUNCH_CHECK();
    } else if (n_new_members > 0) {
        k_fill_i32<<<blocks_for(n_new_members), kThreads, 0, st>>>(w.mhalo, n_new_members, -1);
        HALMA_LAUNCH_CHECK();
        k_fill_i32<<<blocks_for(n_new_members), kThreads, 0, st>>>(w.mrest, n_new_members, -1);
        HALMA_LAUNCH_CHECK();
    }
    // 3. per new halo: the set of old haloes, its size, its two largest, the type
    Table t{w.key, w.head, w.next, w.halo, w.shift, static_cast<unsigned long long>(w.cap - 1), 0};
    const unsigned hb = static_cast<unsigned>(std::min<int64_t>(n_new, 65536));
    k_halo_rank<<<hb, kRankThreads, 0, st>>>(n_new, new_offsets, w.mhalo, w.mrest, t, new_mass,
                                             old_mass, w.flag, pro1, pro2, n_mergers,
                                             merger_type);
    HALMA_LAUNCH_CHECK();
    k_rank_fallback<<<1, kRankThreads, 0, st>>>(n_new, n_old, new_offsets, w.mhalo, w.mrest, t,
                                                w.flag, w.gset, new_mass, old_mass, pro1, pro2,
                                                n_mergers, merger_type);
    HALMA_LAUNCH_CHECK();
    return 0;
}

}  // extern "C"
