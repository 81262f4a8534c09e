// Progenitor matching (pyHALMA.py:1211-1266) for every new halo at once, as a sort join
// on particle IDs instead of one np.in1d per (new, old) halo pair.
//
// 1. IDs are mapped to unsigned keys (sign bit flipped) relative to the smallest ID of
//    both catalogues; the old members are sorted by key (stable radix passes over the low
//    word and, when the IDs span 2^32 or more, the high word).
// 2. Every new member finds its key's run in the sorted old keys (binary search); each
//    old member of the run gives a (new halo, old halo) pair, written at scanned offsets.
// 3. The pairs are sorted by (new halo, old halo) and made unique: the old haloes that
//    share at least one ID with a new halo, in ascending order per new halo.
// 4. One thread per new halo counts them (n_mergers), ranks them by old mass (equal
//    masses: the higher old index first, as np.flip of a stable argsort) and applies
//    the reference's classification: -1 when 1.2 * mass < the main progenitor's mass;
//    with two or more, 2 when 1/20 < second / first < 1/3 and 3 otherwise (the type 1
//    of second / first > 1/3 is overwritten by the following if / else, as there).
#include <algorithm>

#include "halma_common.h"
#include "radix.h"
#include "scan.h"

namespace halma {

namespace {

constexpr int kThreads = 256;

inline unsigned blocks_for(int64_t n) {
    return static_cast<unsigned>((n + kThreads - 1) / kThreads);
}

__device__ __forceinline__ unsigned long long id_key(int64_t v) {
    return static_cast<unsigned long long>(v) ^ 0x8000000000000000ull;
}

__global__ void k_range_init(unsigned long long* r) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        r[0] = ~0ull;
        r[1] = 0ull;
    }
}

// smallest and largest key of ids[0..n)
__global__ void k_range(const int64_t* __restrict__ ids, int64_t n, unsigned long long* r) {
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
    if (__lane_id() == 0) {
        atomicMin(r, lo);
        atomicMax(r + 1, hi);
    }
}

__global__ void k_rel_keys(const int64_t* __restrict__ ids, int64_t n,
                           const unsigned long long* __restrict__ r, uint32_t* __restrict__ lo,
                           uint32_t* __restrict__ hi) {
    const int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const unsigned long long v = id_key(ids[k]) - r[0];
    lo[k] = static_cast<uint32_t>(v & 0xffffffffull);
    hi[k] = static_cast<uint32_t>(v >> 32);
}

__global__ void k_gather_u32(const uint32_t* __restrict__ src, const int32_t* __restrict__ perm,
                             int64_t n, uint32_t* __restrict__ dst) {
    const int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k < n) dst[k] = src[perm[k]];
}

// the last g with off[g] <= k (off[0] = 0 <= k < off[G])
__device__ __forceinline__ int32_t group_of(const int64_t* off, int64_t G, int64_t k) {
    int64_t lo = 0, hi = G - 1;
    while (lo < hi) {
        const int64_t mid = (lo + hi + 1) >> 1;
        if (off[mid] <= k)
            lo = mid;
        else
            hi = mid - 1;
    }
    return static_cast<int32_t>(lo);
}

// sorted old keys (64-bit) and the old halo of each sorted position
__global__ void k_old_sorted(int64_t n, const int32_t* __restrict__ perm,
                             const uint32_t* __restrict__ lo, const uint32_t* __restrict__ hi,
                             const int64_t* __restrict__ off, int64_t G,
                             unsigned long long* __restrict__ skey, int32_t* __restrict__ soh) {
    const int64_t j = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const int32_t p = perm[j];
    skey[j] = (static_cast<unsigned long long>(hi[p]) << 32) | lo[p];
    soh[j] = group_of(off, G, p);
}

// the run of a new member's key among the sorted old keys
__global__ void k_match_count(int64_t nn, const uint32_t* __restrict__ lo,
                              const uint32_t* __restrict__ hi,
                              const unsigned long long* __restrict__ skey, int64_t no,
                              int64_t* __restrict__ lb, int32_t* __restrict__ cnt) {
    const int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= nn) return;
    const unsigned long long key = (static_cast<unsigned long long>(hi[k]) << 32) | lo[k];
    int64_t a = 0, b = no;  // first position with skey >= key
    while (a < b) {
        const int64_t mid = (a + b) >> 1;
        if (skey[mid] < key)
            a = mid + 1;
        else
            b = mid;
    }
    int64_t e = a;
    while (e < no && skey[e] == key) ++e;
    lb[k] = a;
    cnt[k] = static_cast<int32_t>(e - a);
}

__global__ void k_match_emit(int64_t nn, const int64_t* __restrict__ noff, int64_t H,
                             const int64_t* __restrict__ lb, const int32_t* __restrict__ cnt,
                             const int64_t* __restrict__ pos, const int32_t* __restrict__ soh,
                             int obits, int composite, uint32_t* __restrict__ pkey,
                             int32_t* __restrict__ pih, int32_t* __restrict__ poh) {
    const int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= nn || cnt[k] == 0) return;
    const int32_t ih = group_of(noff, H, k);
    for (int t = 0; t < cnt[k]; ++t) {
        const int64_t q = pos[k] + t;
        const int32_t oh = soh[lb[k] + t];
        pih[q] = ih;
        poh[q] = oh;
        pkey[q] = composite ? (static_cast<uint32_t>(ih) << obits) | static_cast<uint32_t>(oh)
                            : static_cast<uint32_t>(oh);
    }
}

__global__ void k_pair_heads(int64_t P, const int32_t* __restrict__ perm,
                             const int32_t* __restrict__ pih, const int32_t* __restrict__ poh,
                             int32_t* __restrict__ head) {
    const int64_t j = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (j >= P) return;
    int h = 1;
    if (j > 0) {
        const int32_t a = perm[j], b = perm[j - 1];
        h = (pih[a] != pih[b] || poh[a] != poh[b]) ? 1 : 0;
    }
    head[j] = h;
}

__global__ void k_pair_unique(int64_t P, const int32_t* __restrict__ perm,
                              const int32_t* __restrict__ head, const int64_t* __restrict__ uidx,
                              const int32_t* __restrict__ pih, const int32_t* __restrict__ poh,
                              int32_t* __restrict__ uih, int32_t* __restrict__ uoh) {
    const int64_t j = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (j >= P || !head[j]) return;
    const int32_t a = perm[j];
    uih[uidx[j]] = pih[a];
    uoh[uidx[j]] = poh[a];
}

// a ranks before b: larger mass, or equal mass and larger old index
__device__ __forceinline__ bool ranks_before(double ma, int32_t a, double mb, int32_t b) {
    return ma > mb || (ma == mb && a > b);
}

__global__ void k_classify(int64_t H, const int32_t* __restrict__ uih,
                           const int32_t* __restrict__ uoh, const int64_t* __restrict__ nu,
                           const double* __restrict__ new_mass,
                           const double* __restrict__ old_mass, int32_t* __restrict__ pro1,
                           int32_t* __restrict__ pro2, int32_t* __restrict__ nmergers,
                           int32_t* __restrict__ mtype) {
    const int64_t ih = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (ih >= H) return;
    const int64_t U = nu ? *nu : 0;
    int64_t a = 0, b = U;  // first unique pair of this new halo
    while (a < b) {
        const int64_t mid = (a + b) >> 1;
        if (uih[mid] < ih)
            a = mid + 1;
        else
            b = mid;
    }
    int n = 0;
    int32_t first = -1, second = -1;
    double m1 = 0.0, m2 = 0.0;
    for (int64_t j = a; j < U && uih[j] == ih; ++j) {
        const int32_t oh = uoh[j];
        const double m = old_mass[oh];
        ++n;
        if (first < 0 || ranks_before(m, oh, m1, first)) {
            second = first;
            m2 = m1;
            first = oh;
            m1 = m;
        } else if (second < 0 || ranks_before(m, oh, m2, second)) {
            second = oh;
            m2 = m;
        }
    }
    int32_t p1 = 0, p2 = 0, t = 0;
    if (n > 0) {
        p1 = first + 1;
        if (1.2 * new_mass[ih] < m1) {
            t = -1;  // the old halo breaking apart
        } else if (n > 1) {
            p2 = second + 1;
            const double frac = m2 / m1;
            t = (1.0 / 20.0 < frac && frac < 1.0 / 3.0) ? 2 : 3;
        }
    }
    pro1[ih] = p1;
    pro2[ih] = p2;
    nmergers[ih] = n;
    mtype[ih] = t;
}

struct ProgWs {
    unsigned long long* range;
    int64_t* totals;  // [0] pairs, [1] unique pairs
    uint32_t* olo;
    uint32_t* ohi;
    uint32_t* nlo;
    uint32_t* nhi;
    uint32_t* k2;
    uint32_t* k3;
    int32_t* o0;
    int32_t* o1;
    unsigned long long* skey;
    int32_t* soh;
    int64_t* lb;
    int32_t* cnt;
    int64_t* pos;
    uint32_t* pkey;
    int32_t* pih;
    int32_t* poh;
    int32_t* q0;
    int32_t* q1;
    int32_t* head;
    int64_t* uidx;
    int32_t* uih;
    int32_t* uoh;
    void* rws;
    void* sws;
    size_t bytes;
};

ProgWs carve_prog(void* base, int64_t n_new, int64_t n_old, int64_t max_pairs) {
    Carver c(base);
    const int64_t nn = std::max<int64_t>(n_new, 1), no = std::max<int64_t>(n_old, 1);
    const int64_t np = std::max<int64_t>(max_pairs, 1);
    const int64_t nmax = std::max(std::max(nn, no), np);
    ProgWs w;
    w.range = c.take<unsigned long long>(2);
    w.totals = c.take<int64_t>(4);
    w.olo = c.take<uint32_t>(no);
    w.ohi = c.take<uint32_t>(no);
    w.nlo = c.take<uint32_t>(nn);
    w.nhi = c.take<uint32_t>(nn);
    w.k2 = c.take<uint32_t>(nmax);
    w.k3 = c.take<uint32_t>(nmax);
    w.o0 = c.take<int32_t>(no);
    w.o1 = c.take<int32_t>(no);
    w.skey = c.take<unsigned long long>(no);
    w.soh = c.take<int32_t>(no);
    w.lb = c.take<int64_t>(nn);
    w.cnt = c.take<int32_t>(nn);
    w.pos = c.take<int64_t>(nn);
    w.pkey = c.take<uint32_t>(np);
    w.pih = c.take<int32_t>(np);
    w.poh = c.take<int32_t>(np);
    w.q0 = c.take<int32_t>(np);
    w.q1 = c.take<int32_t>(np);
    w.head = c.take<int32_t>(np);
    w.uidx = c.take<int64_t>(np);
    w.uih = c.take<int32_t>(np);
    w.uoh = c.take<int32_t>(np);
    w.rws = c.take<char>(radix_workspace_bytes(nmax));
    w.sws = c.take<char>(scan_workspace_bytes(nmax));
    w.bytes = c.used;
    return w;
}

int bits_for(uint64_t top) {  // bits needed for values 0..top (at least 1)
    int b = 1;
    while (b < 64 && (top >> b) != 0) ++b;
    return b;
}

__global__ void k_zero_i32(int32_t* a, int64_t n) {
    const int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k < n) a[k] = 0;
}

}  // namespace

}  // namespace halma

using namespace halma;

extern "C" {

size_t halma_progenitors_workspace_size(int64_t n_new_members, int64_t n_old_members,
                                        int64_t max_pairs) {
    if (n_new_members < 0 || n_old_members < 0 || max_pairs < 0) return 0;
    return carve_prog(nullptr, n_new_members, n_old_members, max_pairs).bytes;
}

int halma_progenitors(int64_t n_new, int64_t n_new_members, const int64_t* new_offsets,
                      const int64_t* new_ids, const double* new_mass, int64_t n_old,
                      int64_t n_old_members, const int64_t* old_offsets, const int64_t* old_ids,
                      const double* old_mass, int64_t max_pairs, int32_t* pro1, int32_t* pro2,
                      int32_t* n_mergers, int32_t* merger_type, int64_t* n_pairs, void* ws,
                      size_t ws_bytes, void* stream) {
    if (n_new < 0 || n_new_members < 0 || n_old < 0 || n_old_members < 0 || max_pairs < 0 ||
        !n_pairs)
        return set_error(HALMA_E_ARG, "bad argument");
    *n_pairs = 0;
    if (n_new == 0) return 0;
    if (!new_offsets || !new_mass || !pro1 || !pro2 || !n_mergers || !merger_type ||
        (n_new_members > 0 && !new_ids) || (n_old > 0 && (!old_offsets || !old_mass)) ||
        (n_old_members > 0 && !old_ids))
        return set_error(HALMA_E_ARG, "null pointer");
    if (n_new > INT32_MAX || n_old > INT32_MAX || n_new_members > INT32_MAX ||
        n_old_members > INT32_MAX)
        return set_error(HALMA_E_LIMIT, "at most 2^31-1 haloes and members per catalogue");
    ProgWs w = carve_prog(ws, n_new_members, n_old_members, max_pairs);
    if (ws_bytes < w.bytes)
        return set_error(HALMA_E_WORKSPACE, "workspace %zu < required %zu", ws_bytes, w.bytes);
    hipStream_t st = static_cast<hipStream_t>(stream);
    if (n_new_members == 0 || n_old_members == 0 || n_old == 0) {
        int32_t* outs[4] = {pro1, pro2, n_mergers, merger_type};
        for (int32_t* o : outs) {
            k_zero_i32<<<blocks_for(n_new), kThreads, 0, st>>>(o, n_new);
            HALMA_LAUNCH_CHECK();
        }
        return 0;
    }
    // 1. key range over both catalogues, then the old members sorted by key
    k_range_init<<<1, 1, 0, st>>>(w.range);
    HALMA_LAUNCH_CHECK();
    const unsigned rb = static_cast<unsigned>(std::min<int64_t>(1024, (std::max(n_new_members, n_old_members) + kThreads - 1) / kThreads));
    k_range<<<rb, kThreads, 0, st>>>(old_ids, n_old_members, w.range);
    HALMA_LAUNCH_CHECK();
    k_range<<<rb, kThreads, 0, st>>>(new_ids, n_new_members, w.range);
    HALMA_LAUNCH_CHECK();
    unsigned long long r[2];
    HALMA_HIP(hipMemcpyAsync(r, w.range, sizeof(r), hipMemcpyDeviceToHost, st));
    HALMA_HIP(hipStreamSynchronize(st));
    const int kbits = bits_for(r[1] - r[0]);
    k_rel_keys<<<blocks_for(n_old_members), kThreads, 0, st>>>(old_ids, n_old_members, w.range,
                                                              w.olo, w.ohi);
    HALMA_LAUNCH_CHECK();
    k_rel_keys<<<blocks_for(n_new_members), kThreads, 0, st>>>(new_ids, n_new_members, w.range,
                                                              w.nlo, w.nhi);
    HALMA_LAUNCH_CHECK();
    int rc = radix_sort_pairs(w.olo, nullptr, w.k2, w.o0, n_old_members, std::min(kbits, 32),
                              w.rws, st);
    if (rc) return rc;
    const int32_t* operm = w.o0;
    if (kbits > 32) {
        k_gather_u32<<<blocks_for(n_old_members), kThreads, 0, st>>>(w.ohi, w.o0, n_old_members,
                                                                     w.k3);
        HALMA_LAUNCH_CHECK();
        rc = radix_sort_pairs(w.k3, w.o0, w.k2, w.o1, n_old_members, kbits - 32, w.rws, st);
        if (rc) return rc;
        operm = w.o1;
    }
    k_old_sorted<<<blocks_for(n_old_members), kThreads, 0, st>>>(
        n_old_members, operm, w.olo, w.ohi, old_offsets, n_old, w.skey, w.soh);
    HALMA_LAUNCH_CHECK();
    // 2. (new halo, old halo) pairs of every shared ID
    k_match_count<<<blocks_for(n_new_members), kThreads, 0, st>>>(
        n_new_members, w.nlo, w.nhi, w.skey, n_old_members, w.lb, w.cnt);
    HALMA_LAUNCH_CHECK();
    rc = exclusive_scan_i32_to_i64(w.cnt, w.pos, n_new_members, w.totals, w.sws, st);
    if (rc) return rc;
    HALMA_HIP(hipMemcpyAsync(n_pairs, w.totals, sizeof(int64_t), hipMemcpyDeviceToHost, st));
    HALMA_HIP(hipStreamSynchronize(st));
    const int64_t P = *n_pairs;
    if (P > max_pairs)
        return set_error(HALMA_E_LIMIT, "%lld (new, old) ID matches exceed max_pairs = %lld",
                         static_cast<long long>(P), static_cast<long long>(max_pairs));
    if (P > INT32_MAX) return set_error(HALMA_E_LIMIT, "at most 2^31-1 ID matches");
    const int hbits = bits_for(static_cast<uint64_t>(n_new - 1));
    const int obits = bits_for(static_cast<uint64_t>(n_old - 1));
    const int composite = hbits + obits <= 32;
    const int32_t* pperm = w.q0;
    if (P > 0) {
        k_match_emit<<<blocks_for(n_new_members), kThreads, 0, st>>>(
            n_new_members, new_offsets, n_new, w.lb, w.cnt, w.pos, w.soh, obits, composite,
            w.pkey, w.pih, w.poh);
        HALMA_LAUNCH_CHECK();
        // 3. sorted by (new halo, old halo), then unique
        if (composite) {
            rc = radix_sort_pairs(w.pkey, nullptr, w.k2, w.q0, P, hbits + obits, w.rws, st);
            if (rc) return rc;
        } else {
            rc = radix_sort_pairs(w.pkey, nullptr, w.k2, w.q1, P, obits, w.rws, st);
            if (rc) return rc;
            k_gather_u32<<<blocks_for(P), kThreads, 0, st>>>(
                reinterpret_cast<const uint32_t*>(w.pih), w.q1, P, w.k3);
            HALMA_LAUNCH_CHECK();
            rc = radix_sort_pairs(w.k3, w.q1, w.k2, w.q0, P, hbits, w.rws, st);
            if (rc) return rc;
        }
        k_pair_heads<<<blocks_for(P), kThreads, 0, st>>>(P, pperm, w.pih, w.poh, w.head);
        HALMA_LAUNCH_CHECK();
        rc = exclusive_scan_i32_to_i64(w.head, w.uidx, P, w.totals + 1, w.sws, st);
        if (rc) return rc;
        k_pair_unique<<<blocks_for(P), kThreads, 0, st>>>(P, pperm, w.head, w.uidx, w.pih, w.poh,
                                                           w.uih, w.uoh);
        HALMA_LAUNCH_CHECK();
    }
    // 4. per new halo: count, rank, classify
    k_classify<<<blocks_for(n_new), kThreads, 0, st>>>(n_new, w.uih, w.uoh,
                                                       P > 0 ? w.totals + 1 : nullptr, new_mass,
                                                       old_mass, pro1, pro2, n_mergers,
                                                       merger_type);
    HALMA_LAUNCH_CHECK();
    return 0;
}

}  // extern "C"
