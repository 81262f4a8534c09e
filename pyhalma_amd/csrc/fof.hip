// Friends-of-friends labels on the GPU: connected components of the graph with an edge
// wherever ((dx*dx + dy*dy) + dz*dz) <= ll*ll in FP64 (fof.py:18-79 through scipy's
// KDTree; inclusive, non-periodic), canonical label = minimum particle index.
//
// 1. bounding box (two-pass reduction);
// 2. sub-cells of side s = ll/sqrt(3) * (1 - 1e-6): two particles in one sub-cell are
//    within ll under the FP64 test (|dx|,|dy|,|dz| <= s(1+1e-9) after rounding of the
//    cell index), so a sub-cell is one vertex of the cell graph and needs no pair test;
//    every link spans at most 2 sub-cells per axis (ll/s < 1.733);
// 3. occupied sub-cells go into an open-addressing hash table (64-bit keys, vector
//    atomics); each particle records its cell's slot, and a new cell sets its bit in a
//    hashed occupancy bitmap;
// 4. a stable radix sort of (slot, particle) lists the particles cell by cell (cells
//    numbered in slot order, particles ascending within a cell); coordinates are
//    gathered into that order;
// 5. links, in two rounds (the 13 neighbours at Chebyshev distance 1, then the 49 at
//    distance 2): one lane per (cell, half-stencil offset) skips empty neighbours (a
//    clear occupancy bit, else the table) and pairs already joined, tests small pairs
//    of cells itself, and lists larger pairs for one wavefront each (64 lanes over the
//    larger cell, early exit); linked cells are joined in a lock-free union-find over
//    cells (the larger root index is hooked under the smaller, so parents only decrease
//    and no cycle can form; finds halve paths);
// 6. roots are read back with atomic reads (see k_roots), each component's minimum
//    particle index is taken over its cells, and every particle gets it as label.
#include <math.h>

#include <algorithm>

#include "halma_common.h"
#include "radix.h"
#include "scan.h"
#include "staging.h"

namespace halma {

namespace {

constexpr uint64_t kEmpty = ~0ull;
constexpr int kCellBits = 21;
constexpr double kCellLimit = 2097152.0;  // 2^21 cells per axis
constexpr int kBoxBlocks = 1024;

// Table slot of a cell key: the (x, y) column is hashed and z is added, so the cells of
// one column sit in consecutive slots and the 5 z-neighbours of a stencil column are one
// or two memory lines apart.
__device__ __forceinline__ uint64_t mix64(uint64_t k);
__device__ __forceinline__ uint64_t slot_of(uint64_t key, uint64_t mask) {
    return (mix64(key >> kCellBits) + (key & ((1ull << kCellBits) - 1))) & mask;
}

// 16-byte table entry: cell key (kEmpty when free), particle count, cell number
struct __attribute__((aligned(16))) Entry {
    uint64_t key;
    int32_t cnt;
    int32_t cell;
};

__device__ __forceinline__ uint64_t mix64(uint64_t k) {
    k ^= k >> 30;
    k *= 0xbf58476d1ce4e5b9ull;
    k ^= k >> 27;
    k *= 0x94d049bb133111ebull;
    k ^= k >> 31;
    return k;
}

template <typename T>
__device__ __forceinline__ int ld_relaxed(const T* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ int uf_find(int* P, int x) {
    int p = __hip_atomic_load(P + x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    while (p != x) {
        const int g = __hip_atomic_load(P + p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (g == p) return p;
        __hip_atomic_store(P + x, g, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        x = g;
        p = __hip_atomic_load(P + x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    return x;
}

__device__ void uf_union(int* P, int a, int b) {
    for (;;) {
        a = uf_find(P, a);
        b = uf_find(P, b);
        if (a == b) return;
        if (a < b) {
            const int t = a;
            a = b;
            b = t;
        }
        // hook root a (> b) under b
        int expected = a;
        if (__hip_atomic_compare_exchange_strong(P + a, &expected, b, __ATOMIC_RELAXED,
                                                 __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
            return;
        a = expected;  // a was hooked meanwhile: continue from its new parent
    }
}

template <typename T>
__global__ __launch_bounds__(256) void k_bbox_partial(const T* __restrict__ x,
                                                      const T* __restrict__ y,
                                                      const T* __restrict__ z, int64_t n,
                                                      double* __restrict__ part) {
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const double v[3] = {static_cast<double>(x[i]), static_cast<double>(y[i]),
                             static_cast<double>(z[i])};
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            lo[d] = fmin(lo[d], v[d]);
            hi[d] = fmax(hi[d], v[d]);
        }
    }
    __shared__ double sh[4][6];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        for (int s = 32; s >= 1; s >>= 1) {
            lo[d] = fmin(lo[d], __shfl_down(lo[d], s, kWave));
            hi[d] = fmax(hi[d], __shfl_down(hi[d], s, kWave));
        }
    }
    const int w = threadIdx.x / kWave;
    if (__lane_id() == 0)
        for (int d = 0; d < 3; ++d) {
            sh[w][d] = lo[d];
            sh[w][3 + d] = hi[d];
        }
    __syncthreads();
    if (threadIdx.x < 6) {
        double v = sh[0][threadIdx.x];
        for (int k = 1; k < 4; ++k)
            v = threadIdx.x < 3 ? fmin(v, sh[k][threadIdx.x]) : fmax(v, sh[k][threadIdx.x]);
        part[blockIdx.x * 6 + threadIdx.x] = v;
    }
}

__global__ __launch_bounds__(256) void k_bbox_final(const double* __restrict__ part, int nb,
                                                    double* __restrict__ box) {
    __shared__ double sh[4][6];
    double v[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
    for (int b = threadIdx.x; b < nb; b += blockDim.x)
        for (int d = 0; d < 6; ++d)
            v[d] = d < 3 ? fmin(v[d], part[b * 6 + d]) : fmax(v[d], part[b * 6 + d]);
    for (int d = 0; d < 6; ++d)
        for (int s = 32; s >= 1; s >>= 1) {
            const double o = __shfl_down(v[d], s, kWave);
            v[d] = d < 3 ? fmin(v[d], o) : fmax(v[d], o);
        }
    const int w = threadIdx.x / kWave;
    if (__lane_id() == 0)
        for (int d = 0; d < 6; ++d) sh[w][d] = v[d];
    __syncthreads();
    if (threadIdx.x < 6) {
        double r = sh[0][threadIdx.x];
        for (int k = 1; k < 4; ++k)
            r = threadIdx.x < 3 ? fmin(r, sh[k][threadIdx.x]) : fmax(r, sh[k][threadIdx.x]);
        box[threadIdx.x] = r;
    }
}

__global__ void k_clear_table(Entry* t, int64_t n) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) {
        t[i].key = kEmpty;
        t[i].cnt = 0;
        t[i].cell = INT32_MAX;
    }
}

__global__ void k_fill_i32n(int* p, int64_t n, int v) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// (x, y, z, 0) in one vector: float4 for FP32 coordinates, double4 for FP64
template <typename T>
struct Packed;
template <>
struct Packed<float> {
    typedef float4 type;
    __device__ static float4 make(float x, float y, float z) { return make_float4(x, y, z, 0.f); }
};
template <>
struct Packed<double> {
    typedef double4 type;
    __device__ static double4 make(double x, double y, double z) {
        return make_double4(x, y, z, 0.0);
    }
};

// Binning: every particle finds (or inserts) its sub-cell in the hash table and records
// the slot; a stable radix sort of (slot, particle index) then lists the particles cell
// by cell, ascending within a cell (so a cell's first particle is its minimum index).
// A newly inserted cell also sets its bit in a hashed occupancy bitmap, which lets the
// link phase skip the hash lookups of empty neighbour cells.
// Occupancy bit of a cell: the low bits of its Morton code, so that the neighbourhood of
// a cell falls on a few bitmap lines (cells far apart may share a bit: a false positive
// only costs a table lookup).
__device__ __forceinline__ uint64_t spread3(uint64_t v) {  // 21 bits -> every third bit
    v &= 0x1fffffull;
    v = (v | (v << 32)) & 0x1f00000000ffffull;
    v = (v | (v << 16)) & 0x1f0000ff0000ffull;
    v = (v | (v << 8)) & 0x100f00f00f00f00full;
    v = (v | (v << 4)) & 0x10c30c30c30c30c3ull;
    v = (v | (v << 2)) & 0x1249249249249249ull;
    return v;
}

__device__ __forceinline__ uint64_t occ_bit(uint64_t key, uint64_t bmask) {
    const uint64_t m = (1ull << kCellBits) - 1;
    return (spread3(key >> (2 * kCellBits)) << 2 | spread3((key >> kCellBits) & m) << 1 |
            spread3(key & m)) & bmask;
}

template <typename T>
__global__ __launch_bounds__(256) void k_fof_keys(
    const T* __restrict__ x, const T* __restrict__ y, const T* __restrict__ z, int64_t n,
    const double* __restrict__ box, double inv_s, Entry* __restrict__ tab, uint64_t mask,
    uint32_t* __restrict__ skey, uint32_t* __restrict__ bits, uint64_t bmask,
    int* __restrict__ err, typename Packed<T>::type* __restrict__ pk) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const T xi = x[i], yi = y[i], zi = z[i];
    // packed copy for the cell-order gather (one line per particle instead of three)
    pk[i] = Packed<T>::make(xi, yi, zi);
    const double q[3] = {(static_cast<double>(xi) - box[0]) * inv_s,
                         (static_cast<double>(yi) - box[1]) * inv_s,
                         (static_cast<double>(zi) - box[2]) * inv_s};
    uint64_t key = 0;
    bool bad = false;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        bad |= !(q[d] >= 0.0 && q[d] < kCellLimit);  // also catches NaN
        const uint64_t c = bad ? 0 : static_cast<uint64_t>(floor(q[d]));
        key = (key << kCellBits) | c;
    }
    if (bad) {
        atomicOr(err, 1);
        key = 0;
    }
    uint64_t h = slot_of(key, mask);
    for (uint64_t probe = 0;; ++probe) {
        // a plain load: a slot's key never changes once set, so a stale value can only
        // be a stale "empty", which the compare-and-swap below corrects
        const uint64_t k = tab[h].key;
        if (k == key) break;
        if (k == kEmpty) {
            uint64_t expected = kEmpty;
            if (__hip_atomic_compare_exchange_strong(&tab[h].key, &expected, key,
                                                     __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                     __HIP_MEMORY_SCOPE_AGENT)) {
                const uint64_t b = occ_bit(key, bmask);  // this thread created the cell
                atomicOr(bits + (b >> 5), 1u << (b & 31));
                break;
            }
            if (expected == key) break;
        }
        if (probe > mask) {  // table full (cannot happen with capacity >= 2n)
            atomicOr(err, 2);
            h = 0;
            break;
        }
        h = (h + 1) & mask;
    }
    skey[i] = static_cast<uint32_t>(h);
}

__global__ void k_run_heads(const uint32_t* __restrict__ s, int64_t n, int* __restrict__ head) {
    const int64_t j = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (j < n) head[j] = (j == 0 || s[j] != s[j - 1]) ? 1 : 0;
}

// per sorted position j: its cell c (cells numbered in slot order), the particle's
// coordinates in cell order, and, at the first particle of a cell, the cell's records
template <typename T>
__global__ void k_cells(const uint32_t* __restrict__ skey, const int32_t* __restrict__ perm,
                        const int64_t* __restrict__ before, const int* __restrict__ head,
                        int64_t n, Entry* __restrict__ tab,
                        const typename Packed<T>::type* __restrict__ pk, T* __restrict__ sx,
                        T* __restrict__ sy, T* __restrict__ sz, int* __restrict__ pcell,
                        int* __restrict__ cell_start, uint64_t* __restrict__ cell_key,
                        int* __restrict__ cell_min, int* __restrict__ parent,
                        int* __restrict__ comp_min) {
    const int64_t j = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const int hd = head[j];
    const int c = static_cast<int>(before[j] + hd - 1);
    const int32_t i = perm[j];
    const typename Packed<T>::type p = pk[i];
    sx[j] = p.x;
    sy[j] = p.y;
    sz[j] = p.z;
    pcell[i] = c;
    if (hd) {
        const uint32_t slot = skey[j];
        tab[slot].cell = c;
        cell_start[c] = static_cast<int>(j);
        cell_key[c] = tab[slot].key;
        cell_min[c] = i;  // stable sort: the cell's smallest particle index comes first
        parent[c] = c;
        comp_min[c] = INT32_MAX;
    }
    if (j == n - 1) cell_start[c + 1] = static_cast<int>(n);
}

// cell number of a key, or -1
__device__ __forceinline__ int table_find(const Entry* __restrict__ tab, uint64_t mask,
                                          uint64_t key) {
    uint64_t h = slot_of(key, mask);
    for (uint64_t probe = 0; probe <= mask; ++probe) {
        const uint64_t k = tab[h].key;
        if (k == key) return tab[h].cell;
        if (k == kEmpty) return -1;
        h = (h + 1) & mask;
    }
    return -1;
}

// Read-only find with plain loads, for the "already joined?" test before a pair search.
// Parent pointers only ever move to ancestors (hooks attach roots, halving skips a
// level), so every node reached here, even through a stale value, is an ancestor of x
// in the current forest: equal results prove the two cells are joined, and a stale
// "different" only costs a pair search.
__device__ __forceinline__ int find_ro(const int* P, int x) {
    for (int p = P[x]; p != x; p = P[x]) x = p;
    return x;
}

// The 62 half-stencil offsets (dx, dy, dz) of a cell: the 13 at Chebyshev distance 1
// first, then the 49 at distance 2, each group in (dz, dy, dx) order.
__constant__ signed char kNbr[62][3] = {
    {1, 0, 0}, {-1, 1, 0}, {0, 1, 0}, {1, 1, 0}, {-1, -1, 1}, {0, -1, 1}, {1, -1, 1},
    {-1, 0, 1}, {0, 0, 1}, {1, 0, 1}, {-1, 1, 1}, {0, 1, 1}, {1, 1, 1},
    {2, 0, 0}, {-2, 1, 0}, {2, 1, 0}, {-2, 2, 0}, {-1, 2, 0}, {0, 2, 0}, {1, 2, 0},
    {2, 2, 0}, {-2, -2, 1}, {-1, -2, 1}, {0, -2, 1}, {1, -2, 1}, {2, -2, 1}, {-2, -1, 1},
    {2, -1, 1}, {-2, 0, 1}, {2, 0, 1}, {-2, 1, 1}, {2, 1, 1}, {-2, 2, 1}, {-1, 2, 1},
    {0, 2, 1}, {1, 2, 1}, {2, 2, 1}, {-2, -2, 2}, {-1, -2, 2}, {0, -2, 2}, {1, -2, 2},
    {2, -2, 2}, {-2, -1, 2}, {-1, -1, 2}, {0, -1, 2}, {1, -1, 2}, {2, -1, 2}, {-2, 0, 2},
    {-1, 0, 2}, {0, 0, 2}, {1, 0, 2}, {2, 0, 2}, {-2, 1, 2}, {-1, 1, 2}, {0, 1, 2},
    {1, 1, 2}, {2, 1, 2}, {-2, 2, 2}, {-1, 2, 2}, {0, 2, 2}, {1, 2, 2}, {2, 2, 2}};
constexpr int kNear = 13;
constexpr int kLaneTest = 64;  // pair tests a lane does alone

template <typename T>
__device__ __forceinline__ bool within(const T* sx, const T* sy, const T* sz, int a, int b,
                                       double ll2) {
    const double dx = static_cast<double>(sx[a]) - static_cast<double>(sx[b]);
    const double dy = static_cast<double>(sy[a]) - static_cast<double>(sy[b]);
    const double dz = static_cast<double>(sz[a]) - static_cast<double>(sz[b]);
    return (dx * dx + dy * dy) + dz * dz <= ll2;
}

// Every cell pointed straight at its root (between the candidate passes, when no union
// runs): the read-only root checks of the next pass then take one or two steps instead
// of walking the chains that hooking by index builds in dense regions.  Concurrent
// writers store each cell's own root, so a reader sees either its old parent or the root.
__global__ void k_flatten(const int64_t* __restrict__ totals, int* __restrict__ parent) {
    const int64_t ncell = totals[0] & 0xffffffff;
    for (int64_t c = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; c < ncell;
         c += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        int r = static_cast<int>(c);
        for (int p = parent[r]; p != r; p = parent[r]) r = p;
        parent[c] = r;
    }
}

// Link candidates, one lane per (cell, neighbour offset) of one distance class (NEAR:
// the 13 at Chebyshev distance 1; otherwise the 49 at distance 2, launched after every
// distance-1 link is in the forest, so that dense regions are recognised as joined and
// need no search).  A lane skips empty neighbours (occupancy bit, then the table) and
// pairs already joined, tests small pairs of cells itself and hooks them when linked,
// and hands larger pairs to k_fof_big through a list (one atomic per wave); a full list
// makes the lane test the pair itself.
template <typename T, bool NEAR>
__global__ __launch_bounds__(256) void k_fof_cand(
    const Entry* __restrict__ tab, uint64_t mask, const uint32_t* __restrict__ bits,
    uint64_t bmask, const int* __restrict__ cell_start,
    const uint64_t* __restrict__ cell_key, const T* __restrict__ sx, const T* __restrict__ sy,
    const T* __restrict__ sz, const int64_t* __restrict__ totals, double ll2,
    int* __restrict__ parent, int2* __restrict__ big, int64_t big_cap,
    unsigned long long* __restrict__ n_big) {
    constexpr int K = NEAR ? kNear : 62 - kNear;
    constexpr int K0 = NEAR ? 0 : kNear;
    const int lane = __lane_id();
    const int64_t total = (totals[0] & 0xffffffff) * K;
    const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
    const uint64_t cmask = (1ull << kCellBits) - 1;
    // the wave iterates uniformly (lanes past the end idle) so that ballots are whole
    for (int64_t i0 = static_cast<int64_t>(blockIdx.x) * blockDim.x + (threadIdx.x & ~(kWave - 1));
         i0 < total; i0 += stride) {
        const int64_t idx = i0 + lane;
        bool push = false;
        int c = 0, d = -1;
        if (idx < total) {
            c = static_cast<int>(idx / K);
            const int k = K0 + static_cast<int>(idx % K);
            const uint64_t key = cell_key[c];
            const int64_t nx = static_cast<int64_t>((key >> (2 * kCellBits)) & cmask) + kNbr[k][0];
            const int64_t ny = static_cast<int64_t>((key >> kCellBits) & cmask) + kNbr[k][1];
            const int64_t nz = static_cast<int64_t>(key & cmask) + kNbr[k][2];
            if (nx >= 0 && ny >= 0 && nz >= 0 && nx <= static_cast<int64_t>(cmask) &&
                ny <= static_cast<int64_t>(cmask) && nz <= static_cast<int64_t>(cmask)) {
                const uint64_t nkey = (static_cast<uint64_t>(nx) << (2 * kCellBits)) |
                                      (static_cast<uint64_t>(ny) << kCellBits) |
                                      static_cast<uint64_t>(nz);
                const uint64_t b = occ_bit(nkey, bmask);
                if ((bits[b >> 5] >> (b & 31)) & 1u) d = table_find(tab, mask, nkey);
            }
            if (d >= 0 && find_ro(parent, c) != find_ro(parent, d)) {
                const int sc = cell_start[c], nc = cell_start[c + 1] - sc;
                const int sd = cell_start[d], nd = cell_start[d + 1] - sd;
                if (static_cast<int64_t>(nc) * nd <= kLaneTest) {
                    bool linked = false;
                    for (int a = 0; a < nc && !linked; ++a)
                        for (int bb = 0; bb < nd; ++bb)
                            if (within(sx, sy, sz, sc + a, sd + bb, ll2)) {
                                linked = true;
                                break;
                            }
                    if (linked) uf_union(parent, c, d);
                } else {
                    push = true;
                }
            }
        }
        const uint64_t pm = __ballot(push);
        if (pm) {
            unsigned long long base = 0;
            if (lane == 0) base = atomicAdd(n_big, static_cast<unsigned long long>(__popcll(pm)));
            base = __shfl(base, 0, kWave);
            if (push) {
                const int64_t at = static_cast<int64_t>(base + __popcll(pm & ((1ull << lane) - 1ull)));
                if (at < big_cap) {
                    big[at] = make_int2(c, d);
                } else {  // list full: test here (correct, slower)
                    const int sc = cell_start[c], nc = cell_start[c + 1] - sc;
                    const int sd = cell_start[d], nd = cell_start[d + 1] - sd;
                    bool linked = false;
                    for (int a = 0; a < nc && !linked; ++a)
                        for (int bb = 0; bb < nd; ++bb)
                            if (within(sx, sy, sz, sc + a, sd + bb, ll2)) {
                                linked = true;
                                break;
                            }
                    if (linked) uf_union(parent, c, d);
                }
            }
        }
    }
}

// Large pairs of cells, one wavefront each: skipped if joined meanwhile, else lanes run
// over the larger cell four blocks of 64 at a time (loads in flight together) and loop
// over the smaller cell, stopping at the first pair within ll.
template <typename T>
__global__ __launch_bounds__(256) void k_fof_big(
    const int2* __restrict__ big, const unsigned long long* __restrict__ n_big,
    int64_t big_cap, const int* __restrict__ cell_start, const T* __restrict__ sx,
    const T* __restrict__ sy, const T* __restrict__ sz, double ll2, int* __restrict__ parent) {
    constexpr int U = 4;
    const int lane = __lane_id();
    const int64_t nb = static_cast<int64_t>(n_big[0]) < big_cap
                           ? static_cast<int64_t>(n_big[0]) : big_cap;
    const int64_t wave0 = (static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x) / kWave;
    const int64_t nwaves = static_cast<int64_t>(gridDim.x) * blockDim.x / kWave;
    for (int64_t e = wave0; e < nb; e += nwaves) {
        const int2 pr = big[e];
        const int c = pr.x, d = pr.y;
        if (find_ro(parent, c) == find_ro(parent, d)) continue;
        const int sc = cell_start[c], nc = cell_start[c + 1] - sc;
        const int sd = cell_start[d], nd = cell_start[d + 1] - sd;
        const bool c_small = nc <= nd;
        const int sA = c_small ? sc : sd, nA = c_small ? nc : nd;
        const int sB = c_small ? sd : sc, nB = c_small ? nd : nc;
        bool linked = false;
        for (int b0 = 0; b0 < nB && !linked; b0 += U * kWave) {
            double xb[U], yb[U], zb[U];
            bool valid[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int b = b0 + u * kWave + lane;
                valid[u] = b < nB;
                const int bb = valid[u] ? sB + b : sB;
                xb[u] = static_cast<double>(sx[bb]);
                yb[u] = static_cast<double>(sy[bb]);
                zb[u] = static_cast<double>(sz[bb]);
            }
            for (int a = 0; a < nA; ++a) {
                const double xa = static_cast<double>(sx[sA + a]);
                const double ya = static_cast<double>(sy[sA + a]);
                const double za = static_cast<double>(sz[sA + a]);
                bool hit = false;
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const double dx = xa - xb[u];
                    const double dy = ya - yb[u];
                    const double dz = za - zb[u];
                    hit |= valid[u] && (dx * dx + dy * dy) + dz * dz <= ll2;
                }
                if (__any(hit)) {
                    linked = true;
                    break;
                }
            }
        }
        if (linked && lane == 0) uf_union(parent, c, d);
    }
}

// Roots after the link phase.  Parents are read with read-modify-write atomics (a
// fetch-or of 0): on a multi-XCD device a relaxed load may return a value cached before
// another XCD's hook, and a stale "I am a root" answer would split a component, while
// an atomic read sees the value at the coherence point.  Results go to root[] (parent is
// not modified here), and every root is re-checked the same way: a non-root target is
// reported through err (bit 4) instead of producing a wrong partition.
__device__ __forceinline__ int coherent_read(int* p) {
    return __hip_atomic_fetch_or(p, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__global__ void k_roots(const int64_t* __restrict__ totals, int* __restrict__ parent,
                        int* __restrict__ root) {
    const int64_t n_cells = totals[0] & 0xffffffff;
    for (int64_t c = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; c < n_cells;
         c += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        int x = static_cast<int>(c);
        for (int p = coherent_read(parent + x); p != x; p = coherent_read(parent + x)) x = p;
        root[c] = x;
    }
}

// also folds every cell's minimum particle index into its component's minimum
__global__ void k_check_roots(const int64_t* __restrict__ totals, int* __restrict__ parent,
                              const int* __restrict__ root, const int* __restrict__ cell_min,
                              int* __restrict__ comp_min, int* __restrict__ err) {
    const int64_t n_cells = totals[0] & 0xffffffff;
    for (int64_t c = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; c < n_cells;
         c += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const int r = root[c];
        if (coherent_read(parent + r) != r) atomicOr(err, 4);
        atomicMin(comp_min + r, cell_min[c]);
    }
}

__global__ void k_labels(int64_t n, const int* __restrict__ pcell, const int* __restrict__ root,
                         const int* __restrict__ comp_min, int* __restrict__ labels) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) labels[i] = comp_min[root[pcell[i]]];
}

struct FofWs {
    double* box_part;
    double* box;
    Entry* tab;
    uint32_t* bits;   // hashed occupancy bitmap of the cells
    int64_t* totals;
    void* scan_ws;
    void* sort_ws;
    uint32_t* skey;   // slot of each particle, then sorted
    uint32_t* skey_sorted;
    int32_t* perm;    // particle of each sorted position
    int* head;
    int64_t* before;  // exclusive scan of head
    int* pcell;
    int* cell_start;  // n_cells + 1 entries
    uint64_t* cell_key;
    int* parent;
    int* root;
    int* comp_min;
    int* cell_min;
    void* sx;
    void* sy;
    void* sz;
    void* pk;         // (x, y, z, 0) per particle in input order
    int2* big;        // large pairs of cells left for k_fof_big
    unsigned long long* n_big;
    int64_t big_cap;
    int* err;
    size_t bytes;
    int64_t cap;
    int64_t nbits;
    int slot_bits;
};

int64_t table_capacity(int64_t n) {
    int64_t cap = 1024;
    while (cap < 2 * n) cap <<= 1;
    return cap;
}

// occupancy bits: the power of two >= 16 n (>= 2^15, <= 2^31), so that at most one
// empty neighbour in 16 passes the bit test
int64_t bitmap_bits(int64_t n) {
    int64_t b = int64_t{1} << 15;
    while (b < 16 * n && b < (int64_t{1} << 31)) b <<= 1;
    return b;
}

FofWs carve_fof(void* base, int64_t n, size_t elt) {
    Carver c(base);
    FofWs w;
    const int64_t nn = n > 0 ? n : 1;
    w.cap = table_capacity(n);
    w.slot_bits = 0;
    while ((int64_t{1} << w.slot_bits) < w.cap) ++w.slot_bits;
    w.nbits = bitmap_bits(n);
    w.box_part = c.take<double>(kBoxBlocks * 6);
    w.box = c.take<double>(8);
    w.tab = c.take<Entry>(w.cap);
    w.bits = c.take<uint32_t>(w.nbits / 32);
    w.totals = c.take<int64_t>(4);
    w.scan_ws = c.take<char>(scan_workspace_bytes(nn));
    w.sort_ws = c.take<char>(radix_workspace_bytes(nn));
    w.skey = c.take<uint32_t>(nn);
    w.skey_sorted = c.take<uint32_t>(nn);
    w.perm = c.take<int32_t>(nn);
    w.head = c.take<int>(nn);
    w.before = c.take<int64_t>(nn);
    w.pcell = c.take<int>(nn);
    w.cell_start = c.take<int>(nn + 1);
    w.cell_key = c.take<uint64_t>(nn);
    w.parent = c.take<int>(nn);
    w.root = c.take<int>(nn);
    w.comp_min = c.take<int>(nn);
    w.cell_min = c.take<int>(nn);
    w.sx = c.take<char>(nn * elt);
    w.sy = c.take<char>(nn * elt);
    w.sz = c.take<char>(nn * elt);
    w.pk = c.take<char>(nn * 4 * elt);
    w.big_cap = 2 * nn;
    w.big = c.take<int2>(w.big_cap);
    w.n_big = c.take<unsigned long long>(4);
    w.err = c.take<int>(64);
    w.bytes = c.used;
    return w;
}

__global__ void k_clear_words(uint32_t* __restrict__ p, int64_t n) {
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += static_cast<int64_t>(gridDim.x) * blockDim.x)
        p[i] = 0u;
}

inline unsigned blocks_for(int64_t n, int t = 256) {
    return static_cast<unsigned>((n + t - 1) / t);
}

template <typename T>
int fof_run(const T* x, const T* y, const T* z, int64_t n, double ll, int* labels,
            const FofWs& w, hipStream_t st) {
    const int grid_cells = device_cu_count() > 0 ? device_cu_count() * 8 : 2048;
    k_fill_i32n<<<1, 64, 0, st>>>(w.err, 1, 0);
    HALMA_LAUNCH_CHECK();
    k_bbox_partial<T><<<kBoxBlocks, 256, 0, st>>>(x, y, z, n, w.box_part);
    HALMA_LAUNCH_CHECK();
    k_bbox_final<<<1, 256, 0, st>>>(w.box_part, kBoxBlocks, w.box);
    HALMA_LAUNCH_CHECK();
    k_clear_table<<<blocks_for(w.cap), 256, 0, st>>>(w.tab, w.cap);
    HALMA_LAUNCH_CHECK();
    k_clear_words<<<1024, 256, 0, st>>>(w.bits, w.nbits / 32);
    HALMA_LAUNCH_CHECK();
    // sub-cell side: ll/sqrt(3) with a 1e-6 relative margin (see the file header)
    const double side = ll / sqrt(3.0) * (1.0 - 1e-6);
    const double inv_s = 1.0 / side;
    const uint64_t bmask = static_cast<uint64_t>(w.nbits - 1);
    auto* pk = static_cast<typename Packed<T>::type*>(w.pk);
    k_fof_keys<T><<<blocks_for(n), 256, 0, st>>>(x, y, z, n, w.box, inv_s, w.tab,
                                                 static_cast<uint64_t>(w.cap - 1), w.skey,
                                                 w.bits, bmask, w.err, pk);
    HALMA_LAUNCH_CHECK();
    int rc = radix_sort_pairs(w.skey, nullptr, w.skey_sorted, w.perm, n, w.slot_bits,
                              w.sort_ws, st);
    if (rc) return rc;
    k_run_heads<<<blocks_for(n), 256, 0, st>>>(w.skey_sorted, n, w.head);
    HALMA_LAUNCH_CHECK();
    rc = exclusive_scan_i32_to_i64(w.head, w.before, n, w.totals, w.scan_ws, st);
    if (rc) return rc;
    k_cells<T><<<blocks_for(n), 256, 0, st>>>(
        w.skey_sorted, w.perm, w.before, w.head, n, w.tab, pk, static_cast<T*>(w.sx),
        static_cast<T*>(w.sy), static_cast<T*>(w.sz), w.pcell, w.cell_start, w.cell_key,
        w.cell_min, w.parent, w.comp_min);
    HALMA_LAUNCH_CHECK();
    k_clear_words<<<1, 64, 0, st>>>(reinterpret_cast<uint32_t*>(w.n_big), 4);
    HALMA_LAUNCH_CHECK();
    const T* sx = static_cast<const T*>(w.sx);
    const T* sy = static_cast<const T*>(w.sy);
    const T* sz = static_cast<const T*>(w.sz);
    const unsigned cand_grid = static_cast<unsigned>(std::max(1, device_cu_count()) * 32);
    k_fof_cand<T, true><<<cand_grid, 256, 0, st>>>(
        w.tab, static_cast<uint64_t>(w.cap - 1), w.bits, bmask, w.cell_start, w.cell_key, sx,
        sy, sz, w.totals, ll * ll, w.parent, w.big, w.big_cap, w.n_big);
    HALMA_LAUNCH_CHECK();
    k_fof_big<T><<<grid_cells, 256, 0, st>>>(w.big, w.n_big, w.big_cap, w.cell_start, sx, sy,
                                             sz, ll * ll, w.parent);
    HALMA_LAUNCH_CHECK();
    k_flatten<<<grid_cells, 256, 0, st>>>(w.totals, w.parent);
    HALMA_LAUNCH_CHECK();
    k_fof_cand<T, false><<<cand_grid, 256, 0, st>>>(
        w.tab, static_cast<uint64_t>(w.cap - 1), w.bits, bmask, w.cell_start, w.cell_key, sx,
        sy, sz, w.totals, ll * ll, w.parent, w.big, w.big_cap, w.n_big + 1);
    HALMA_LAUNCH_CHECK();
    k_fof_big<T><<<grid_cells, 256, 0, st>>>(w.big, w.n_big + 1, w.big_cap, w.cell_start, sx,
                                             sy, sz, ll * ll, w.parent);
    HALMA_LAUNCH_CHECK();
    k_roots<<<grid_cells, 256, 0, st>>>(w.totals, w.parent, w.root);
    HALMA_LAUNCH_CHECK();
    k_check_roots<<<grid_cells, 256, 0, st>>>(w.totals, w.parent, w.root, w.cell_min,
                                              w.comp_min, w.err);
    HALMA_LAUNCH_CHECK();
    k_labels<<<blocks_for(n), 256, 0, st>>>(n, w.pcell, w.root, w.comp_min, labels);
    HALMA_LAUNCH_CHECK();
    return 0;
}

}  // namespace

// Shared with the group builder: runs FoF into `labels` and leaves the cell structures in
// the workspace.  Asynchronous; *err (device) is non-zero on a range or table failure.
int fof_labels_async(const void* x, const void* y, const void* z, int dtype, int64_t n,
                     double ll, int* labels, void* ws, size_t ws_bytes, int** err_out,
                     int64_t** ncells_out, hipStream_t st) {
    const size_t elt = dtype == HALMA_F64 ? sizeof(double) : sizeof(float);
    FofWs w = carve_fof(ws, n, elt);
    if (ws_bytes < w.bytes)
        return set_error(HALMA_E_WORKSPACE, "workspace %zu < required %zu", ws_bytes, w.bytes);
    if (err_out) *err_out = w.err;
    if (ncells_out) *ncells_out = w.totals;
    if (dtype == HALMA_F64)
        return fof_run<double>(static_cast<const double*>(x), static_cast<const double*>(y),
                               static_cast<const double*>(z), n, ll, labels, w, st);
    return fof_run<float>(static_cast<const float*>(x), static_cast<const float*>(y),
                          static_cast<const float*>(z), n, ll, labels, w, st);
}

size_t fof_workspace_bytes(int64_t n, int dtype) {
    return carve_fof(nullptr, n, dtype == HALMA_F64 ? sizeof(double) : sizeof(float)).bytes;
}

int fof_check_error(const int* d_err, hipStream_t st) {
    long long* h = pinned_words(2);
    if (!h) return set_error(HALMA_E_STATE, "pinned host buffer unavailable");
    h[0] = 0;
    HALMA_HIP(hipMemcpyAsync(h, d_err, sizeof(int), hipMemcpyDeviceToHost, st));
    HALMA_HIP(hipStreamSynchronize(st));
    const int e = static_cast<int>(h[0] & 0xffffffff);
    if (e & 1)
        return set_error(HALMA_E_LIMIT,
                         "friends-of-friends: coordinates span more than 2^21 sub-cells of "
                         "side ll/sqrt(3) on an axis, or are not finite");
    if (e & 2) return set_error(HALMA_E_STATE, "friends-of-friends: cell table overflow");
    if (e & 4) return set_error(HALMA_E_STATE, "friends-of-friends: inconsistent union-find roots");
    return 0;
}

}  // namespace halma

using namespace halma;

extern "C" {

size_t halma_fof_workspace_size(int64_t n, int dtype) {
    if (n < 0 || (dtype != HALMA_F32 && dtype != HALMA_F64)) return 0;
    return fof_workspace_bytes(n, dtype);
}

int halma_fof_labels(const void* x, const void* y, const void* z, int dtype, int64_t n,
                     double ll, int32_t* labels, void* ws, size_t ws_bytes, void* stream) {
    if (n < 0) return set_error(HALMA_E_ARG, "negative particle count");
    if (n == 0) return 0;
    if (n > (int64_t{1} << 30)) return set_error(HALMA_E_LIMIT, "at most 2^30 particles");
    if (!x || !y || !z || !labels) return set_error(HALMA_E_ARG, "null pointer");
    if (dtype != HALMA_F32 && dtype != HALMA_F64) return set_error(HALMA_E_ARG, "bad dtype");
    if (!(ll > 0.0) || !isfinite(ll))
        return set_error(HALMA_E_ARG, "linking length must be positive and finite");
    hipStream_t st = static_cast<hipStream_t>(stream);
    int* err = nullptr;
    int rc = fof_labels_async(x, y, z, dtype, n, ll, labels, ws, ws_bytes, &err, nullptr, st);
    if (rc) return rc;
    return fof_check_error(err, st);
}

}  // extern "C"
