// Friends-of-friends labels on the GPU: connected components of the graph with an edge
// wherever ((dx*dx + dy*dy) + dz*dz) <= ll*ll in FP64 (fof.py:18-79 through scipy's
// KDTree; inclusive, non-periodic), canonical label = minimum particle index.
//
// 1. bounding box (two-pass reduction), read back once: it fixes the bits of each cell
//    coordinate;
// 2. sub-cells of side s = ll/sqrt(3) * (1 - 1e-6): two particles in one sub-cell are
//    within ll under the FP64 test (|dx|,|dy|,|dz| <= s(1+1e-9) after rounding of the
//    cell index), so a sub-cell is one vertex of the cell graph and needs no pair test;
//    every link spans at most 2 sub-cells per axis (ll/s < 1.733);
// 3. a stable LSD radix sort of the particles by cell key (the mixed-radix index
//    (x * ny + y) * nz + z) lists them cell by cell: cells of one column (x, y) are
//    consecutive and ordered by z, columns are in (x, y) order, particles ascending within
//    a cell; coordinates are gathered into that order as one vector per particle, each
//    cell is one 8-byte record
//    (first position, z), and each column enters a hash table whose 16-byte entry holds
//    its key, cell count, first cell and (between the rounds) shared root, so that a
//    neighbour column costs one line; the table is sized from the column count;
// 4. links, in two rounds (Chebyshev distance 1, then 2), one lane per (segment of up to
//    4 cells of a column, neighbour column): the half-stencil's 12 columns (dx > 0, or
//    dx = 0 and dy > 0) and the column itself (cells above); one table probe finds the
//    neighbour column, then the two z-ordered cell lists are merged for the cell pairs
//    within the round's dz.  Pairs
//    of cells with at most 8 particle pairs are tested at once (roots are looked up only
//    for a link); larger ones already joined are skipped, the rest tested by the lane or,
//    beyond 64 particle pairs, listed for one wavefront each (64 lanes over the larger
//    cell, early exit); linked cells are joined in a lock-free union-find over cells (the
//    larger root index is hooked under the smaller, so parents only decrease and no cycle
//    can form; finds halve paths);
// 5. roots are read back with atomic reads (see k_roots), each component's minimum
//    particle index is taken over its cells, and every particle gets it as label (in
//    sorted order, where its cell is known).
#include <math.h>

#include <algorithm>

#include "halma_common.h"
#include "radix.h"
#include "scan.h"
#include "staging.h"

namespace halma {

namespace {

constexpr uint64_t kEmpty = ~0ull;
constexpr double kCellLimit = 2097152.0;  // 2^21 cells per axis
constexpr int kBoxBlocks = 1024;

// A column's word: its key (x | y, at most 42 bits) in the low bits and its cell count
// (at most 2^21) above them.
constexpr int kKeyBits = 42;
constexpr uint64_t kKeyMask = (uint64_t{1} << kKeyBits) - 1;

// 16-byte column table entry, everything a lane needs of a neighbour column in one line:
// the word (kEmpty when free), the first cell, and between the rounds the root shared by
// all its cells (-1 when they differ)
struct __attribute__((aligned(16))) ColEntry {
    uint64_t word;
    int32_t c0;
    int32_t root;
};

// A column segment in link-lane order: its word (the column's key in the low bits, the
// segment's cell count above them, then how many cells of the column follow the segment,
// at most 2: all that an own-column pair can reach), its first cell and the column's
// table slot
struct __attribute__((aligned(16))) ColRec {
    uint64_t word;
    int32_t c0;
    int32_t slot;
};
constexpr int kSegLenShift = kKeyBits;       // 10 bits
constexpr int kSegMoreShift = kKeyBits + 10;  // 2 bits

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


// Cell grid of one call: origin, inverse side, the cell count of each axis, and the bits
// of the x and y cell coordinates in a column's packed table key (x << by | y).  The sort
// key of a cell is its mixed-radix index (x * ny + y) * nz + z: the same (x, y, z) order
// as packed bit fields, in as few bits as the grid needs (31 instead of 33 for a box of
// 1114^3 cells, so one 32-bit sort instead of two).
struct Grid {
    double lo[3];
    double inv_s;
    uint64_t n[3];
    int bx, by;
};

template <typename T>
__device__ __forceinline__ uint64_t cell_key(T xi, T yi, T zi, const Grid& g, bool& bad) {
    const double q[3] = {(static_cast<double>(xi) - g.lo[0]) * g.inv_s,
                         (static_cast<double>(yi) - g.lo[1]) * g.inv_s,
                         (static_cast<double>(zi) - g.lo[2]) * g.inv_s};
    uint64_t c[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const bool ok = q[d] >= 0.0 && q[d] < kCellLimit;  // also catches NaN
        c[d] = ok ? static_cast<uint64_t>(floor(q[d])) : 0;
        if (c[d] >= g.n[d]) {  // beyond the grid the box fixed: cannot happen for finite data
            bad = true;
            c[d] = 0;
        }
        bad |= !ok;
    }
    return (c[0] * g.n[1] + c[1]) * g.n[2] + c[2];
}

// per particle: the low and high 32 bits of its cell key, and a packed copy of its
// coordinates for the cell-order gather (one line per particle instead of three)
template <typename T>
__global__ __launch_bounds__(256) void k_fof_keys(const T* __restrict__ x,
                                                  const T* __restrict__ y,
                                                  const T* __restrict__ z, int64_t n, Grid g,
                                                  uint32_t* __restrict__ klo,
                                                  uint32_t* __restrict__ khi,
                                                  int* __restrict__ err,
                                                  typename Packed<T>::type* __restrict__ pk) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const T xi = x[i], yi = y[i], zi = z[i];
    pk[i] = Packed<T>::make(xi, yi, zi);
    bool bad = false;
    const uint64_t key = cell_key(xi, yi, zi, g, bad);
    if (bad) atomicOr(err, 1);
    klo[i] = static_cast<uint32_t>(key);
    khi[i] = static_cast<uint32_t>(key >> 32);
}

__global__ void k_gather_u32(const uint32_t* __restrict__ src, const int32_t* __restrict__ perm,
                             int64_t n, uint32_t* __restrict__ dst) {
    const int64_t j = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (j < n) dst[j] = src[perm[j]];
}

// per sorted position j: the particle's packed coordinates in cell order (the link
// kernels load a particle as one vector) and its key (recomputed from the same
// coordinates with the same arithmetic)
template <typename T>
__global__ __launch_bounds__(256) void k_fof_sorted(const int32_t* __restrict__ perm, int64_t n,
                                                    Grid g,
                                                    const typename Packed<T>::type* __restrict__ pk,
                                                    typename Packed<T>::type* __restrict__ spk,
                                                    uint64_t* __restrict__ skey) {
    const int64_t j = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const typename Packed<T>::type p = pk[perm[j]];
    spk[j] = p;
    bool bad = false;
    skey[j] = cell_key(p.x, p.y, p.z, g, bad);
}

// per sorted position j: the cell / column heads (bit 0 / bit 32: the key / the column
// part of the key, key / nz, differs from position j - 1)
__global__ void k_fof_heads(const uint64_t* __restrict__ skey, int64_t n, uint64_t nz,
                            int64_t* __restrict__ heads) {
    const int64_t j = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const uint64_t k = skey[j];
    const uint64_t kp = j > 0 ? skey[j - 1] : ~k;
    const int64_t cell = k != kp ? 1 : 0;
    const int64_t col = j == 0 || (k / nz) != (kp / nz) ? 1 : 0;
    heads[j] = cell | (col << 32);
}

// the column table's size from the column count (power of two >= 2 columns, >= 1024,
// <= cap), kept as its mask in totals[1]; then cleared
__global__ void k_col_table_size(int64_t* __restrict__ totals, int64_t cap) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const int64_t ncol = totals[0] >> 32;
    int64_t c = 1024;
    while (c < 2 * ncol && c < cap) c <<= 1;
    totals[1] = c - 1;
}

__global__ void k_clear_cols(const int64_t* __restrict__ totals, ColEntry* __restrict__ tab) {
    const int64_t m = totals[1] + 1;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < m;
         i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        tab[i].word = kEmpty;
        tab[i].c0 = -1;
        tab[i].root = -1;
    }
}

// Cells and columns from the sorted keys.  Cell c: its record (first sorted position,
// z = key mod nz), its smallest particle index (the sort is stable, so the first one),
// the union-find root; column q: its first cell and its packed key (x << by | y, from
// key / nz = x * ny + y).
__global__ __launch_bounds__(256) void k_fof_cells(
    const uint64_t* __restrict__ skey, const int32_t* __restrict__ perm,
    const int64_t* __restrict__ heads, const int64_t* __restrict__ before, int64_t n, Grid g,
    int2* __restrict__ crec, int* __restrict__ cell_min, int* __restrict__ parent,
    int* __restrict__ comp_min, int* __restrict__ col_cell, uint64_t* __restrict__ col_key) {
    const int64_t j = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const int64_t h = heads[j], b = before[j];
    const int c = static_cast<int>((b & 0xffffffff) + (h & 0xffffffff)) - 1;
    const int q = static_cast<int>((b >> 32) + (h >> 32)) - 1;
    const uint64_t k = skey[j];
    const uint64_t ck = k / g.n[2];  // x * ny + y
    if (h & 0xffffffff) {
        crec[c] = make_int2(static_cast<int>(j), static_cast<int>(k - ck * g.n[2]));
        cell_min[c] = perm[j];
        parent[c] = c;
        comp_min[c] = INT32_MAX;
    }
    if (h >> 32) {
        col_cell[q] = c;
        const uint64_t cx = ck / g.n[1];
        col_key[q] = (cx << g.by) | (ck - cx * g.n[1]);
    }
    if (j == n - 1) {
        crec[c + 1] = make_int2(static_cast<int>(n), INT32_MAX);
        col_cell[q + 1] = c + 1;
    }
}

// The table entry of a column key, or an entry whose word is kEmpty when the column has
// no cells.  Every word is final before the link kernels run.
__device__ __forceinline__ ColEntry col_find(const ColEntry* __restrict__ tab, uint64_t mask,
                                             uint64_t key) {
    uint64_t s = mix64(key) & mask;
    for (uint64_t probe = 0; probe <= mask; ++probe) {
        const ColEntry e = tab[s];
        if (e.word == kEmpty || (e.word & kKeyMask) == key) return e;
        s = (s + 1) & mask;
    }
    ColEntry e;
    e.word = kEmpty;
    e.c0 = -1;
    e.root = -1;
    return e;
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

template <typename V>
__device__ __forceinline__ bool within(const V& a, const V& b, double ll2) {
    const double dx = static_cast<double>(a.x) - static_cast<double>(b.x);
    const double dy = static_cast<double>(a.y) - static_cast<double>(b.y);
    const double dz = static_cast<double>(a.z) - static_cast<double>(b.z);
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


// Link lanes: a column of more than kSegMin cells is cut into segments of kSegCells
// cells, and each segment takes its own lanes (a long column's cells are then linked by
// several lanes at once instead of one long chain of finds and unions).  Lane order:
// segments grouped by the log2 of their particle count, largest first, so that the lanes
// of a wavefront walk segments of similar length (a wavefront lasts as long as its
// longest lane) and the longest start first.  A counting sort over kColClasses classes;
// each block takes one contiguous stretch of columns, counts their segments, reserves
// their places with one atomic per class, and writes them (the order within a class is
// not fixed; the links do not depend on it).
constexpr int kColClasses = 8;
constexpr int kClassBlocks = 256;
constexpr int kSegMin = 4;
constexpr int kSegCells = 4;
static_assert(kSegCells < 1024 && kSegMin < 1024, "a segment's cell count has 10 bits");

// cells per segment of a column of n cells
__device__ __forceinline__ int seg_step(int n) { return n > kSegMin ? kSegCells : (n > 0 ? n : 1); }

// a segment's class: log2 of its particle count
__device__ __forceinline__ int seg_class(const int2* __restrict__ crec, int c0, int c1) {
    const int np = crec[c1].x - crec[c0].x;
    const int lg = 31 - __clz(np > 0 ? np : 1);
    return lg < kColClasses - 1 ? lg : kColClasses - 1;
}

// counts[0..7]: segments per class; counts[8..15]: the class cursors (zero here).  Also
// enters each column in the table (word, first cell) and keeps its slot.
__global__ __launch_bounds__(256) void k_col_class_count(const int64_t* __restrict__ totals,
                                                         const int* __restrict__ col_cell,
                                                         const uint64_t* __restrict__ col_key,
                                                         int* __restrict__ col_slot,
                                                         const int2* __restrict__ crec,
                                                         ColEntry* __restrict__ tab,
                                                         int* __restrict__ counts,
                                                         int* __restrict__ err) {
    __shared__ int h[kColClasses];
    if (threadIdx.x < kColClasses) h[threadIdx.x] = 0;
    __syncthreads();
    const int64_t ncol = totals[0] >> 32;
    const int64_t per = (ncol + gridDim.x - 1) / gridDim.x;
    const int64_t q0 = blockIdx.x * per, q1 = q0 + per < ncol ? q0 + per : ncol;
    const uint64_t mask = static_cast<uint64_t>(totals[1]);
    for (int64_t q = q0 + threadIdx.x; q < q1; q += blockDim.x) {
        const int a0 = col_cell[q], a1 = col_cell[q + 1];
        const int step = seg_step(a1 - a0);
        for (int c = a0; c < a1; c += step) atomicAdd(&h[seg_class(crec, c, min(c + step, a1))], 1);
        const uint64_t key = col_key[q];
        const uint64_t word = key | (static_cast<uint64_t>(a1 - a0) << kKeyBits);
        uint64_t s = mix64(key) & mask;
        for (uint64_t probe = 0;; ++probe) {
            uint64_t expected = kEmpty;
            if (__hip_atomic_compare_exchange_strong(&tab[s].word, &expected, word,
                                                     __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                     __HIP_MEMORY_SCOPE_AGENT)) {
                tab[s].c0 = a0;
                col_slot[q] = static_cast<int>(s);
                break;
            }
            if (probe > mask) {  // table full (cannot happen: capacity >= 2 columns)
                atomicOr(err, 2);
                col_slot[q] = 0;
                break;
            }
            s = (s + 1) & mask;
        }
    }
    __syncthreads();
    if (threadIdx.x < kColClasses && h[threadIdx.x] > 0) atomicAdd(counts + threadIdx.x, h[threadIdx.x]);
}

__global__ __launch_bounds__(256) void k_col_class_scatter(const int64_t* __restrict__ totals,
                                                           const int* __restrict__ col_cell,
                                                           const uint64_t* __restrict__ col_key,
                                                           const int* __restrict__ col_slot,
                                                           const int2* __restrict__ crec,
                                                           int* __restrict__ counts,
                                                           ColRec* __restrict__ lanes) {
    __shared__ int h[kColClasses];
    __shared__ int base[kColClasses];
    if (threadIdx.x < kColClasses) h[threadIdx.x] = 0;
    __syncthreads();
    const int64_t ncol = totals[0] >> 32;
    const int64_t per = (ncol + gridDim.x - 1) / gridDim.x;
    const int64_t q0 = blockIdx.x * per, q1 = q0 + per < ncol ? q0 + per : ncol;
    for (int64_t q = q0 + threadIdx.x; q < q1; q += blockDim.x) {
        const int a0 = col_cell[q], a1 = col_cell[q + 1];
        const int step = seg_step(a1 - a0);
        for (int c = a0; c < a1; c += step) atomicAdd(&h[seg_class(crec, c, min(c + step, a1))], 1);
    }
    __syncthreads();
    if (threadIdx.x < kColClasses) {
        int before = 0;  // segments of the higher classes: the longest come first
        for (int k = threadIdx.x + 1; k < kColClasses; ++k) before += counts[k];
        base[threadIdx.x] =
            h[threadIdx.x] > 0 ? before + atomicAdd(counts + kColClasses + threadIdx.x, h[threadIdx.x]) : 0;
        h[threadIdx.x] = 0;
    }
    __syncthreads();
    for (int64_t q = q0 + threadIdx.x; q < q1; q += blockDim.x) {
        const int a0 = col_cell[q], a1 = col_cell[q + 1];
        const int step = seg_step(a1 - a0);
        for (int c = a0; c < a1; c += step) {
            const int c1 = min(c + step, a1);
            const int k = seg_class(crec, c, c1);
            ColRec r;
            r.word = col_key[q] | (static_cast<uint64_t>(c1 - c) << kSegLenShift) |
                     (static_cast<uint64_t>(min(a1 - c1, 2)) << kSegMoreShift);
            r.c0 = c;
            r.slot = col_slot[q];
            lanes[base[k] + atomicAdd(&h[k], 1)] = r;
        }
    }
}

// Between the rounds, after k_flatten (parent[c] is then c's root): the root shared by
// all cells of a column, or -1 when they differ, into the column's table entry (the
// table is walked in slot order).  Two columns with the same shared root are joined
// throughout, so round 2 skips the pair without looking at its cells.
__global__ void k_col_roots(const int64_t* __restrict__ totals, const int* __restrict__ parent,
                            ColEntry* __restrict__ tab) {
    const int64_t m = totals[1] + 1;
    for (int64_t s = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; s < m;
         s += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const uint64_t word = tab[s].word;
        if (word == kEmpty) continue;
        const int a0 = tab[s].c0, a1 = a0 + static_cast<int>(word >> kKeyBits);
        const int r = parent[a0];
        int u = r;
        for (int c = a0 + 1; c < a1; ++c)
            if (parent[c] != r) {
                u = -1;
                break;
            }
        tab[s].root = u;
    }
}

// The neighbour columns of the half stencil, (dx, dy) with dx > 0 or dx = 0 and dy > 0:
// the four within Chebyshev distance 1 first, then the eight at distance 2.
__constant__ signed char kCol[12][2] = {{0, 1}, {1, -1}, {1, 0}, {1, 1},  // near
                                        {0, 2}, {1, -2}, {1, 2}, {2, -2},
                                        {2, -1}, {2, 0}, {2, 1}, {2, 2}};
constexpr int kLaneTest = 64;  // pair tests a lane does alone
// Round 1: pairs of cells this small are tested before any root lookup.  Round 2 looks
// the roots up first (after round 1 most of its pairs are already joined).
constexpr int kTestFirst = 8;

// One candidate pair of cells c (particles [sc, sc + nc)) and d (particles [sd, sd + nd)).
// Pairs of at most TF particle pairs are tested first, and the roots looked at only when
// a pair is within ll (a test costs less than the root lookups).  Larger pairs are skipped
// when already joined, else tested here when below kLaneTest pairs, else listed for
// k_fof_big (tested here too when the list is full).  rc caches c's root (-1: not looked
// up yet; from a read-only find, possibly out of date: equal roots still prove the cells
// joined); it is refreshed after a link.
template <int TF, typename V>
__device__ __forceinline__ void candidate(int c, int& rc, int sc, int nc, int d, int sd,
                                          int nd, const V* __restrict__ spk, double ll2,
                                          int* __restrict__ parent, int2* __restrict__ big,
                                          int64_t big_cap,
                                          unsigned long long* __restrict__ n_big) {
    const int64_t pairs = static_cast<int64_t>(nc) * nd;
    if (pairs > TF) {
        if (rc < 0) rc = find_ro(parent, c);
        if (rc == find_ro(parent, d)) return;
        if (pairs > kLaneTest) {
            // neighbouring dense cells nearly always hold a close pair among their first
            // members: a 2 x 2 sample first, the whole pair to a wavefront only without one
            const int sa = nc < 2 ? nc : 2, sb = nd < 2 ? nd : 2;
            for (int b = 0; b < sb; ++b) {
                const V pb = spk[sd + b];
                for (int a = 0; a < sa; ++a)
                    if (within(spk[sc + a], pb, ll2)) {
                        uf_union(parent, c, d);
                        rc = find_ro(parent, c);
                        return;
                    }
            }
            const unsigned long long at = atomicAdd(n_big, 1ull);
            if (static_cast<int64_t>(at) < big_cap) {
                big[at] = make_int2(c, d);
                return;
            }
        }
    }
    for (int b = 0; b < nd; ++b) {
        const V pb = spk[sd + b];
        for (int a = 0; a < nc; ++a)
            if (within(spk[sc + a], pb, ll2)) {
                if (rc < 0) rc = find_ro(parent, c);
                if (rc != find_ro(parent, d)) {
                    uf_union(parent, c, d);
                    rc = find_ro(parent, c);
                }
                return;
            }
    }
}

// Link candidates, one lane per (column segment, column offset) of one round: the
// segment's cells against the cells ROUND apart above them in their own column, or
// against a neighbour column.  ROUND 1: the own cells one z apart, and the four near
// columns with |dz| <= 1 (the 13 cell offsets at Chebyshev distance 1).  ROUND 2,
// launched after every distance-1 link is in the forest (so that dense regions are
// recognised as joined and need no search): the own cells two apart, the near columns
// with |dz| = 2, the far columns with |dz| <= 2 (the other 49 offsets).  The lanes of
// one segment are consecutive and read its record together; a neighbour column costs
// one table line (first cell, cell count, shared root), its cell records and its
// particles.
template <typename T, int ROUND>
__global__ __launch_bounds__(256) void k_fof_cols(
    const ColEntry* __restrict__ tab, const ColRec* __restrict__ lanes,
    const int2* __restrict__ crec, const typename Packed<T>::type* __restrict__ spk,
    const int64_t* __restrict__ totals, int bx, int by, double ll2, int* __restrict__ parent,
    int2* __restrict__ big, int64_t big_cap, unsigned long long* __restrict__ n_big,
    const int* __restrict__ seg_counts) {
    constexpr int K = ROUND == 1 ? 5 : 13;  // own column + neighbour columns
    int64_t nseg = 0;
#pragma unroll
    for (int k = 0; k < kColClasses; ++k) nseg += seg_counts[k];
    const uint64_t mask = static_cast<uint64_t>(totals[1]);
    const int64_t total = nseg * K;
    const int64_t ymask = (int64_t{1} << by) - 1;
    const int64_t nxc = int64_t{1} << bx, nyc = int64_t{1} << by;
    for (int64_t idx = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
         idx < total; idx += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const ColRec col = lanes[idx / K];
        const int o = static_cast<int>(idx % K);
        const int a0 = col.c0;  // the segment's cells [a0, a1)
        const int a1 = a0 + static_cast<int>((col.word >> kSegLenShift) & 1023);
        const int ra = ROUND == 2 ? tab[col.slot].root : -1;  // the column's shared root
        if (o == 0) {  // the segment's cells against those ROUND above them in the column
            if (ra >= 0) continue;  // one component already
            const int e1 = a1 + static_cast<int>((col.word >> kSegMoreShift) & 3);
            int2 cc = crec[a0];
            for (int c = a0; c < a1 && c + 1 < e1; ++c) {
                const int2 cn = crec[c + 1];
                int rc = -1;
                const int sc = cc.x, nc = cn.x - cc.x;
                int2 dd = cn;
                for (int d = c + 1; d < e1 && dd.y - cc.y <= ROUND; ++d) {
                    const int2 dn = crec[d + 1];
                    if (dd.y - cc.y == ROUND)
                        candidate<ROUND == 1 ? kTestFirst : 0>(c, rc, sc, nc, d, dd.x,
                                                               dn.x - dd.x, spk, ll2, parent,
                                                               big, big_cap, n_big);
                    dd = dn;
                }
                cc = cn;
            }
            continue;
        }
        const int k = o - 1;  // round 1: the near columns; round 2: all twelve
        const uint64_t key = col.word & kKeyMask;
        const int64_t nx = static_cast<int64_t>(key >> by) + kCol[k][0];
        const int64_t ny = static_cast<int64_t>(key & ymask) + kCol[k][1];
        if (nx < 0 || ny < 0 || nx >= nxc || ny >= nyc) continue;
        const ColEntry e = col_find(tab, mask, (static_cast<uint64_t>(nx) << by) |
                                                   static_cast<uint64_t>(ny));
        if (e.word == kEmpty) continue;
        if (ra >= 0 && ra == e.root) continue;  // joined throughout
        const int b0 = e.c0, b1 = b0 + static_cast<int>(e.word >> kKeyBits);
        // |dz| range of this (round, column): near columns take |dz| <= 1 in round 1 and
        // |dz| = 2 in round 2, far columns |dz| <= 2
        const bool near = k < 4;
        const int lo = (ROUND == 2 && near) ? 2 : 0;
        const int hi = ROUND == 1 ? 1 : 2;
        int p = b0;
        int2 pp = crec[b0];
        int2 cc = crec[a0];
        for (int c = a0; c < a1; ++c) {
            const int2 cn = crec[c + 1];
            const int zc = cc.y;
            int rc = -1;
            while (p < b1 && pp.y < zc - hi) pp = crec[++p];
            int2 dd = pp;
            for (int d = p; d < b1 && dd.y <= zc + hi; ++d) {
                const int2 dn = crec[d + 1];
                const int dz = dd.y > zc ? dd.y - zc : zc - dd.y;
                if (dz >= lo)
                    candidate<ROUND == 1 ? kTestFirst : 0>(c, rc, cc.x, cn.x - cc.x, d, dd.x,
                                                           dn.x - dd.x, spk, ll2, parent, big,
                                                           big_cap, n_big);
                dd = dn;
            }
            cc = cn;
        }
    }
}

// Large pairs of cells, one wavefront each: skipped if joined meanwhile, else lanes run
// over the larger cell four blocks of 64 at a time (loads in flight together) and loop
// over the smaller cell, stopping at the first pair within ll.
template <typename T>
__global__ __launch_bounds__(256) void k_fof_big(
    const int2* __restrict__ big, const unsigned long long* __restrict__ n_big,
    int64_t big_cap, const int2* __restrict__ crec,
    const typename Packed<T>::type* __restrict__ spk, double ll2, int* __restrict__ parent) {
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
        const int sc = crec[c].x, nc = crec[c + 1].x - sc;
        const int sd = crec[d].x, nd = crec[d + 1].x - sd;
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
                const auto pb = spk[valid[u] ? sB + b : sB];
                xb[u] = static_cast<double>(pb.x);
                yb[u] = static_cast<double>(pb.y);
                zb[u] = static_cast<double>(pb.z);
            }
            for (int a = 0; a < nA; ++a) {
                const auto pa = spk[sA + a];
                const double xa = static_cast<double>(pa.x);
                const double ya = static_cast<double>(pa.y);
                const double za = static_cast<double>(pa.z);
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

// Also folds every cell's minimum particle index into its component's minimum.  The
// lanes of a wavefront whose root is the first lane's (consecutive cells mostly share
// one) are merged into one check and one atomic; the others act alone.
__global__ __launch_bounds__(256) void k_check_roots(const int64_t* __restrict__ totals,
                                                     int* __restrict__ parent,
                                                     const int* __restrict__ root,
                                                     const int* __restrict__ cell_min,
                                                     int* __restrict__ comp_min,
                                                     int* __restrict__ err) {
    const int64_t n_cells = totals[0] & 0xffffffff;
    const int lane = __lane_id();
    for (int64_t base = static_cast<int64_t>(blockIdx.x) * blockDim.x + (threadIdx.x & ~(kWave - 1));
         base < n_cells; base += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const int64_t c = base + lane;
        const bool valid = c < n_cells;  // lane 0 always is
        const int r = valid ? root[c] : -1;
        const int m = valid ? cell_min[c] : INT32_MAX;
        const int r0 = __shfl(r, 0, kWave);
        const bool same = r == r0;
        int v = same ? m : INT32_MAX;
        for (int s = 32; s >= 1; s >>= 1) v = min(v, __shfl_xor(v, s, kWave));
        if (lane == 0) {
            if (coherent_read(parent + r0) != r0) atomicOr(err, 4);
            atomicMin(comp_min + r0, v);
        }
        if (valid && !same) {
            if (coherent_read(parent + r) != r) atomicOr(err, 4);
            atomicMin(comp_min + r, m);
        }
    }
}

// labels in sorted order: position j lies in cell (cell heads before and at j) - 1, and
// its particle perm[j] gets the minimum of that cell's component
__global__ void k_labels(int64_t n, const int32_t* __restrict__ perm,
                         const int64_t* __restrict__ heads, const int64_t* __restrict__ before,
                         const int* __restrict__ root, const int* __restrict__ comp_min,
                         int* __restrict__ labels) {
    const int64_t j = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const int c = static_cast<int>((before[j] & 0xffffffff) + (heads[j] & 0xffffffff)) - 1;
    labels[perm[j]] = comp_min[root[c]];
}

struct FofWs {
    double* box_part;
    double* box;
    ColEntry* tab;
    int64_t* totals;  // [0] cells | columns << 32, [1] column table mask
    void* scan_ws;
    void* sort_ws;
    uint32_t* klo;    // low / high 32 bits of each particle's cell key
    uint32_t* khi;
    uint32_t* klo_s;  // sorted keys of the passes (unused beyond the sort)
    uint32_t* khi_g;
    uint32_t* khi_s;
    int32_t* perm1;   // order after the low-bits pass
    int32_t* perm;    // particle of each sorted position
    uint64_t* skey;   // cell key of each sorted position
    int64_t* heads;   // cell / column heads of each sorted position
    int64_t* before;  // their exclusive scan (cell and column numbers)
    int2* crec;       // per cell: first sorted position, z (n_cells + 1 entries)
    int* col_cell;    // first cell of each column (n_cols + 1 entries)
    uint64_t* col_key;
    int* col_slot;    // table slot of each column
    ColRec* lanes;    // the column segments in link-lane order (by class)
    int* class_counts;
    int* parent;
    int* root;
    int* comp_min;
    int* cell_min;
    void* spk;        // (x, y, z, 0) per sorted position
    void* pk;         // (x, y, z, 0) per particle in input order
    int2* big;        // large pairs of cells left for k_fof_big
    unsigned long long* n_big;
    int64_t big_cap;
    int* err;
    size_t bytes;
    int64_t cap;      // column table capacity for n particles (the largest it gets)
};

int64_t table_capacity(int64_t n) {
    int64_t cap = 1024;
    while (cap < 2 * n) cap <<= 1;
    return cap;
}

FofWs carve_fof(void* base, int64_t n, size_t elt) {
    Carver c(base);
    FofWs w;
    const int64_t nn = n > 0 ? n : 1;
    w.cap = table_capacity(n);
    w.box_part = c.take<double>(kBoxBlocks * 6);
    w.box = c.take<double>(8);
    w.tab = c.take<ColEntry>(w.cap);
    w.totals = c.take<int64_t>(4);
    w.scan_ws = c.take<char>(scan_workspace_bytes(nn));
    w.sort_ws = c.take<char>(radix_workspace_bytes(nn));
    w.klo = c.take<uint32_t>(nn);
    w.khi = c.take<uint32_t>(nn);
    w.klo_s = c.take<uint32_t>(nn);
    w.khi_g = c.take<uint32_t>(nn);
    w.khi_s = c.take<uint32_t>(nn);
    w.perm1 = c.take<int32_t>(nn);
    w.perm = c.take<int32_t>(nn);
    w.skey = c.take<uint64_t>(nn);
    w.heads = c.take<int64_t>(nn);
    w.before = c.take<int64_t>(nn);
    w.crec = c.take<int2>(nn + 1);
    w.col_cell = c.take<int>(nn + 1);
    w.col_key = c.take<uint64_t>(nn);
    w.col_slot = c.take<int>(nn);
    w.lanes = c.take<ColRec>(nn + nn / kSegCells + 1);  // segments <= columns + cells / 4
    w.class_counts = c.take<int>(2 * kColClasses);
    w.parent = c.take<int>(nn);
    w.root = c.take<int>(nn);
    w.comp_min = c.take<int>(nn);
    w.cell_min = c.take<int>(nn);
    w.spk = c.take<char>(nn * 4 * elt);
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

int bits_for(int64_t cells) {  // bits of the values 0 .. cells - 1
    int b = 0;
    while ((int64_t{1} << b) < cells) ++b;
    return b;
}

template <typename T>
int fof_run(const T* x, const T* y, const T* z, int64_t n, double ll, int* labels,
            const FofWs& w, hipStream_t st) {
    const int cus = std::max(1, device_cu_count());
    const int grid_cells = cus * 8;
    k_fill_i32n<<<1, 64, 0, st>>>(w.err, 1, 0);
    HALMA_LAUNCH_CHECK();
    k_bbox_partial<T><<<kBoxBlocks, 256, 0, st>>>(x, y, z, n, w.box_part);
    HALMA_LAUNCH_CHECK();
    k_bbox_final<<<1, 256, 0, st>>>(w.box_part, kBoxBlocks, w.box);
    HALMA_LAUNCH_CHECK();
    // the box fixes the grid (one small read-back)
    long long* hw = pinned_words(8);
    if (!hw) return set_error(HALMA_E_STATE, "pinned host buffer unavailable");
    double* hb = reinterpret_cast<double*>(hw);
    HALMA_HIP(hipMemcpyAsync(hb, w.box, 6 * sizeof(double), hipMemcpyDeviceToHost, st));
    HALMA_HIP(hipStreamSynchronize(st));
    // sub-cell side: ll/sqrt(3) with a 1e-6 relative margin (see the file header)
    const double side = ll / sqrt(3.0) * (1.0 - 1e-6);
    Grid g;
    g.inv_s = 1.0 / side;
    for (int d = 0; d < 3; ++d) {
        g.lo[d] = hb[d];
        const double span = (hb[3 + d] - hb[d]) * g.inv_s;
        if (!(span >= 0.0 && span < kCellLimit))  // also catches NaN and infinities
            return set_error(HALMA_E_LIMIT,
                             "friends-of-friends: coordinates span more than 2^21 sub-cells "
                             "of side ll/sqrt(3) on an axis, or are not finite");
        g.n[d] = static_cast<uint64_t>(floor(span)) + 1;  // at most 2^21
    }
    g.bx = bits_for(static_cast<int64_t>(g.n[0]));
    g.by = bits_for(static_cast<int64_t>(g.n[1]));
    // bits of the mixed-radix keys 0 .. nx * ny * nz - 1 (the product is at most 2^63)
    const uint64_t ncells = g.n[0] * g.n[1] * g.n[2];
    int kbits = 0;
    while (kbits < 64 && (ncells - 1) >> kbits) ++kbits;
    auto* pk = static_cast<typename Packed<T>::type*>(w.pk);
    k_fof_keys<T><<<blocks_for(n), 256, 0, st>>>(x, y, z, n, g, w.klo, w.khi, w.err, pk);
    HALMA_LAUNCH_CHECK();
    // stable LSD sort by the key: the low 32 bits, then the high ones
    int rc = radix_sort_pairs(w.klo, nullptr, w.klo_s, w.perm1, n, std::max(1, std::min(kbits, 32)),
                              w.sort_ws, st);
    if (rc) return rc;
    const int32_t* perm = w.perm1;
    if (kbits > 32) {
        k_gather_u32<<<blocks_for(n), 256, 0, st>>>(w.khi, w.perm1, n, w.khi_g);
        HALMA_LAUNCH_CHECK();
        rc = radix_sort_pairs(w.khi_g, w.perm1, w.khi_s, w.perm, n, kbits - 32, w.sort_ws, st);
        if (rc) return rc;
        perm = w.perm;
    }
    typedef typename Packed<T>::type V;
    V* spk = static_cast<V*>(w.spk);
    k_fof_sorted<T><<<blocks_for(n), 256, 0, st>>>(perm, n, g, pk, spk, w.skey);
    HALMA_LAUNCH_CHECK();
    k_fof_heads<<<blocks_for(n), 256, 0, st>>>(w.skey, n, g.n[2], w.heads);
    HALMA_LAUNCH_CHECK();
    rc = exclusive_scan_i64(w.heads, w.before, n, w.totals, w.scan_ws, st);
    if (rc) return rc;
    k_col_table_size<<<1, 64, 0, st>>>(w.totals, w.cap);
    HALMA_LAUNCH_CHECK();
    k_clear_cols<<<grid_cells, 256, 0, st>>>(w.totals, w.tab);
    HALMA_LAUNCH_CHECK();
    k_fof_cells<<<blocks_for(n), 256, 0, st>>>(w.skey, perm, w.heads, w.before, n, g,
                                               w.crec, w.cell_min, w.parent, w.comp_min,
                                               w.col_cell, w.col_key);
    HALMA_LAUNCH_CHECK();
    k_clear_words<<<1, 64, 0, st>>>(reinterpret_cast<uint32_t*>(w.n_big), 4);
    HALMA_LAUNCH_CHECK();
    k_clear_words<<<1, 64, 0, st>>>(reinterpret_cast<uint32_t*>(w.class_counts), 2 * kColClasses);
    HALMA_LAUNCH_CHECK();
    k_col_class_count<<<kClassBlocks, 256, 0, st>>>(w.totals, w.col_cell, w.col_key,
                                                     w.col_slot, w.crec, w.tab,
                                                     w.class_counts, w.err);
    HALMA_LAUNCH_CHECK();
    k_col_class_scatter<<<kClassBlocks, 256, 0, st>>>(w.totals, w.col_cell, w.col_key,
                                                       w.col_slot, w.crec, w.class_counts,
                                                       w.lanes);
    HALMA_LAUNCH_CHECK();
    const unsigned cand_grid = static_cast<unsigned>(cus * 32);
    const double ll2 = ll * ll;
    k_fof_cols<T, 1><<<cand_grid, 256, 0, st>>>(w.tab, w.lanes, w.crec, spk, w.totals, g.bx,
                                                g.by, ll2, w.parent, w.big, w.big_cap,
                                                w.n_big, w.class_counts);
    HALMA_LAUNCH_CHECK();
    k_fof_big<T><<<grid_cells, 256, 0, st>>>(w.big, w.n_big, w.big_cap, w.crec, spk, ll2,
                                             w.parent);
    HALMA_LAUNCH_CHECK();
    k_flatten<<<grid_cells, 256, 0, st>>>(w.totals, w.parent);
    HALMA_LAUNCH_CHECK();
    k_col_roots<<<grid_cells, 256, 0, st>>>(w.totals, w.parent, w.tab);
    HALMA_LAUNCH_CHECK();
    k_fof_cols<T, 2><<<cand_grid, 256, 0, st>>>(w.tab, w.lanes, w.crec, spk, w.totals, g.bx,
                                                g.by, ll2, w.parent, w.big, w.big_cap,
                                                w.n_big + 1, w.class_counts);
    HALMA_LAUNCH_CHECK();
    k_fof_big<T><<<grid_cells, 256, 0, st>>>(w.big, w.n_big + 1, w.big_cap, w.crec, spk, ll2,
                                             w.parent);
    HALMA_LAUNCH_CHECK();
    k_roots<<<grid_cells, 256, 0, st>>>(w.totals, w.parent, w.root);
    HALMA_LAUNCH_CHECK();
    k_check_roots<<<grid_cells, 256, 0, st>>>(w.totals, w.parent, w.root, w.cell_min,
                                              w.comp_min, w.err);
    HALMA_LAUNCH_CHECK();
    k_labels<<<blocks_for(n), 256, 0, st>>>(n, perm, w.heads, w.before, w.root, w.comp_min,
                                            labels);
    HALMA_LAUNCH_CHECK();
    return 0;
}

}  // namespace

// Runs FoF into `labels` (the cell structures stay in the workspace).  Asynchronous
// after one read-back of the bounding box; *err (device) is non-zero on a range or table
// failure.
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
