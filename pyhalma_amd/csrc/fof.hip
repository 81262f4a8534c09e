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
//    atomics); counts, per-particle ranks and each cell's minimum particle index come
//    from the same pass, aggregated per block in LDS;
// 4. an exclusive scan of the table numbers the cells and gives their particle ranges;
//    particles are scattered into cell order with their coordinates;
// 5. one wavefront per cell looks up its 62 half-stencil neighbours (one lane each) and,
//    for every neighbour not already in the same component, searches for one linking
//    pair with 64 lanes in parallel (early exit) and hooks the two roots (lock-free
//    union-find over cells: the larger root index is hooked under the smaller, so parents
//    only decrease and no cycle can form; finds halve paths);
// 6. roots are read back with atomic reads (see k_roots), each component's minimum
//    particle index is taken over its cells, and every particle gets it as label.
#include <math.h>

#include "halma_common.h"
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

// Binning, aggregated per block: a block takes kBinChunk particles, finds (or inserts)
// each particle's cell in the global table, and counts its particles per cell in an LDS
// table first; one global atomic per distinct cell and block then reserves the block's
// ranks (and folds the block's minimum particle index into the entry's `cell` field,
// which holds the cell minimum until k_cells numbers the cells).  Dense haloes put 10^4
// particles into one sub-cell, and one global atomic per particle on that entry would
// serialise them all.
constexpr int kBinThreads = 256;
constexpr int kBinPerThread = 8;
constexpr int kBinChunk = kBinThreads * kBinPerThread;
constexpr int kBinLds = 2 * kBinChunk;  // LDS slots, load factor <= 1/2

template <typename T>
__global__ __launch_bounds__(kBinThreads) void k_fof_bin(
    const T* __restrict__ x, const T* __restrict__ y, const T* __restrict__ z, int64_t n,
    const double* __restrict__ box, double inv_s, Entry* __restrict__ tab, uint64_t mask,
    int* __restrict__ pslot, int* __restrict__ prank, int* __restrict__ err) {
    __shared__ int l_slot[kBinLds];
    __shared__ int l_cnt[kBinLds];
    __shared__ int l_min[kBinLds];
    for (int k = threadIdx.x; k < kBinLds; k += kBinThreads) {
        l_slot[k] = -1;
        l_cnt[k] = 0;
        l_min[k] = INT32_MAX;
    }
    __syncthreads();
    const int64_t base = static_cast<int64_t>(blockIdx.x) * kBinChunk + threadIdx.x;
    int loc[kBinPerThread], rk[kBinPerThread];
#pragma unroll
    for (int j = 0; j < kBinPerThread; ++j) {
        loc[j] = -1;
        rk[j] = 0;
        const int64_t i = base + j * kBinThreads;
        if (i >= n) continue;
        const double q[3] = {(static_cast<double>(x[i]) - box[0]) * inv_s,
                             (static_cast<double>(y[i]) - box[1]) * inv_s,
                             (static_cast<double>(z[i]) - box[2]) * inv_s};
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
        bool full = false;
        for (uint64_t probe = 0;; ++probe) {
            const uint64_t k = __hip_atomic_load(&tab[h].key, __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_AGENT);
            if (k == key) break;
            if (k == kEmpty) {
                uint64_t expected = kEmpty;
                if (__hip_atomic_compare_exchange_strong(&tab[h].key, &expected, key,
                                                         __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                         __HIP_MEMORY_SCOPE_AGENT) ||
                    expected == key)
                    break;
            }
            if (probe > mask) {  // table full (cannot happen with capacity >= 2n)
                full = true;
                break;
            }
            h = (h + 1) & mask;
        }
        if (full) {
            atomicOr(err, 2);
            pslot[i] = 0;
            prank[i] = 0;
            continue;
        }
        pslot[i] = static_cast<int>(h);
        unsigned l = (static_cast<unsigned>(h) * 0x9e3779b1u) >> 20;  // 12 bits
        for (;;) {
            const int old = atomicCAS(&l_slot[l], -1, static_cast<int>(h));
            if (old == -1 || old == static_cast<int>(h)) break;
            l = (l + 1) & (kBinLds - 1);
        }
        loc[j] = static_cast<int>(l);
        rk[j] = atomicAdd(&l_cnt[l], 1);
        atomicMin(&l_min[l], static_cast<int>(i));
    }
    __syncthreads();
    for (int k = threadIdx.x; k < kBinLds; k += kBinThreads) {
        const int h = l_slot[k];
        if (h >= 0) {
            l_cnt[k] = atomicAdd(&tab[h].cnt, l_cnt[k]);
            atomicMin(&tab[h].cell, l_min[k]);
        }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kBinPerThread; ++j)
        if (loc[j] >= 0) prank[base + j * kBinThreads] = l_cnt[loc[j]] + rk[j];
}
static_assert(kBinLds == 4096, "the LDS slot hash takes the top 12 bits");

// packed (count << 32 | occupied) per table slot for one scan that yields both the
// particle offset (high word) and the cell number (low word) of every occupied slot
__global__ void k_slot_pack(const Entry* __restrict__ tab, int64_t cap, int64_t* __restrict__ v) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= cap) return;
    const int c = tab[i].cnt;
    v[i] = (static_cast<int64_t>(c) << 32) | (c > 0 ? 1 : 0);
}

__global__ void k_cells(Entry* __restrict__ tab, const int64_t* __restrict__ scan, int64_t cap,
                        int* __restrict__ cell_start, int* __restrict__ cell_cnt,
                        uint64_t* __restrict__ cell_key, int* __restrict__ cell_min,
                        int* __restrict__ parent, int* __restrict__ comp_min) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= cap) return;
    const int c = tab[i].cnt;
    if (c <= 0) return;
    const int64_t v = scan[i];
    const int cell = static_cast<int>(v & 0xffffffff);
    cell_min[cell] = tab[i].cell;  // the binning left the cell's minimum index here
    tab[i].cell = cell;
    cell_start[cell] = static_cast<int>(v >> 32);
    cell_cnt[cell] = c;
    cell_key[cell] = tab[i].key;
    parent[cell] = cell;
    comp_min[cell] = INT32_MAX;
}

template <typename T>
__global__ void k_scatter(const T* __restrict__ x, const T* __restrict__ y,
                          const T* __restrict__ z, int64_t n, const int* __restrict__ pslot,
                          const int* __restrict__ prank, const int64_t* __restrict__ scan,
                          T* __restrict__ sx, T* __restrict__ sy,
                          T* __restrict__ sz, int* __restrict__ pcell) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int64_t v = scan[pslot[i]];
    const int64_t pos = (v >> 32) + prank[i];
    sx[pos] = x[i];
    sy[pos] = y[i];
    sz[pos] = z[i];
    pcell[i] = static_cast<int>(v & 0xffffffff);
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

template <typename T>
__global__ __launch_bounds__(256) void k_fof_link(
    const Entry* __restrict__ tab, uint64_t mask,
    const int* __restrict__ cell_start, const int* __restrict__ cell_cnt,
    const uint64_t* __restrict__ cell_key, const T* __restrict__ sx, const T* __restrict__ sy,
    const T* __restrict__ sz, const int64_t* __restrict__ totals, double ll2,
    int* __restrict__ parent) {
    const int lane = __lane_id();
    const int64_t n_cells = totals[0] & 0xffffffff;
    const int64_t wave0 = (static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x) / kWave;
    const int64_t nwaves = static_cast<int64_t>(gridDim.x) * blockDim.x / kWave;
    const uint64_t cmask = (1ull << kCellBits) - 1;
    for (int64_t c64 = wave0; c64 < n_cells; c64 += nwaves) {
        const int c = static_cast<int>(c64);
        const uint64_t key = cell_key[c];
        const int64_t cx = static_cast<int64_t>((key >> (2 * kCellBits)) & cmask);
        const int64_t cy = static_cast<int64_t>((key >> kCellBits) & cmask);
        const int64_t cz = static_cast<int64_t>(key & cmask);
        // lane k < 62 -> the k-th offset after the centre in (dz, dy, dx) order
        int dcell = -1;
        if (lane < 62) {
            const int idx = 63 + lane;
            const int64_t nz = cz + idx / 25 - 2;
            const int64_t ny = cy + (idx / 5) % 5 - 2;
            const int64_t nx = cx + idx % 5 - 2;
            if (nx >= 0 && ny >= 0 && nz >= 0 && nx <= static_cast<int64_t>(cmask) &&
                ny <= static_cast<int64_t>(cmask) && nz <= static_cast<int64_t>(cmask)) {
                const uint64_t nkey = (static_cast<uint64_t>(nx) << (2 * kCellBits)) |
                                      (static_cast<uint64_t>(ny) << kCellBits) |
                                      static_cast<uint64_t>(nz);
                dcell = table_find(tab, mask, nkey);
            }
        }
        uint64_t found = __ballot(dcell >= 0);
        const int sc = cell_start[c];
        const int nc = cell_cnt[c];
        while (found) {
            const int l = __builtin_ctzll(found);
            found &= found - 1;
            const int d = __shfl(dcell, l, kWave);
            if (uf_find(parent, c) == uf_find(parent, d)) continue;
            const int sd = cell_start[d];
            const int nd = cell_cnt[d];
            // lanes over the larger cell, loop over the smaller one
            const bool c_small = nc <= nd;
            const int sA = c_small ? sc : sd, nA = c_small ? nc : nd;
            const int sB = c_small ? sd : sc, nB = c_small ? nd : nc;
            bool linked = false;
            for (int b0 = 0; b0 < nB && !linked; b0 += kWave) {
                const int b = b0 + lane;
                const bool valid = b < nB;
                const int bb = valid ? sB + b : sB;
                const double xb = static_cast<double>(sx[bb]);
                const double yb = static_cast<double>(sy[bb]);
                const double zb = static_cast<double>(sz[bb]);
                for (int a = 0; a < nA; ++a) {
                    const double dx = static_cast<double>(sx[sA + a]) - xb;
                    const double dy = static_cast<double>(sy[sA + a]) - yb;
                    const double dz = static_cast<double>(sz[sA + a]) - zb;
                    const double d2 = (dx * dx + dy * dy) + dz * dz;
                    if (__any(valid && d2 <= ll2)) {
                        linked = true;
                        break;
                    }
                }
            }
            if (linked && lane == 0) uf_union(parent, c, d);
        }
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
    int64_t* scan;
    int64_t* totals;
    void* scan_ws;
    int* pslot;
    int* prank;
    int* pcell;
    int* cell_start;
    int* cell_cnt;
    uint64_t* cell_key;
    int* parent;
    int* root;
    int* comp_min;
    int* cell_min;
    void* sx;
    void* sy;
    void* sz;
    int* err;
    size_t bytes;
    int64_t cap;
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
    w.tab = c.take<Entry>(w.cap);
    w.scan = c.take<int64_t>(w.cap);
    w.totals = c.take<int64_t>(4);
    w.scan_ws = c.take<char>(scan_workspace_bytes(w.cap));
    w.pslot = c.take<int>(nn);
    w.prank = c.take<int>(nn);
    w.pcell = c.take<int>(nn);
    w.cell_start = c.take<int>(nn);
    w.cell_cnt = c.take<int>(nn);
    w.cell_key = c.take<uint64_t>(nn);
    w.parent = c.take<int>(nn);
    w.root = c.take<int>(nn);
    w.comp_min = c.take<int>(nn);
    w.cell_min = c.take<int>(nn);
    w.sx = c.take<char>(nn * elt);
    w.sy = c.take<char>(nn * elt);
    w.sz = c.take<char>(nn * elt);
    w.err = c.take<int>(64);
    w.bytes = c.used;
    return w;
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
    // sub-cell side: ll/sqrt(3) with a 1e-6 relative margin (see the file header)
    const double side = ll / sqrt(3.0) * (1.0 - 1e-6);
    const double inv_s = 1.0 / side;
    k_fof_bin<T><<<blocks_for(n, kBinChunk), kBinThreads, 0, st>>>(
        x, y, z, n, w.box, inv_s, w.tab, static_cast<uint64_t>(w.cap - 1), w.pslot, w.prank,
        w.err);
    HALMA_LAUNCH_CHECK();
    k_slot_pack<<<blocks_for(w.cap), 256, 0, st>>>(w.tab, w.cap, w.scan);
    HALMA_LAUNCH_CHECK();
    int rc = exclusive_scan_i64(w.scan, w.scan, w.cap, w.totals, w.scan_ws, st);
    if (rc) return rc;
    k_cells<<<blocks_for(w.cap), 256, 0, st>>>(w.tab, w.scan, w.cap, w.cell_start, w.cell_cnt,
                                               w.cell_key, w.cell_min, w.parent, w.comp_min);
    HALMA_LAUNCH_CHECK();
    k_scatter<T><<<blocks_for(n), 256, 0, st>>>(x, y, z, n, w.pslot, w.prank, w.scan,
                                                static_cast<T*>(w.sx), static_cast<T*>(w.sy),
                                                static_cast<T*>(w.sz), w.pcell);
    HALMA_LAUNCH_CHECK();
    k_fof_link<T><<<grid_cells, 256, 0, st>>>(
        w.tab, static_cast<uint64_t>(w.cap - 1), w.cell_start, w.cell_cnt, w.cell_key,
        static_cast<const T*>(w.sx), static_cast<const T*>(w.sy), static_cast<const T*>(w.sz),
        w.totals, ll * ll, w.parent);
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
