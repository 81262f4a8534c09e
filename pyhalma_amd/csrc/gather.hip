// Source gathering for many haloes at once: the particles within a radius of each centre
// (halo_gas.py:223-277 queries its KD-trees one halo at a time, pyHALMA.py:1172-1185 sums
// the dark-matter mass inside a sphere the same way).
//
// The points are binned once into a uniform grid (cell keys (iz ny + iy) nx + ix, stable
// radix sort, packed FP64 copies in cell order, dense cell start offsets), so the points
// of one (iy, iz) row of cells between two ix are contiguous.  One block per centre walks
// the rows of the cells that can hold a point within R (the box of the sphere widened by
// one cell against rounding), each wavefront a contiguous range of rows: a first kernel
// counts the hits per wavefront (and sums their masses in double-double), a scan gives
// each centre its output range, a second kernel writes every wavefront's hits after the
// counts of the wavefronts before it (hits in row order), and an optional pair of stable
// radix passes puts every centre's indices in ascending order.
//
// Predicates: (dx^2 + dy^2) + dz^2 <= R^2 (the KD-tree's query_ball_point) or
// sqrt((dx^2 + dy^2) + dz^2) < R (parallel_inside, halo_gas.py:218-220), in FP64.
#include <math.h>

#include <algorithm>

#include "halma_common.h"
#include "radix.h"
#include "scan.h"

namespace halma {

namespace {

constexpr int kThreads = 256;
constexpr int kRB = 2048;                  // reduction blocks of the bounding box
constexpr int kWavesPerBlock = kThreads / kWave;

inline unsigned blocks_for(int64_t n) {
    return static_cast<unsigned>((n + kThreads - 1) / kThreads);
}

struct Pts {
    const void* x;
    const void* y;
    const void* z;
    int t;
};

// bounding box: (min x, min y, min z, max x, max y, max z) partial per block ...
__global__ __launch_bounds__(kThreads) void k_bbox(Pts p, int64_t n, double* __restrict__ part) {
    __shared__ double sh[6 * kWavesPerBlock];
    double v[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
    for (int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; k < n;
         k += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const double c[3] = {ldf(p.x, p.t, k), ldf(p.y, p.t, k), ldf(p.z, p.t, k)};
        for (int d = 0; d < 3; ++d) {
            v[d] = fmin(v[d], c[d]);
            v[3 + d] = fmax(v[3 + d], c[d]);
        }
    }
    for (int s = kWave / 2; s >= 1; s >>= 1)
        for (int d = 0; d < 3; ++d) {
            v[d] = fmin(v[d], __shfl_down(v[d], s, kWave));
            v[3 + d] = fmax(v[3 + d], __shfl_down(v[3 + d], s, kWave));
        }
    const int w = threadIdx.x / kWave;
    if (__lane_id() == 0)
        for (int d = 0; d < 6; ++d) sh[6 * w + d] = v[d];
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int q = 1; q < kWavesPerBlock; ++q)
            for (int d = 0; d < 3; ++d) {
                v[d] = fmin(v[d], sh[6 * q + d]);
                v[3 + d] = fmax(v[3 + d], sh[6 * q + 3 + d]);
            }
        for (int d = 0; d < 6; ++d) part[6 * blockIdx.x + d] = v[d];
    }
}

// ... then over the partials (one wavefront)
__global__ void k_bbox_final(const double* __restrict__ part, int np, double* __restrict__ out) {
    double v[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
    for (int q = __lane_id(); q < np; q += kWave)
        for (int d = 0; d < 3; ++d) {
            v[d] = fmin(v[d], part[6 * q + d]);
            v[3 + d] = fmax(v[3 + d], part[6 * q + 3 + d]);
        }
    for (int s = kWave / 2; s >= 1; s >>= 1)
        for (int d = 0; d < 3; ++d) {
            v[d] = fmin(v[d], __shfl_down(v[d], s, kWave));
            v[3 + d] = fmax(v[3 + d], __shfl_down(v[3 + d], s, kWave));
        }
    if (__lane_id() == 0)
        for (int d = 0; d < 6; ++d) out[d] = v[d];
}

struct Grid {
    double o[3];   // origin (the bounding box minimum)
    double cs;     // cell side
    int64_t n[3];  // cells per axis
};

// cell index of coordinate v on axis d, clamped to the grid (non-finite -> 0)
__device__ __forceinline__ int64_t cell_of(const Grid& g, int d, double v) {
    const double t = floor((v - g.o[d]) / g.cs);
    if (!(t >= 0.0)) return 0;
    const double top = static_cast<double>(g.n[d] - 1);
    return t > top ? g.n[d] - 1 : static_cast<int64_t>(t);
}

__device__ __forceinline__ int64_t cell_key(const Grid& g, int64_t ix, int64_t iy, int64_t iz) {
    return (iz * g.n[1] + iy) * g.n[0] + ix;
}

__global__ void k_keys(Pts p, int64_t n, Grid g, uint32_t* __restrict__ key) {
    const int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= n) return;
    key[k] = static_cast<uint32_t>(cell_key(g, cell_of(g, 0, ldf(p.x, p.t, k)),
                                            cell_of(g, 1, ldf(p.y, p.t, k)),
                                            cell_of(g, 2, ldf(p.z, p.t, k))));
}

// start[c] = first sorted position of cell c, a lower bound in the sorted keys
// (start[ncell] = n)
__global__ void k_starts(const uint32_t* __restrict__ skey, int64_t n, int64_t ncell,
                         int32_t* __restrict__ start) {
    const int64_t c = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (c > ncell) return;
    int64_t a = 0, b = n;
    while (a < b) {
        const int64_t mid = (a + b) >> 1;
        if (static_cast<int64_t>(skey[mid]) < c)
            a = mid + 1;
        else
            b = mid;
    }
    start[c] = static_cast<int32_t>(a);
}

__global__ void k_zero_out(int64_t H, int64_t* __restrict__ off, double* __restrict__ mass) {
    const int64_t h = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (h > H) return;
    off[h] = 0;
    if (mass && h < H) mass[h] = 0.0;
}

// packed FP64 copies in input order (coalesced) ...
__global__ void k_pack(Pts p, int64_t n, double4* __restrict__ tmp) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    tmp[i] = make_double4(ldf(p.x, p.t, i), ldf(p.y, p.t, i), ldf(p.z, p.t, i), 0.0);
}

// ... then into cell order, one 32-byte record per point
__global__ void k_pack_sorted(const double4* __restrict__ tmp, const int32_t* __restrict__ perm,
                              int64_t n, double4* __restrict__ pk) {
    const int64_t j = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (j < n) pk[j] = tmp[perm[j]];
}

struct Query {
    const double* centre;  // 3 per centre
    const double* radius;
    int strict;            // 1: sqrt(d2) < R, 0: d2 <= R^2
    const void* m;         // masses to sum (may be null)
    int mt;
};

struct CellRows {
    int64_t ix0, ix1, iy0, iy1, iz0, iz1;
    __device__ int64_t rows() const {
        return ix1 < ix0 ? 0 : (iy1 - iy0 + 1) * (iz1 - iz0 + 1);
    }
};

__device__ CellRows rows_of(const Grid& g, const double* c, double R) {
    CellRows r;
    if (!(R >= 0.0) || !isfinite(c[0]) || !isfinite(c[1]) || !isfinite(c[2])) {
        r.ix0 = r.iy0 = r.iz0 = 1;
        r.ix1 = r.iy1 = r.iz1 = 0;
        return r;
    }
    int64_t lo[3], hi[3];
    for (int d = 0; d < 3; ++d) {
        lo[d] = cell_of(g, d, c[d] - R) - 1;
        hi[d] = cell_of(g, d, c[d] + R) + 1;
        lo[d] = lo[d] < 0 ? 0 : lo[d];
        hi[d] = hi[d] > g.n[d] - 1 ? g.n[d] - 1 : hi[d];
    }
    r.ix0 = lo[0];
    r.ix1 = hi[0];
    r.iy0 = lo[1];
    r.iy1 = hi[1];
    r.iz0 = lo[2];
    r.iz1 = hi[2];
    return r;
}

__device__ __forceinline__ bool is_hit(const double4& q, const double* c, double R, int strict) {
    const double dx = q.x - c[0];
    const double dy = q.y - c[1];
    const double dz = q.z - c[2];
    const double d2 = (dx * dx + dy * dy) + dz * dz;
    return strict ? sqrt(d2) < R : d2 <= R * R;
}

// sorted-position range of row r of a centre's cell box
__device__ __forceinline__ void row_range(const Grid& g, const CellRows& cr, const int32_t* start,
                                          int64_t r, int64_t& a, int64_t& b) {
    const int64_t niy = cr.iy1 - cr.iy0 + 1;
    const int64_t iz = cr.iz0 + r / niy, iy = cr.iy0 + r % niy;
    a = start[cell_key(g, cr.ix0, iy, iz)];
    b = start[cell_key(g, cr.ix1, iy, iz) + 1];
}

// the rows of a centre's cell box are split into kWavesPerBlock contiguous ranges, one
// per wavefront (the same split in both kernels)
__device__ __forceinline__ void wave_rows(int64_t nrows, int w, int64_t& r0, int64_t& r1) {
    const int64_t per = (nrows + kWavesPerBlock - 1) / kWavesPerBlock;
    r0 = w * per < nrows ? w * per : nrows;
    r1 = r0 + per < nrows ? r0 + per : nrows;
}

// one block per centre: hits per wavefront's rows (wcount, kWavesPerBlock per centre), their
// total, and their mass (double-double, fixed order)
__global__ __launch_bounds__(kThreads) void k_ball_count(int64_t H, Grid g, Query q,
                                                         const int32_t* __restrict__ start,
                                                         const int32_t* __restrict__ perm,
                                                         const double4* __restrict__ pk,
                                                         int64_t* __restrict__ count,
                                                         int64_t* __restrict__ wcount,
                                                         double* __restrict__ mass) {
    __shared__ int64_t shc[kWavesPerBlock];
    __shared__ double shm[2 * kWavesPerBlock];
    const int w = threadIdx.x / kWave;
    for (int64_t h = blockIdx.x; h < H; h += gridDim.x) {
        const double* c = q.centre + 3 * h;
        const double R = q.radius[h];
        const CellRows cr = rows_of(g, c, R);
        int64_t r0, r1;
        wave_rows(cr.rows(), w, r0, r1);
        int64_t cnt = 0;
        double mh = 0.0, ml = 0.0;
        for (int64_t r = r0; r < r1; ++r) {
            int64_t a, b;
            row_range(g, cr, start, r, a, b);
            for (int64_t j = a + __lane_id(); j < b; j += kWave)
                if (is_hit(pk[j], c, R, q.strict)) {
                    ++cnt;
                    if (q.m) dd_add(mh, ml, ldf(q.m, q.mt, perm[j]));
                }
        }
        for (int s = kWave / 2; s >= 1; s >>= 1) cnt += __shfl_down(cnt, s, kWave);
        if (q.m) {
            for (int s = kWave / 2; s >= 1; s >>= 1) {
                const double h2 = __shfl_down(mh, s, kWave), l2 = __shfl_down(ml, s, kWave);
                dd_add_dd(mh, ml, h2, l2);
            }
        }
        if (__lane_id() == 0) {
            shc[w] = cnt;
            wcount[kWavesPerBlock * h + w] = cnt;
            shm[2 * w] = mh;
            shm[2 * w + 1] = ml;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            int64_t t = shc[0];
            double th = shm[0], tl = shm[1];
            for (int v = 1; v < kWavesPerBlock; ++v) {
                t += shc[v];
                dd_add_dd(th, tl, shm[2 * v], shm[2 * v + 1]);
            }
            count[h] = t;
            if (mass) mass[h] = th + tl;
        }
        __syncthreads();
    }
}

// one block per centre: every wavefront writes the hits of its rows, in row order, after
// the counts of the wavefronts before it (rows in order -> hits in row order)
__global__ __launch_bounds__(kThreads) void k_ball_emit(int64_t H, Grid g, Query q,
                                                        const int32_t* __restrict__ start,
                                                        const int32_t* __restrict__ perm,
                                                        const double4* __restrict__ pk,
                                                        const int64_t* __restrict__ off,
                                                        const int64_t* __restrict__ wcount,
                                                        int32_t* __restrict__ out) {
    const int w = threadIdx.x / kWave;
    const int lane = __lane_id();
    const unsigned long long below = lane == 0 ? 0ull : (~0ull >> (kWave - lane));
    for (int64_t h = blockIdx.x; h < H; h += gridDim.x) {
        const double* c = q.centre + 3 * h;
        const double R = q.radius[h];
        const CellRows cr = rows_of(g, c, R);
        int64_t r0, r1;
        wave_rows(cr.rows(), w, r0, r1);
        int64_t base = off[h];
        for (int v = 0; v < w; ++v) base += wcount[kWavesPerBlock * h + v];
        for (int64_t r = r0; r < r1; ++r) {
            int64_t a, b;
            row_range(g, cr, start, r, a, b);
            for (int64_t j0 = a; j0 < b; j0 += kWave) {
                const int64_t j = j0 + lane;
                const bool hit = j < b && is_hit(pk[j], c, R, q.strict);
                const unsigned long long mask = __ballot(hit);
                if (hit) out[base + __popcll(mask & below)] = perm[j];
                base += __popcll(mask);
            }
        }
    }
}

__global__ void k_entry_centre(const int64_t* __restrict__ off, int64_t H,
                               uint32_t* __restrict__ cid) {
    for (int64_t h = blockIdx.x; h < H; h += gridDim.x)
        for (int64_t k = off[h] + threadIdx.x; k < off[h + 1]; k += blockDim.x)
            cid[k] = static_cast<uint32_t>(h);
}

__global__ void k_gather_u32(const uint32_t* __restrict__ src, const int32_t* __restrict__ perm,
                             int64_t n, uint32_t* __restrict__ dst) {
    const int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k < n) dst[k] = src[perm[k]];
}

__global__ void k_gather_i32(const int32_t* __restrict__ src, const int32_t* __restrict__ perm,
                             int64_t n, int32_t* __restrict__ dst) {
    const int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k < n) dst[k] = src[perm[k]];
}

int bits_for(uint64_t top) {  // bits needed for 0..top (at least 1)
    int b = 1;
    while (b < 64 && (top >> b) != 0) ++b;
    return b;
}

struct BallWs {
    double* part;
    double* bbox;
    uint32_t* key;
    uint32_t* skey;
    int32_t* perm;
    double4* pk;
    double4* tmp;
    int32_t* start;
    int64_t* wcount;
    int64_t* count;
    int64_t* total;
    int32_t* raw;
    uint32_t* cid;
    uint32_t* k2;
    uint32_t* k3;
    int32_t* p0;
    int32_t* p1;
    void* rws;
    void* sws;
    int64_t max_cells;
    size_t bytes;
};

BallWs carve_ball(void* base, int64_t n, int64_t H, int64_t max_found) {
    Carver c(base);
    BallWs w;
    const int64_t n1 = std::max<int64_t>(n, 1), h1 = std::max<int64_t>(H, 1);
    const int64_t f1 = std::max<int64_t>(max_found, 1);
    w.max_cells = std::min<int64_t>(std::max<int64_t>(4 * n1, int64_t{1} << 16), INT32_MAX - 1);
    w.part = c.take<double>(6 * kRB);
    w.bbox = c.take<double>(8);
    w.key = c.take<uint32_t>(n1);
    w.skey = c.take<uint32_t>(n1);
    w.perm = c.take<int32_t>(n1);
    w.pk = c.take<double4>(n1);
    w.tmp = c.take<double4>(n1);
    w.wcount = c.take<int64_t>(kWavesPerBlock * h1);
    w.start = c.take<int32_t>(w.max_cells + 1);
    w.count = c.take<int64_t>(h1 + 1);
    w.total = c.take<int64_t>(2);
    w.raw = c.take<int32_t>(f1);
    w.cid = c.take<uint32_t>(f1);
    w.k2 = c.take<uint32_t>(f1);
    w.k3 = c.take<uint32_t>(f1);
    w.p0 = c.take<int32_t>(f1);
    w.p1 = c.take<int32_t>(f1);
    w.rws = c.take<char>(radix_workspace_bytes(std::max(n1, f1)));
    w.sws = c.take<char>(scan_workspace_bytes(h1 + 1));
    w.bytes = c.used;
    return w;
}

// ---------------------------------------------------------------- AMR gas cells ----
// Gas particles from AMR cells (halo_gas.py:9-141, with the sphere test and mass scaling of
// :236-246 when asked): for every query, every level >= 1 patch, the cells whose centre
// x0 + ix res (x0 = rx - res/2) lies strictly inside the query's box per axis, that are
// neither refined nor overlapped (cr0amr and solapst) and, with a sphere, lie within
// sqrt(d2) < R.  One block per halo culls the patches
// 256 at a time (a box-overlap test, widened by a cell against rounding), then walks the
// candidate patches in order, the cells of each patch's clip box a block-width at a time
// in (ix, iy, iz) order -- the reference's loop order, kept by the chunk scan of the
// emitting kernel.
constexpr int kMaxLevels = 40;

struct AmrPatches {
    int64_t P;
    const int32_t* level;
    const int32_t* nx;
    const int32_t* ny;
    const int32_t* nz;
    const double* rx;
    const double* ry;
    const double* rz;
    const int64_t* cell_off;
    double res[kMaxLevels];   // cell side per level
    double res3[kMaxLevels];  // its cube, as the caller computes it
    int n_levels;
};

struct AmrCells {
    const void* delta;
    const void* vx;
    const void* vy;
    const void* vz;
    const void* temp;
    int t;
    const uint8_t* cr0amr;
    const uint8_t* solapst;
    double rho_B;
    double mass_scale;  // rete^3 after the sphere test, 1 without
    double vel_scale;   // 3e5 (c = 1 units to km/s), 1 for the raw patch values
};

struct PatchClip {
    int64_t ix0, ix1, iy0, iy1, iz0, iz1, base;
    int32_t ny, nz;
    double x0, y0, z0, res, res3;
    __device__ int64_t rows() const {
        return (ix1 < ix0 || iy1 < iy0 || iz1 < iz0) ? 0 : (ix1 - ix0 + 1) * (iy1 - iy0 + 1);
    }
};

// cell indices k with x0 + k res possibly strictly inside (lo, hi), widened by one
// (k0 > k1 when none; non-finite bounds give none)
__device__ __forceinline__ void clip_axis(double x0, double res, int32_t n, double lo, double hi,
                                          int64_t& k0, int64_t& k1) {
    const double a = ceil((lo - x0) / res) - 1.0;
    const double b = floor((hi - x0) / res) + 1.0;
    const double top = static_cast<double>(n - 1);
    if (!isfinite(a) || !isfinite(b) || n < 1 || b < 0.0 || a > top) {
        k0 = 1;
        k1 = 0;
        return;
    }
    k0 = a < 0.0 ? 0 : static_cast<int64_t>(a);
    k1 = b > top ? n - 1 : static_cast<int64_t>(b);
}

__device__ PatchClip clip_patch(const AmrPatches& pa, int64_t p, const double* box) {
    PatchClip c;
    c.ix0 = 1;
    c.ix1 = 0;
    c.iy0 = c.iz0 = 0;
    c.iy1 = c.iz1 = 0;
    const int l = pa.level[p];
    if (l < 1 || l >= pa.n_levels || !(pa.res[l] > 0.0)) return c;
    c.res = pa.res[l];
    c.res3 = pa.res3[l];
    c.x0 = pa.rx[p] - c.res / 2;
    c.y0 = pa.ry[p] - c.res / 2;
    c.z0 = pa.rz[p] - c.res / 2;
    c.ny = pa.ny[p];
    c.nz = pa.nz[p];
    c.base = pa.cell_off[p];
    clip_axis(c.x0, c.res, pa.nx[p], box[0], box[1], c.ix0, c.ix1);
    clip_axis(c.y0, c.res, c.ny, box[2], box[3], c.iy0, c.iy1);
    clip_axis(c.z0, c.res, c.nz, box[4], box[5], c.iz0, c.iz1);
    return c;
}

__device__ __forceinline__ bool amr_hit(const PatchClip& c, const AmrCells& g, const double* box,
                                        const double* cen, double R, int64_t ix, int64_t iy,
                                        int64_t iz, int64_t& cell, double& x, double& y,
                                        double& z) {  // cen null: the box alone
    x = c.x0 + static_cast<double>(ix) * c.res;
    y = c.y0 + static_cast<double>(iy) * c.res;
    z = c.z0 + static_cast<double>(iz) * c.res;
    if (!(x > box[0] && x < box[1] && y > box[2] && y < box[3] && z > box[4] && z < box[5]))
        return false;
    cell = c.base + (ix * c.ny + iy) * c.nz + iz;
    if (!(g.cr0amr[cell] && g.solapst[cell])) return false;
    if (!cen) return true;
    const double dx = x - cen[0], dy = y - cen[1], dz = z - cen[2];
    return sqrt((dx * dx + dy * dy) + dz * dz) < R;
}

// the patches of a halo's box, in order, kThreads at a time into an LDS list
template <typename F>
__device__ void for_candidate_patches(const AmrPatches& pa, const double* box, int32_t* list,
                                      int* nlist, F&& f) {
    for (int64_t p0 = 0; p0 < pa.P; p0 += blockDim.x) {
        const int64_t p = p0 + threadIdx.x;
        const bool cand = p < pa.P && clip_patch(pa, p, box).rows() > 0;
        const unsigned long long mask = __ballot(cand);
        const int w = threadIdx.x / kWave, lane = __lane_id();
        __shared__ int wcnt[kWavesPerBlock];
        if (lane == 0) wcnt[w] = __popcll(mask);
        __syncthreads();
        int base = 0;
        for (int v = 0; v < w; ++v) base += wcnt[v];
        if (cand) {
            const unsigned long long below = lane == 0 ? 0ull : (~0ull >> (kWave - lane));
            list[base + __popcll(mask & below)] = static_cast<int32_t>(p);
        }
        if (threadIdx.x == 0) {
            int t = 0;
            for (int v = 0; v < kWavesPerBlock; ++v) t += wcnt[v];
            *nlist = t;
        }
        __syncthreads();
        const int n = *nlist;
        for (int i = 0; i < n; ++i) f(static_cast<int64_t>(list[i]));
        __syncthreads();
    }
}

struct AmrHalo {
    const double* box;     // 6 per query: x lo, x hi, y lo, y hi, z lo, z hi
    const double* centre;  // 3 per query, null: no sphere test
    const double* radius;
};

// cell q (0 .. cells()) of a patch's clip box in (ix, iy, iz) order
__device__ __forceinline__ void clip_cell(const PatchClip& c, int64_t q, int64_t& ix, int64_t& iy,
                                          int64_t& iz) {
    const int64_t niz = c.iz1 - c.iz0 + 1, niy = c.iy1 - c.iy0 + 1;
    iz = c.iz0 + q % niz;
    iy = c.iy0 + (q / niz) % niy;
    ix = c.ix0 + q / (niz * niy);
}

__device__ __forceinline__ int64_t clip_cells(const PatchClip& c) {
    return c.rows() * (c.iz1 - c.iz0 + 1);
}

__global__ __launch_bounds__(kThreads) void k_amr_count(int64_t H, AmrPatches pa, AmrCells g,
                                                        AmrHalo hq, int64_t* __restrict__ count) {
    __shared__ int32_t list[kThreads];
    __shared__ int nlist;
    __shared__ int64_t shc[kWavesPerBlock];
    const int w = threadIdx.x / kWave;
    for (int64_t h = blockIdx.x; h < H; h += gridDim.x) {
        const double* cen = hq.centre ? hq.centre + 3 * h : nullptr;
        const double R = hq.centre ? hq.radius[h] : 0.0;
        const double* box = hq.box + 6 * h;
        int64_t cnt = 0;
        if (!cen || R >= 0.0)
            for_candidate_patches(pa, box, list, &nlist, [&](int64_t p) {
                const PatchClip c = clip_patch(pa, p, box);
                const int64_t nc = clip_cells(c);
                for (int64_t q = threadIdx.x; q < nc; q += blockDim.x) {
                    int64_t ix, iy, iz, cell;
                    double x, y, z;
                    clip_cell(c, q, ix, iy, iz);
                    if (amr_hit(c, g, box, cen, R, ix, iy, iz, cell, x, y, z)) ++cnt;
                }
            });
        for (int s = kWave / 2; s >= 1; s >>= 1) cnt += __shfl_down(cnt, s, kWave);
        if (__lane_id() == 0) shc[w] = cnt;
        __syncthreads();
        if (threadIdx.x == 0) {
            int64_t t = 0;
            for (int v = 0; v < kWavesPerBlock; ++v) t += shc[v];
            count[h] = t;
        }
        __syncthreads();
    }
}

// the same traversal, the clip box's cells a block-width at a time in (ix, iy, iz) order:
// each chunk's hits are written after the chunk's earlier hits (ballot + wave counts)
__global__ __launch_bounds__(kThreads) void k_amr_emit(int64_t H, AmrPatches pa, AmrCells g,
                                                       AmrHalo hq, const int64_t* __restrict__ off,
                                                       int64_t stride, double* __restrict__ out) {
    __shared__ int32_t list[kThreads];
    __shared__ int nlist;
    __shared__ int64_t shc[kWavesPerBlock];
    const int w = threadIdx.x / kWave;
    const int lane = __lane_id();
    const unsigned long long below = lane == 0 ? 0ull : (~0ull >> (kWave - lane));
    for (int64_t h = blockIdx.x; h < H; h += gridDim.x) {
        const double* cen = hq.centre ? hq.centre + 3 * h : nullptr;
        const double R = hq.centre ? hq.radius[h] : 0.0;
        const double* box = hq.box + 6 * h;
        int64_t running = off[h];
        if (cen && !(R >= 0.0)) continue;  // block-uniform
        for_candidate_patches(pa, box, list, &nlist, [&](int64_t p) {
            const PatchClip c = clip_patch(pa, p, box);
            const int64_t nc = clip_cells(c);
            for (int64_t q0 = 0; q0 < nc; q0 += blockDim.x) {
                const int64_t q = q0 + threadIdx.x;
                int64_t ix = 0, iy = 0, iz = 0, cell = 0;
                double x = 0, y = 0, z = 0;
                bool hit = false;
                if (q < nc) {
                    clip_cell(c, q, ix, iy, iz);
                    hit = amr_hit(c, g, box, cen, R, ix, iy, iz, cell, x, y, z);
                }
                const unsigned long long mask = __ballot(hit);
                if (lane == 0) shc[w] = __popcll(mask);
                __syncthreads();
                int64_t base = running;
                for (int v = 0; v < w; ++v) base += shc[v];
                int64_t total = 0;
                for (int v = 0; v < kWavesPerBlock; ++v) total += shc[v];
                if (hit) {
                    const int64_t o = base + __popcll(mask & below);
                    out[o] = x;
                    out[stride + o] = y;
                    out[2 * stride + o] = z;
                    out[3 * stride + o] = ldf(g.vx, g.t, cell) * g.vel_scale;
                    out[4 * stride + o] = ldf(g.vy, g.t, cell) * g.vel_scale;
                    out[5 * stride + o] = ldf(g.vz, g.t, cell) * g.vel_scale;
                    out[6 * stride + o] =
                        (1.0 + ldf(g.delta, g.t, cell)) * g.rho_B * c.res3 * g.mass_scale;
                    out[7 * stride + o] = ldf(g.temp, g.t, cell);
                }
                running += total;
                __syncthreads();
            }
        });
    }
}

}  // namespace

}  // namespace halma

using namespace halma;

extern "C" {

size_t halma_ball_query_workspace_size(int64_t n_points, int64_t n_centres, int64_t max_found) {
    if (n_points < 0 || n_centres < 0 || max_found < 0) return 0;
    return carve_ball(nullptr, n_points, n_centres, max_found).bytes;
}

int halma_ball_query(int64_t n_points, const void* x, const void* y, const void* z,
                     int pos_dtype, const void* m, int mass_dtype, int64_t n_centres,
                     const double* centres, const double* radii, int strict, double cell,
                     int sorted, int64_t max_found, int64_t* offsets, int32_t* indices,
                     double* mass_sums, int64_t* n_found, void* ws, size_t ws_bytes,
                     void* stream) {
    if (n_points < 0 || n_centres < 0 || max_found < 0 || !n_found)
        return set_error(HALMA_E_ARG, "bad argument");
    *n_found = 0;
    if (n_centres == 0) return 0;
    if (!centres || !radii || !offsets || (n_points > 0 && (!x || !y || !z)) ||
        (mass_sums && n_points > 0 && !m))
        return set_error(HALMA_E_ARG, "null pointer");
    if ((pos_dtype != HALMA_F32 && pos_dtype != HALMA_F64) ||
        (m && mass_dtype != HALMA_F32 && mass_dtype != HALMA_F64))
        return set_error(HALMA_E_ARG, "bad dtype");
    if (n_points > INT32_MAX || n_centres > INT32_MAX)
        return set_error(HALMA_E_LIMIT, "at most 2^31-1 points and centres");
    BallWs w = carve_ball(ws, n_points, n_centres, max_found);
    if (ws_bytes < w.bytes)
        return set_error(HALMA_E_WORKSPACE, "workspace %zu < required %zu", ws_bytes, w.bytes);
    hipStream_t st = static_cast<hipStream_t>(stream);
    Pts p{x, y, z, pos_dtype};
    Grid g{};
    if (n_points > 0) {
        k_bbox<<<kRB, kThreads, 0, st>>>(p, n_points, w.part);
        HALMA_LAUNCH_CHECK();
        k_bbox_final<<<1, kWave, 0, st>>>(w.part, kRB, w.bbox);
        HALMA_LAUNCH_CHECK();
        double bb[6];
        HALMA_HIP(hipMemcpyAsync(bb, w.bbox, sizeof(bb), hipMemcpyDeviceToHost, st));
        HALMA_HIP(hipStreamSynchronize(st));
        for (int d = 0; d < 3; ++d)
            if (!std::isfinite(bb[d]) || !std::isfinite(bb[3 + d]))
                return set_error(HALMA_E_ARG, "non-finite point coordinates");
        double ext = 0.0;
        for (int d = 0; d < 3; ++d) ext = std::max(ext, bb[3 + d] - bb[d]);
        double cs = cell > 0.0 && std::isfinite(cell) ? cell : 0.0;
        if (cs == 0.0) cs = ext > 0.0 ? ext / std::cbrt(static_cast<double>(n_points)) : 1.0;
        if (!(cs > 0.0)) cs = 1.0;
        for (;;) {  // a coarser grid when the cells would not fit the start table
            int64_t tot = 1;
            bool ok = true;
            for (int d = 0; d < 3; ++d) {
                const double nd = std::floor((bb[3 + d] - bb[d]) / cs) + 1.0;
                if (!(nd < 1e9)) ok = false;
                g.n[d] = ok ? static_cast<int64_t>(nd) : 1;
                tot = ok ? tot * g.n[d] : tot;
                if (tot > w.max_cells) ok = false;
            }
            if (ok) break;
            cs *= 1.25;
        }
        g.cs = cs;
        for (int d = 0; d < 3; ++d) g.o[d] = bb[d];
        const int64_t ncell = g.n[0] * g.n[1] * g.n[2];
        k_keys<<<blocks_for(n_points), kThreads, 0, st>>>(p, n_points, g, w.key);
        HALMA_LAUNCH_CHECK();
        int rc = radix_sort_pairs(w.key, nullptr, w.skey, w.perm, n_points,
                                  bits_for(static_cast<uint64_t>(ncell - 1)), w.rws, st);
        if (rc) return rc;
        k_starts<<<blocks_for(ncell + 1), kThreads, 0, st>>>(w.skey, n_points, ncell, w.start);
        HALMA_LAUNCH_CHECK();
        k_pack<<<blocks_for(n_points), kThreads, 0, st>>>(p, n_points, w.tmp);
        HALMA_LAUNCH_CHECK();
        k_pack_sorted<<<blocks_for(n_points), kThreads, 0, st>>>(w.tmp, w.perm, n_points, w.pk);
        HALMA_LAUNCH_CHECK();
    } else {
        g.cs = 1.0;
        g.n[0] = g.n[1] = g.n[2] = 1;
    }
    Query q{centres, radii, strict, mass_sums ? m : nullptr, mass_dtype};
    const unsigned hb = static_cast<unsigned>(std::min<int64_t>(n_centres, 65536));
    if (n_points > 0) {
        k_ball_count<<<hb, kThreads, 0, st>>>(n_centres, g, q, w.start, w.perm, w.pk, w.count,
                                              w.wcount, mass_sums);
        HALMA_LAUNCH_CHECK();
        int rc = exclusive_scan_i64(w.count, offsets, n_centres + 1, w.total, w.sws, st);
        if (rc) return rc;
        HALMA_HIP(hipMemcpyAsync(n_found, offsets + n_centres, sizeof(int64_t),
                                 hipMemcpyDeviceToHost, st));
        HALMA_HIP(hipStreamSynchronize(st));
    } else {  // no points: empty ranges, zero masses
        k_zero_out<<<blocks_for(n_centres + 1), kThreads, 0, st>>>(n_centres, offsets, mass_sums);
        HALMA_LAUNCH_CHECK();
    }
    const int64_t F = *n_found;
    if (!indices) return 0;  // counting only
    if (F > max_found)
        return set_error(HALMA_E_LIMIT, "%lld points found exceed max_found = %lld",
                         static_cast<long long>(F), static_cast<long long>(max_found));
    if (F > INT32_MAX) return set_error(HALMA_E_LIMIT, "at most 2^31-1 points found");
    if (F == 0) return 0;
    k_ball_emit<<<hb, kThreads, 0, st>>>(n_centres, g, q, w.start, w.perm, w.pk, offsets,
                                         w.wcount, sorted ? w.raw : indices);
    HALMA_LAUNCH_CHECK();
    if (sorted) {  // ascending indices per centre: by index, then stably by centre
        k_entry_centre<<<hb, kThreads, 0, st>>>(offsets, n_centres, w.cid);
        HALMA_LAUNCH_CHECK();
        int rc = radix_sort_pairs(reinterpret_cast<const uint32_t*>(w.raw), nullptr, w.k2, w.p0,
                                  F, bits_for(static_cast<uint64_t>(n_points - 1)), w.rws, st);
        if (rc) return rc;
        k_gather_u32<<<blocks_for(F), kThreads, 0, st>>>(w.cid, w.p0, F, w.k3);
        HALMA_LAUNCH_CHECK();
        rc = radix_sort_pairs(w.k3, w.p0, w.k2, w.p1, F,
                              bits_for(static_cast<uint64_t>(n_centres - 1)), w.rws, st);
        if (rc) return rc;
        k_gather_i32<<<blocks_for(F), kThreads, 0, st>>>(w.raw, w.p1, F, indices);
        HALMA_LAUNCH_CHECK();
    }
    return 0;
}

size_t halma_amr_particles_workspace_size(int64_t n_haloes) {
    if (n_haloes < 0) return 0;
    Carver c(nullptr);
    c.take<int64_t>(n_haloes + 1);
    c.take<int64_t>(2);
    c.take<char>(scan_workspace_bytes(n_haloes + 1));
    return c.used;
}

int halma_amr_particles(int64_t n_patches, const int32_t* level, const int32_t* nx,
                        const int32_t* ny, const int32_t* nz, const double* rx, const double* ry,
                        const double* rz, const int64_t* cell_offsets, const void* delta,
                        const void* vx, const void* vy, const void* vz, const void* temp,
                        int gas_dtype, const uint8_t* cr0amr, const uint8_t* solapst,
                        int n_levels, const double* res, const double* res3, double rho_B,
                        double mass_scale, double vel_scale, int64_t n_haloes,
                        const double* boxes, const double* centres, const double* radii,
                        int64_t max_found, int64_t* offsets, double* out, int64_t* n_found,
                        void* ws, size_t ws_bytes, void* stream) {
    if (n_patches < 0 || n_haloes < 0 || max_found < 0 || !n_found || n_levels < 0 ||
        n_levels > kMaxLevels)
        return set_error(HALMA_E_ARG, "bad argument (at most %d levels)", kMaxLevels);
    *n_found = 0;
    if (n_haloes == 0) return 0;
    if (!boxes || (centres && !radii) || !offsets || (n_levels > 0 && (!res || !res3)) ||
        (n_patches > 0 && (!level || !nx || !ny || !nz || !rx || !ry || !rz || !cell_offsets ||
                           !delta || !vx || !vy || !vz || !temp || !cr0amr || !solapst)))
        return set_error(HALMA_E_ARG, "null pointer");
    if (gas_dtype != HALMA_F32 && gas_dtype != HALMA_F64) return set_error(HALMA_E_ARG, "bad dtype");
    if (n_patches > INT32_MAX || n_haloes > INT32_MAX)
        return set_error(HALMA_E_LIMIT, "at most 2^31-1 patches and haloes");
    Carver cv(ws);
    int64_t* count = cv.take<int64_t>(n_haloes + 1);
    int64_t* total = cv.take<int64_t>(2);
    void* sws = cv.take<char>(scan_workspace_bytes(n_haloes + 1));
    if (ws_bytes < cv.used)
        return set_error(HALMA_E_WORKSPACE, "workspace %zu < required %zu", ws_bytes, cv.used);
    hipStream_t st = static_cast<hipStream_t>(stream);
    AmrPatches pa{};
    pa.P = n_patches;
    pa.level = level;
    pa.nx = nx;
    pa.ny = ny;
    pa.nz = nz;
    pa.rx = rx;
    pa.ry = ry;
    pa.rz = rz;
    pa.cell_off = cell_offsets;
    pa.n_levels = n_levels;
    for (int l = 0; l < n_levels; ++l) {
        pa.res[l] = res[l];
        pa.res3[l] = res3[l];
    }
    AmrCells g{delta, vx, vy, vz, temp, gas_dtype, cr0amr, solapst, rho_B, mass_scale, vel_scale};
    AmrHalo hq{boxes, centres, radii};
    const unsigned hb = static_cast<unsigned>(std::min<int64_t>(n_haloes, 65536));
    k_amr_count<<<hb, kThreads, 0, st>>>(n_haloes, pa, g, hq, count);
    HALMA_LAUNCH_CHECK();
    int rc = exclusive_scan_i64(count, offsets, n_haloes + 1, total, sws, st);
    if (rc) return rc;
    HALMA_HIP(hipMemcpyAsync(n_found, offsets + n_haloes, sizeof(int64_t), hipMemcpyDeviceToHost,
                             st));
    HALMA_HIP(hipStreamSynchronize(st));
    const int64_t F = *n_found;
    if (!out) return 0;  // counting only
    if (F > max_found)
        return set_error(HALMA_E_LIMIT, "%lld gas cells found exceed max_found = %lld",
                         static_cast<long long>(F), static_cast<long long>(max_found));
    if (F == 0) return 0;
    k_amr_emit<<<hb, kThreads, 0, st>>>(n_haloes, pa, g, hq, offsets, max_found, out);
    HALMA_LAUNCH_CHECK();
    return 0;
}

}  // extern "C"
