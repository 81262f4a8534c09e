// Phase-space cleaning (halo_properties.py:113-278, driver pyHALMA.py:789-843): the cell
// statistics and the per-member test of one round, for every group being cleaned.
//
// A round works on a set S of members (each group's members still flagged, positions
// ascending).  halma_clean_cells puts each member of S in the nearest node of its group's
// grid np.arange(-(R + ll), R + ll, 2 ll) per axis (about the group's centre of mass;
// the first node among equal distances, as np.argmin), sorts S by (group, cell) with
// stable radix passes (member order kept inside a cell), and gives every cell its
// mass-weighted velocity, member count and 3D dispersion sqrt(sum |v_cell - v|^2 / 3 / q)
// from double-double sums (one wavefront per cell; the reference sums in member order in
// FP64, which these match to within a few ulp).  halma_clean_decide then applies the
// reference's test to every member of S: a member of a cell with more than q_fil members
// is removed when |v_cell - v| / sigma_cell >= fac * sig_fil (|v_cell - v| counts as 0 in
// cells of at most two members in a group's first round, of fewer than two later), a
// member of a cell with at most q_fil members is removed.
#include <math.h>

#include <algorithm>

#include "halma_common.h"
#include "radix.h"
#include "scan.h"

namespace halma {

namespace {

constexpr int kThreads = 256;
constexpr int32_t kMaxGridNodes = 1 << 21;  // cell keys (ix n + iy) n + iz fit in 63 bits

inline unsigned blocks_for(int64_t n) {
    return static_cast<unsigned>((n + kThreads - 1) / kThreads);
}

__device__ __forceinline__ double ldf(const void* p, int dtype, int64_t i) {
    return dtype == HALMA_F64 ? static_cast<const double*>(p)[i]
                              : static_cast<double>(static_cast<const float*>(p)[i]);
}

// v * m as the reference forms it: in FP32 when both arrays are FP32 (numba keeps the
// product of two float32 values in float32), in FP64 otherwise
__device__ __forceinline__ double vm_product(const void* v, int vt, const void* m, int mt,
                                             int64_t i) {
    if (vt == HALMA_F32 && mt == HALMA_F32)
        return static_cast<double>(static_cast<const float*>(v)[i] *
                                   static_cast<const float*>(m)[i]);
    return ldf(v, vt, i) * ldf(m, mt, i);
}

struct Particles {
    const void* x;
    const void* y;
    const void* z;
    int pos_t;
    const void* vx;
    const void* vy;
    const void* vz;
    int vel_t;
    const void* m;
    int mass_t;
};

// grid of group h: node k = start (k = 0), start + step (k = 1), start + k * d (k >= 2)
struct GridArgs {
    const double* start;
    const double* g1;
    const double* d;
    const int32_t* n;
    const double* centre;  // 3 per group
};

__device__ __forceinline__ double node(const GridArgs& g, int h, int64_t k) {
    return k == 0 ? g.start[h] : (k == 1 ? g.g1[h] : g.start[h] + static_cast<double>(k) * g.d[h]);
}

__device__ int nearest(const GridArgs& g, int h, double xh) {
    const int n = g.n[h];
    const double est = floor((xh - g.start[h]) / g.d[h] + 0.5);
    int64_t kc = est < 0.0 ? 0 : (est > static_cast<double>(n - 1) ? n - 1 : static_cast<int64_t>(est));
    int best = -1;
    double bd = 0.0;
    for (int64_t k = kc - 2; k <= kc + 2; ++k) {
        if (k < 0 || k >= n) continue;
        const double dist = fabs(node(g, h, k) - xh);
        if (best < 0 || dist < bd) {
            best = static_cast<int>(k);
            bd = dist;
        }
    }
    return best;
}

// per member of S: its three node indices and its cell key within the group,
// (ix n + iy) n + iz as two 32-bit words
__global__ void k_clean_cells(int64_t ns, const int32_t* __restrict__ s_mem,
                              const int32_t* __restrict__ s_grp, Particles p, GridArgs g,
                              int32_t* __restrict__ ijk, uint32_t* __restrict__ klo,
                              uint32_t* __restrict__ khi) {
    const int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= ns) return;
    const int64_t i = s_mem[k];
    const int h = s_grp[k];
    const double* c = g.centre + 3 * static_cast<int64_t>(h);
    const int ix = nearest(g, h, ldf(p.x, p.pos_t, i) - c[0]);
    const int iy = nearest(g, h, ldf(p.y, p.pos_t, i) - c[1]);
    const int iz = nearest(g, h, ldf(p.z, p.pos_t, i) - c[2]);
    ijk[3 * k + 0] = ix;
    ijk[3 * k + 1] = iy;
    ijk[3 * k + 2] = iz;
    const uint64_t n = static_cast<uint64_t>(g.n[h]);
    const uint64_t key = (static_cast<uint64_t>(ix) * n + static_cast<uint64_t>(iy)) * n +
                         static_cast<uint64_t>(iz);
    klo[k] = static_cast<uint32_t>(key & 0xffffffffull);
    khi[k] = static_cast<uint32_t>(key >> 32);
}

__global__ void k_gather_u32(const uint32_t* __restrict__ src, const int32_t* __restrict__ perm,
                             int64_t n, uint32_t* __restrict__ dst) {
    const int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k < n) dst[k] = src[perm[k]];
}

// head flags of the (group, cell) runs in sorted order
__global__ void k_cell_heads(int64_t ns, const int32_t* __restrict__ order,
                             const uint32_t* __restrict__ klo, const uint32_t* __restrict__ khi,
                             const int32_t* __restrict__ s_grp, int32_t* __restrict__ head) {
    const int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= ns) return;
    int h = 1;
    if (k > 0) {
        const int32_t a = order[k], b = order[k - 1];
        h = (klo[a] != klo[b] || khi[a] != khi[b] || s_grp[a] != s_grp[b]) ? 1 : 0;
    }
    head[k] = h;
}

__global__ void k_cell_begin(int64_t ns, const int32_t* __restrict__ head,
                             const int64_t* __restrict__ segid, int64_t* __restrict__ begin,
                             int32_t* __restrict__ seg_of) {
    const int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= ns) return;
    const int64_t s = segid[k] + head[k] - 1;  // inclusive id
    if (head[k]) begin[s] = k;
    seg_of[k] = static_cast<int32_t>(s);
}

__device__ __forceinline__ void wave_dd(double& hi, double& lo) {
    // fixed butterfly: the lower lane of each pair adds the upper one's value
    for (int m = kWave / 2; m >= 1; m >>= 1) {
        const double oh = __shfl_xor(hi, m, kWave), ol = __shfl_xor(lo, m, kWave);
        if (__lane_id() & m) {
            dd_add_dd(hi, lo, oh, ol);
        } else {
            double th = oh, tl = ol;
            dd_add_dd(th, tl, hi, lo);
            hi = th;
            lo = tl;
        }
    }
}

// one wavefront per cell: sums of v m (3) and m, the count, then the dispersion
__global__ __launch_bounds__(kThreads) void k_cell_stats(
    int64_t n_seg, int64_t ns, const int64_t* __restrict__ begin,
    const int32_t* __restrict__ order, const int32_t* __restrict__ s_mem, Particles p,
    double* __restrict__ stats /* 6 per cell: vcm x, y, z, mass, count, sigma */) {
    const int lane = __lane_id();
    const int64_t seg = (static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x) / kWave;
    if (seg >= n_seg) return;  // wave-uniform
    const int64_t b = begin[seg];
    const int64_t e = seg + 1 < n_seg ? begin[seg + 1] : ns;
    double acc[4][2] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}};
    for (int64_t k = b + lane; k < e; k += kWave) {
        const int64_t i = s_mem[order[k]];
        dd_add(acc[0][0], acc[0][1], vm_product(p.vx, p.vel_t, p.m, p.mass_t, i));
        dd_add(acc[1][0], acc[1][1], vm_product(p.vy, p.vel_t, p.m, p.mass_t, i));
        dd_add(acc[2][0], acc[2][1], vm_product(p.vz, p.vel_t, p.m, p.mass_t, i));
        dd_add(acc[3][0], acc[3][1], ldf(p.m, p.mass_t, i));
    }
    for (int q = 0; q < 4; ++q) wave_dd(acc[q][0], acc[q][1]);
    const double mass = acc[3][0] + acc[3][1];
    double vc[3];
    for (int d = 0; d < 3; ++d) vc[d] = (acc[d][0] + acc[d][1]) / mass;
    double sh = 0.0, sl = 0.0;
    for (int64_t k = b + lane; k < e; k += kWave) {
        const int64_t i = s_mem[order[k]];
        const double dx = vc[0] - ldf(p.vx, p.vel_t, i);
        const double dy = vc[1] - ldf(p.vy, p.vel_t, i);
        const double dz = vc[2] - ldf(p.vz, p.vel_t, i);
        dd_add(sh, sl, (dx * dx + dy * dy) + dz * dz);
    }
    wave_dd(sh, sl);
    if (lane == 0) {
        const double q = static_cast<double>(e - b);
        double* o = stats + 6 * seg;
        o[0] = vc[0];
        o[1] = vc[1];
        o[2] = vc[2];
        o[3] = mass;
        o[4] = q;
        o[5] = sqrt((sh + sl) / 3.0 / q);
    }
}

// the test of every member of S (in sorted order); phase per group: 0 = first round
// ("> 2" rule), 1 = later round (">= 2"), other = no test.  A removed member clears
// control[target] and counts in cleaned[group].
__global__ void k_clean_decide(int64_t ns, const int32_t* __restrict__ order,
                               const int32_t* __restrict__ seg_of,
                               const int32_t* __restrict__ s_mem,
                               const int32_t* __restrict__ s_grp,
                               const int64_t* __restrict__ target, Particles p,
                               const double* __restrict__ stats,
                               const int32_t* __restrict__ phase,
                               const double* __restrict__ fac, double q_fil, double sig_fil,
                               uint8_t* __restrict__ control, int32_t* __restrict__ cleaned) {
    const int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= ns) return;
    const int32_t s = order[k];
    const int h = s_grp[s];
    const int ph = phase[h];
    if (ph != 0 && ph != 1) return;
    const double* c = stats + 6 * static_cast<int64_t>(seg_of[k]);
    const double q = c[4];
    bool removed;
    if (q > q_fil) {
        const int64_t i = s_mem[s];
        const double dx = c[0] - ldf(p.vx, p.vel_t, i);
        const double dy = c[1] - ldf(p.vy, p.vel_t, i);
        const double dz = c[2] - ldf(p.vz, p.vel_t, i);
        double bas = sqrt((dx * dx + dy * dy) + dz * dz);
        const bool scaled = ph == 0 ? q > 2.0 : q >= 2.0;
        bas = scaled ? bas / c[5] : 0.0;
        removed = bas >= fac[h] * sig_fil;
    } else {
        removed = true;
    }
    if (removed) {
        control[target[s]] = 0;
        atomicAdd(cleaned + h, 1);
    }
}

struct CleanWs {
    uint32_t* klo;
    uint32_t* khi;
    uint32_t* k2;
    uint32_t* k3;
    int32_t* p0;
    int32_t* p1;  // the (group, cell, position) order, read by halma_clean_decide
    int32_t* p2;
    int32_t* head;
    int64_t* segid;
    int64_t* begin;
    int32_t* seg_of;
    int64_t* nseg;
    void* rws;
    void* sws;
    size_t bytes;
};

CleanWs carve_clean(void* base, int64_t n) {
    Carver c(base);
    const int64_t n1 = n > 0 ? n : 1;
    CleanWs w;
    w.klo = c.take<uint32_t>(n1);
    w.khi = c.take<uint32_t>(n1);
    w.k2 = c.take<uint32_t>(n1);
    w.k3 = c.take<uint32_t>(n1);
    w.p0 = c.take<int32_t>(n1);
    w.p1 = c.take<int32_t>(n1);
    w.p2 = c.take<int32_t>(n1);
    w.head = c.take<int32_t>(n1);
    w.segid = c.take<int64_t>(n1);
    w.begin = c.take<int64_t>(n1);
    w.seg_of = c.take<int32_t>(n1);
    w.nseg = c.take<int64_t>(4);
    w.rws = c.take<char>(radix_workspace_bytes(n1));
    w.sws = c.take<char>(scan_workspace_bytes(n1));
    w.bytes = c.used;
    return w;
}

}  // namespace

}  // namespace halma

using namespace halma;

extern "C" {

size_t halma_clean_workspace_size(int64_t n_set) {
    if (n_set < 0) return 0;
    return carve_clean(nullptr, n_set).bytes;
}

int halma_clean_cells(int64_t n_set, int64_t n_groups, const int32_t* set_members,
                      const int32_t* set_groups, const void* x, const void* y, const void* z,
                      int pos_dtype, const void* vx, const void* vy, const void* vz,
                      int vel_dtype, const void* m, int mass_dtype, const double* centre,
                      const double* grid_start, const double* grid_g1, const double* grid_d,
                      const int32_t* grid_n, int32_t max_grid_n, int32_t* ijk,
                      double* cell_stats, int64_t* n_cells, void* ws, size_t ws_bytes,
                      void* stream) {
    if (n_set < 0 || n_groups < 0 || !n_cells) return set_error(HALMA_E_ARG, "bad argument");
    *n_cells = 0;
    if (n_set == 0) return 0;
    if (!set_members || !set_groups || !x || !y || !z || !vx || !vy || !vz || !m || !centre ||
        !grid_start || !grid_g1 || !grid_d || !grid_n || !ijk || !cell_stats)
        return set_error(HALMA_E_ARG, "null pointer");
    if (n_set > INT32_MAX || n_groups > INT32_MAX)
        return set_error(HALMA_E_LIMIT, "at most 2^31-1 members and groups");
    if (max_grid_n < 1 || max_grid_n > kMaxGridNodes)
        return set_error(HALMA_E_LIMIT, "grid of %d nodes per axis (1..%d)", max_grid_n,
                         kMaxGridNodes);
    if ((pos_dtype != HALMA_F32 && pos_dtype != HALMA_F64) ||
        (vel_dtype != HALMA_F32 && vel_dtype != HALMA_F64) ||
        (mass_dtype != HALMA_F32 && mass_dtype != HALMA_F64))
        return set_error(HALMA_E_ARG, "bad dtype");
    CleanWs w = carve_clean(ws, n_set);
    if (ws_bytes < w.bytes)
        return set_error(HALMA_E_WORKSPACE, "workspace %zu < required %zu", ws_bytes, w.bytes);
    hipStream_t st = static_cast<hipStream_t>(stream);
    Particles p{x, y, z, pos_dtype, vx, vy, vz, vel_dtype, m, mass_dtype};
    GridArgs g{grid_start, grid_g1, grid_d, grid_n, centre};
    k_clean_cells<<<blocks_for(n_set), kThreads, 0, st>>>(n_set, set_members, set_groups, p, g,
                                                          ijk, w.klo, w.khi);
    HALMA_LAUNCH_CHECK();
    // (group, cell, position) order by stable passes: low key word, high key word (only
    // when n^3 needs more than 32 bits), then group
    const uint64_t n = static_cast<uint64_t>(max_grid_n);
    const uint64_t top = n * n * n - 1;
    int kbits = 1;
    while (kbits < 64 && (top >> kbits) != 0) ++kbits;
    int gbits = 1;
    while (gbits < 32 && (int64_t{1} << gbits) < n_groups) ++gbits;
    int rc = radix_sort_pairs(w.klo, nullptr, w.k2, w.p0, n_set, std::min(kbits, 32), w.rws, st);
    if (rc) return rc;
    const int32_t* by_key = w.p0;
    if (kbits > 32) {
        k_gather_u32<<<blocks_for(n_set), kThreads, 0, st>>>(w.khi, w.p0, n_set, w.k3);
        HALMA_LAUNCH_CHECK();
        rc = radix_sort_pairs(w.k3, w.p0, w.k2, w.p2, n_set, kbits - 32, w.rws, st);
        if (rc) return rc;
        by_key = w.p2;
    }
    k_gather_u32<<<blocks_for(n_set), kThreads, 0, st>>>(
        reinterpret_cast<const uint32_t*>(set_groups), by_key, n_set, w.k3);
    HALMA_LAUNCH_CHECK();
    rc = radix_sort_pairs(w.k3, by_key, w.k2, w.p1, n_set, gbits, w.rws, st);
    if (rc) return rc;
    k_cell_heads<<<blocks_for(n_set), kThreads, 0, st>>>(n_set, w.p1, w.klo, w.khi, set_groups,
                                                         w.head);
    HALMA_LAUNCH_CHECK();
    rc = exclusive_scan_i32_to_i64(w.head, w.segid, n_set, w.nseg, w.sws, st);
    if (rc) return rc;
    k_cell_begin<<<blocks_for(n_set), kThreads, 0, st>>>(n_set, w.head, w.segid, w.begin,
                                                         w.seg_of);
    HALMA_LAUNCH_CHECK();
    HALMA_HIP(hipMemcpyAsync(n_cells, w.nseg, sizeof(int64_t), hipMemcpyDeviceToHost, st));
    HALMA_HIP(hipStreamSynchronize(st));
    const int64_t nseg = *n_cells;
    const int64_t threads = nseg * kWave;
    k_cell_stats<<<static_cast<unsigned>((threads + kThreads - 1) / kThreads), kThreads, 0, st>>>(
        nseg, n_set, w.begin, w.p1, set_members, p, cell_stats);
    HALMA_LAUNCH_CHECK();
    return 0;
}

int halma_clean_decide(int64_t n_set, const int32_t* set_members, const int32_t* set_groups,
                       const int64_t* set_target, const void* vx, const void* vy,
                       const void* vz, int vel_dtype, const double* cell_stats,
                       const int32_t* phase, const double* fac, double q_fil, double sig_fil,
                       uint8_t* control, int32_t* cleaned, void* ws, size_t ws_bytes,
                       void* stream) {
    if (n_set < 0) return set_error(HALMA_E_ARG, "bad argument");
    if (n_set == 0) return 0;
    if (!set_members || !set_groups || !set_target || !vx || !vy || !vz || !cell_stats ||
        !phase || !fac || !control || !cleaned)
        return set_error(HALMA_E_ARG, "null pointer");
    CleanWs w = carve_clean(ws, n_set);
    if (ws_bytes < w.bytes)
        return set_error(HALMA_E_WORKSPACE, "workspace %zu < required %zu", ws_bytes, w.bytes);
    Particles p{nullptr, nullptr, nullptr, HALMA_F64, vx, vy, vz, vel_dtype, nullptr, HALMA_F64};
    k_clean_decide<<<blocks_for(n_set), kThreads, 0, static_cast<hipStream_t>(stream)>>>(
        n_set, w.p1, w.seg_of, set_members, set_groups, set_target, p, cell_stats, phase, fac,
        q_fil, sig_fil, control, cleaned);
    HALMA_LAUNCH_CHECK();
    return 0;
}

}  // extern "C"
