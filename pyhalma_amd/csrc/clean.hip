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

// first k minimising |node(k) - xh| (np.argmin: the first of equal distances).  The
// estimate's window is refined by descending the distances, which are unimodal in k
// (the nodes increase); a non-finite xh has every distance non-finite: node 0.
__device__ int nearest(const GridArgs& g, int h, double xh) {
    const int64_t n = g.n[h];
    if (!isfinite(xh)) return 0;
    const double est = floor((xh - g.start[h]) / g.d[h] + 0.5);
    const int64_t kc = est >= 0.0 ? (est <= static_cast<double>(n - 1) ? static_cast<int64_t>(est)
                                                                        : n - 1)
                                  : 0;
    int64_t best = kc - 2 < 0 ? 0 : kc - 2;
    double bd = fabs(node(g, h, best) - xh);
    const int64_t hi = kc + 2 < n ? kc + 2 : n - 1;
    for (int64_t k = best + 1; k <= hi; ++k) {
        const double dist = fabs(node(g, h, k) - xh);
        if (dist < bd) {
            best = k;
            bd = dist;
        }
    }
    while (best > 0) {
        const double dist = fabs(node(g, h, best - 1) - xh);
        if (!(dist <= bd)) break;
        --best;
        bd = dist;
    }
    while (best < n - 1) {
        const double dist = fabs(node(g, h, best + 1) - xh);
        if (!(dist < bd)) break;
        ++best;
        bd = dist;
    }
    return static_cast<int>(best);
}

// per member of S: its three node indices and its cell key within the group,
// (ix n + iy) n + iz as two 32-bit words
__global__ void k_clean_cells(int64_t ns, const int32_t* __restrict__ s_mem,
                              const int32_t* __restrict__ s_grp, Particles p, GridArgs g,
                              int kbits, int composite, int32_t* __restrict__ ijk,
                              uint32_t* __restrict__ klo, uint32_t* __restrict__ khi) {
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
    if (composite) {  // group and cell in one 32-bit key
        klo[k] = (static_cast<uint32_t>(h) << kbits) | static_cast<uint32_t>(key);
        khi[k] = 0u;
    } else {
        klo[k] = static_cast<uint32_t>(key & 0xffffffffull);
        khi[k] = static_cast<uint32_t>(key >> 32);
    }
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

// Cell statistics over work items: item = (cell, chunk of kChunk consecutive members of
// the cell in sorted order), one wavefront per item, double-double partial sums folded
// per cell in item order (fixed order: deterministic).
constexpr int kChunk = 256;

__device__ __forceinline__ int64_t cell_end(const int64_t* begin, int64_t c, int64_t nseg,
                                            int64_t ns) {
    return c + 1 < nseg ? begin[c + 1] : ns;
}

__global__ void k_cell_chunks(int64_t nseg, int64_t ns, const int64_t* __restrict__ begin,
                              int32_t* __restrict__ nch) {
    const int64_t c = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (c >= nseg) return;
    nch[c] = static_cast<int32_t>((cell_end(begin, c, nseg, ns) - begin[c] + kChunk - 1) / kChunk);
}

__global__ void k_cell_items(int64_t nseg, const int64_t* __restrict__ ioff,
                             const int32_t* __restrict__ nch, int32_t* __restrict__ item_cell) {
    const int64_t c = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (c >= nseg) return;
    for (int t = 0; t < nch[c]; ++t) item_cell[ioff[c] + t] = static_cast<int32_t>(c);
}

struct CellArgs {
    int64_t nseg;
    int64_t ns;
    int64_t nitems;
    const int64_t* begin;
    const int64_t* ioff;
    const int32_t* nch;
    const int32_t* item_cell;
    const int32_t* order;
    const int32_t* s_mem;
    double* part;   // 8 doubles per item
    double* stats;  // 6 per cell: vcm x, y, z, mass, count, sigma
};

__device__ __forceinline__ void item_range(const CellArgs& a, int64_t it, int64_t& c, int64_t& b,
                                           int64_t& e) {
    c = a.item_cell[it];
    b = a.begin[c] + (it - a.ioff[c]) * kChunk;
    const int64_t ce = cell_end(a.begin, c, a.nseg, a.ns);
    e = b + kChunk < ce ? b + kChunk : ce;
}

// pass 1: sums of v m (3) and m per item
__global__ __launch_bounds__(kThreads) void k_cell_part1(CellArgs a, Particles p) {
    const int64_t it = (static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x) / kWave;
    if (it >= a.nitems) return;  // wave-uniform
    int64_t c, b, e;
    item_range(a, it, c, b, e);
    double acc[4][2] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}};
    for (int64_t k = b + __lane_id(); k < e; k += kWave) {
        const int64_t i = a.s_mem[a.order[k]];
        dd_add(acc[0][0], acc[0][1], vm_product(p.vx, p.vel_t, p.m, p.mass_t, i));
        dd_add(acc[1][0], acc[1][1], vm_product(p.vy, p.vel_t, p.m, p.mass_t, i));
        dd_add(acc[2][0], acc[2][1], vm_product(p.vz, p.vel_t, p.m, p.mass_t, i));
        dd_add(acc[3][0], acc[3][1], ldf(p.m, p.mass_t, i));
    }
    for (int q = 0; q < 4; ++q) wave_dd(acc[q][0], acc[q][1]);
    if (__lane_id() == 0) {
        double* o = a.part + 8 * it;
        for (int q = 0; q < 4; ++q) {
            o[2 * q] = acc[q][0];
            o[2 * q + 1] = acc[q][1];
        }
    }
}

// per cell: fold the items' sums, then the mass-weighted velocity, mass and count
__global__ void k_cell_fold1(CellArgs a) {
    const int64_t c = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (c >= a.nseg) return;
    const double* q0 = a.part + 8 * a.ioff[c];
    double s[4][2];
    for (int q = 0; q < 4; ++q) {
        s[q][0] = q0[2 * q];
        s[q][1] = q0[2 * q + 1];
    }
    for (int t = 1; t < a.nch[c]; ++t)
        for (int q = 0; q < 4; ++q) dd_add_dd(s[q][0], s[q][1], q0[8 * t + 2 * q], q0[8 * t + 2 * q + 1]);
    const double mass = s[3][0] + s[3][1];
    double* o = a.stats + 6 * c;
    for (int d = 0; d < 3; ++d) o[d] = (s[d][0] + s[d][1]) / mass;
    o[3] = mass;
    o[4] = static_cast<double>(cell_end(a.begin, c, a.nseg, a.ns) - a.begin[c]);
}

// pass 2: sums of |v_cell - v|^2 per item
__global__ __launch_bounds__(kThreads) void k_cell_part2(CellArgs a, Particles p) {
    const int64_t it = (static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x) / kWave;
    if (it >= a.nitems) return;  // wave-uniform
    int64_t c, b, e;
    item_range(a, it, c, b, e);
    const double* vc = a.stats + 6 * c;
    double sh = 0.0, sl = 0.0;
    for (int64_t k = b + __lane_id(); k < e; k += kWave) {
        const int64_t i = a.s_mem[a.order[k]];
        const double dx = vc[0] - ldf(p.vx, p.vel_t, i);
        const double dy = vc[1] - ldf(p.vy, p.vel_t, i);
        const double dz = vc[2] - ldf(p.vz, p.vel_t, i);
        dd_add(sh, sl, (dx * dx + dy * dy) + dz * dz);
    }
    wave_dd(sh, sl);
    if (__lane_id() == 0) {
        a.part[2 * it] = sh;
        a.part[2 * it + 1] = sl;
    }
}

__global__ void k_cell_fold2(CellArgs a) {
    const int64_t c = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (c >= a.nseg) return;
    const double* q0 = a.part + 2 * a.ioff[c];
    double h = q0[0], l = q0[1];
    for (int t = 1; t < a.nch[c]; ++t) dd_add_dd(h, l, q0[2 * t], q0[2 * t + 1]);
    double* o = a.stats + 6 * c;
    o[5] = sqrt((h + l) / 3.0 / o[4]);
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

// ------------------------------------------------------------ set selection ----
// flag of every member position (still flagged or forced, group live) and its group
// (the last g with off[g] <= k, by binary search over the offsets)
__global__ void k_sel_flags(const int64_t* __restrict__ off, int64_t G, int64_t nm,
                            const uint8_t* __restrict__ control, const uint8_t* __restrict__ live,
                            const uint8_t* __restrict__ force, int32_t* __restrict__ flag,
                            int32_t* __restrict__ grp) {
    const int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= nm) return;
    int64_t lo = 0, hi = G - 1;  // off[lo] <= k < off[hi + 1]
    while (lo < hi) {
        const int64_t mid = (lo + hi + 1) >> 1;
        if (off[mid] <= k)
            lo = mid;
        else
            hi = mid - 1;
    }
    const bool all = force && force[lo] != 0;
    flag[k] = live[lo] != 0 && (all || control[k] != 0) ? 1 : 0;
    grp[k] = static_cast<int32_t>(lo);
}

__global__ void k_sel_offsets(const int64_t* __restrict__ off, int64_t G, int64_t nm,
                              const int64_t* __restrict__ pos, const int64_t* __restrict__ total,
                              int64_t* __restrict__ set_off) {
    const int64_t g = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (g > G) return;
    set_off[g] = off[g] < nm ? pos[off[g]] : (total ? *total : 0);
}

__global__ void k_sel_scatter(int64_t nm, const int64_t* __restrict__ off,
                              const int32_t* __restrict__ members,
                              const int32_t* __restrict__ flag, const int32_t* __restrict__ grp,
                              const int64_t* __restrict__ pos,
                              const int64_t* __restrict__ set_off,
                              int32_t* __restrict__ set_members, int32_t* __restrict__ set_groups,
                              int64_t* __restrict__ set_target) {
    const int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= nm || !flag[k]) return;
    const int64_t q = pos[k];
    const int32_t g = grp[k];
    set_members[q] = members ? members[k] : static_cast<int32_t>(k);
    set_groups[q] = g;
    set_target[q] = off[g] + (q - set_off[g]);
}

struct SelWs {
    int32_t* flag;
    int32_t* grp;
    int64_t* pos;
    int64_t* total;
    void* sws;
    size_t bytes;
};

SelWs carve_sel(void* base, int64_t n) {
    Carver c(base);
    const int64_t n1 = n > 0 ? n : 1;
    SelWs w;
    w.flag = c.take<int32_t>(n1);
    w.grp = c.take<int32_t>(n1);
    w.pos = c.take<int64_t>(n1);
    w.total = c.take<int64_t>(4);
    w.sws = c.take<char>(scan_workspace_bytes(n1));
    w.bytes = c.used;
    return w;
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
    int32_t* nch;
    int64_t* ioff;
    int32_t* item_cell;
    double* part;
    int64_t* nseg;  // [0] cells, [1] items
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
    const int64_t max_items = n1 + n1 / kChunk + 1;
    w.nch = c.take<int32_t>(n1);
    w.ioff = c.take<int64_t>(n1);
    w.item_cell = c.take<int32_t>(max_items);
    w.part = c.take<double>(8 * max_items);
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
    // (group, cell, position) order by stable passes: one composite 32-bit key when the
    // group and cell bits fit, else low key word, high key word, group
    const uint64_t n = static_cast<uint64_t>(max_grid_n);
    const uint64_t top = n * n * n - 1;
    int kbits = 1;
    while (kbits < 64 && (top >> kbits) != 0) ++kbits;
    int gbits = 1;
    while (gbits < 32 && (int64_t{1} << gbits) < n_groups) ++gbits;
    const int composite = kbits + gbits <= 32;
    k_clean_cells<<<blocks_for(n_set), kThreads, 0, st>>>(n_set, set_members, set_groups, p, g,
                                                          kbits, composite, ijk, w.klo, w.khi);
    HALMA_LAUNCH_CHECK();
    int rc;
    if (composite) {
        rc = radix_sort_pairs(w.klo, nullptr, w.k2, w.p1, n_set, kbits + gbits, w.rws, st);
        if (rc) return rc;
    } else {
        rc = radix_sort_pairs(w.klo, nullptr, w.k2, w.p0, n_set, std::min(kbits, 32), w.rws, st);
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
    }
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
    k_cell_chunks<<<blocks_for(nseg), kThreads, 0, st>>>(nseg, n_set, w.begin, w.nch);
    HALMA_LAUNCH_CHECK();
    rc = exclusive_scan_i32_to_i64(w.nch, w.ioff, nseg, w.nseg + 1, w.sws, st);
    if (rc) return rc;
    int64_t nitems = 0;
    HALMA_HIP(hipMemcpyAsync(&nitems, w.nseg + 1, sizeof(int64_t), hipMemcpyDeviceToHost, st));
    HALMA_HIP(hipStreamSynchronize(st));
    k_cell_items<<<blocks_for(nseg), kThreads, 0, st>>>(nseg, w.ioff, w.nch, w.item_cell);
    HALMA_LAUNCH_CHECK();
    CellArgs a{nseg, n_set, nitems, w.begin, w.ioff, w.nch, w.item_cell, w.p1, set_members,
               w.part, cell_stats};
    const unsigned item_blocks = static_cast<unsigned>((nitems * kWave + kThreads - 1) / kThreads);
    k_cell_part1<<<item_blocks, kThreads, 0, st>>>(a, p);
    HALMA_LAUNCH_CHECK();
    k_cell_fold1<<<blocks_for(nseg), kThreads, 0, st>>>(a);
    HALMA_LAUNCH_CHECK();
    k_cell_part2<<<item_blocks, kThreads, 0, st>>>(a, p);
    HALMA_LAUNCH_CHECK();
    k_cell_fold2<<<blocks_for(nseg), kThreads, 0, st>>>(a);
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

size_t halma_clean_select_workspace_size(int64_t n_members) {
    if (n_members < 0) return 0;
    return carve_sel(nullptr, n_members).bytes;
}

int halma_clean_select(int64_t n_groups, int64_t n_members, const int64_t* offsets,
                       const int32_t* members, const uint8_t* control, const uint8_t* live,
                       const uint8_t* force, int32_t* set_members, int32_t* set_groups,
                       int64_t* set_target, int64_t* set_offsets, int64_t* n_set, void* ws,
                       size_t ws_bytes, void* stream) {
    if (n_groups < 0 || n_members < 0 || !n_set) return set_error(HALMA_E_ARG, "bad argument");
    *n_set = 0;
    if (n_groups == 0) return 0;
    if (!offsets || !live || !set_offsets || (n_members > 0 && (!control || !set_members ||
                                                                !set_groups || !set_target)))
        return set_error(HALMA_E_ARG, "null pointer");
    if (n_members > INT32_MAX || n_groups > INT32_MAX)
        return set_error(HALMA_E_LIMIT, "at most 2^31-1 members and groups");
    SelWs w = carve_sel(ws, n_members);
    if (ws_bytes < w.bytes)
        return set_error(HALMA_E_WORKSPACE, "workspace %zu < required %zu", ws_bytes, w.bytes);
    hipStream_t st = static_cast<hipStream_t>(stream);
    if (n_members > 0) {
        k_sel_flags<<<blocks_for(n_members), kThreads, 0, st>>>(offsets, n_groups, n_members,
                                                                 control, live, force, w.flag,
                                                                 w.grp);
        HALMA_LAUNCH_CHECK();
        int rc = exclusive_scan_i32_to_i64(w.flag, w.pos, n_members, w.total, w.sws, st);
        if (rc) return rc;
    }
    k_sel_offsets<<<blocks_for(n_groups + 1), kThreads, 0, st>>>(
        offsets, n_groups, n_members, w.pos, n_members > 0 ? w.total : nullptr, set_offsets);
    HALMA_LAUNCH_CHECK();
    if (n_members == 0) return 0;
    k_sel_scatter<<<blocks_for(n_members), kThreads, 0, st>>>(
        n_members, offsets, members, w.flag, w.grp, w.pos, set_offsets, set_members, set_groups,
        set_target);
    HALMA_LAUNCH_CHECK();
    HALMA_HIP(hipMemcpyAsync(n_set, w.total, sizeof(int64_t), hipMemcpyDeviceToHost, st));
    HALMA_HIP(hipStreamSynchronize(st));
    return 0;
}

}  // extern "C"
