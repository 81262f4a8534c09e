// Group lists and per-group reductions.
//
// groups_csr: FoF labels -> groups of at least min_size members in CSR form, ordered by
//   label (the minimum member index, so in the order fof.py:34-77 emits them) with the
//   members of each group ascending (fof.py:76 sorts them): a stable radix sort of
//   (label, index) pairs, segment heads, a size filter and two scans.
// group_reduce: one workgroup per group, FP64 with double-double accumulation (so the
//   sums are within an ulp or so of correctly rounded, independent of thread count):
//   M, centre of mass, mass-weighted velocity, furthest member from the centre, bounding
//   box, and a flag saying whether every per-cell mass sum of the group is exact in FP64
//   (halo_properties.py:16-60, 98-111).
// density_peak: halo_properties.py:518-556.  Nearest node of np.linspace(min - ll,
//   max + ll, 50) per axis (first index among equal distances), masses summed per node in
//   an LDS hash table (a dense 50^3 grid in the workspace when a group has too many
//   distinct nodes), the first node in (ix, iy, iz) order with the largest mass strictly
//   above zero wins (centre node 25 if none).  When the group's sums are not exact in
//   FP64 every node within the rounding band of the maximum is re-summed sequentially in
//   member order, which is the reference's own order, so the argmax agrees bit-for-bit.
#include <math.h>

#include "halma_common.h"
#include "radix.h"
#include "scan.h"
#include "staging.h"

namespace halma {

namespace {

constexpr int kThreads = 256;
constexpr int kBig = 1024;                  // threads of the per-group kernels
constexpr int kMaxWaves = kBig / kWave;     // LDS slots of the block reductions
constexpr int kNodes = 50;
constexpr int kCells = kNodes * kNodes * kNodes;
constexpr int kTable = 4096;      // LDS hash slots per group
constexpr int kTableFill = 3072;  // distinct nodes accepted before falling back
constexpr int kDenseSlots = 32;   // concurrent dense-grid fallbacks
constexpr int kMaxCandidates = 256;

__global__ void k_fill_ints(int* p, int n) {
    if (static_cast<int>(threadIdx.x) < n) p[threadIdx.x] = 0;
}

inline unsigned blocks_for(int64_t n, int t = kThreads) {
    return static_cast<unsigned>((n + t - 1) / t);
}

// The ldt<DT> form for a call's arrays: 1 all FP32, 2 all FP64, 0 mixed.
inline int common_dtype(int a, int b, int c) {
    return a == b && b == c ? (a == HALMA_F32 ? 1 : 2) : 0;
}

// ------------------------------------------------------------------ CSR build ----
__global__ void k_heads(const uint32_t* __restrict__ sk, int64_t n, int32_t* __restrict__ head) {
    const int64_t p = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (p < n) head[p] = (p == 0 || sk[p] != sk[p - 1]) ? 1 : 0;
}

__global__ void k_seg_begin(const int32_t* __restrict__ head, const int64_t* __restrict__ segid,
                            int64_t n, int64_t* __restrict__ seg_begin) {
    const int64_t p = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (p < n && head[p]) seg_begin[segid[p]] = p;
}

__global__ void k_seg_keep(const int64_t* __restrict__ seg_begin, const int64_t* __restrict__ nseg_p,
                           int64_t n, int min_size, int32_t* __restrict__ keep,
                           int64_t* __restrict__ ksize) {
    const int64_t s = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    const int64_t nseg = *nseg_p;
    if (s >= n) return;
    if (s >= nseg) {
        keep[s] = 0;
        ksize[s] = 0;
        return;
    }
    const int64_t end = s + 1 < nseg ? seg_begin[s + 1] : n;
    const int64_t size = end - seg_begin[s];
    const bool k = size >= min_size;
    keep[s] = k ? 1 : 0;
    ksize[s] = k ? size : 0;
}

// segid is the exclusive scan of the head flags, so the segment holding position p is
// segid[p] + head[p] - 1 (the inclusive count of heads up to p, minus one)
__global__ void k_emit(const int32_t* __restrict__ sv, const int32_t* __restrict__ head,
                       const int64_t* __restrict__ segid, const int64_t* __restrict__ seg_begin,
                       const int32_t* __restrict__ keep, const int64_t* __restrict__ moff,
                       int64_t n, int32_t* __restrict__ members) {
    const int64_t p = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const int64_t s = segid[p] + head[p] - 1;
    if (keep[s]) members[moff[s] + (p - seg_begin[s])] = sv[p];
}

__global__ void k_offsets(const int32_t* __restrict__ keep, const int64_t* __restrict__ gidx,
                          const int64_t* __restrict__ moff, const int64_t* __restrict__ totals,
                          int64_t n, int64_t* __restrict__ offsets) {
    const int64_t s = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (s == 0) offsets[totals[1]] = totals[2];
    if (s < n && keep[s]) offsets[gidx[s]] = moff[s];
}

__global__ void k_copy_i64(const int64_t* __restrict__ src, int64_t* __restrict__ dst, int k) {
    if (threadIdx.x < k) dst[threadIdx.x] = src[threadIdx.x];
}

struct CsrWs {
    uint32_t* sk;
    int32_t* sv;
    void* radix_ws;
    int32_t* head;
    int64_t* segid;
    int64_t* seg_begin;
    int32_t* keep;
    int64_t* ksize;
    int64_t* gidx;
    int64_t* moff;
    int64_t* totals;  // [0] nseg, [1] n_groups, [2] n_members
    void* scan_ws;
    size_t bytes;
};

CsrWs carve_csr(void* base, int64_t n) {
    Carver c(base);
    const int64_t nn = n > 0 ? n : 1;
    CsrWs w;
    w.sk = c.take<uint32_t>(nn);
    w.sv = c.take<int32_t>(nn);
    w.radix_ws = c.take<char>(radix_workspace_bytes(nn));
    w.head = c.take<int32_t>(nn);
    w.segid = c.take<int64_t>(nn);
    w.seg_begin = c.take<int64_t>(nn);
    w.keep = c.take<int32_t>(nn);
    w.ksize = c.take<int64_t>(nn);
    w.gidx = c.take<int64_t>(nn);
    w.moff = c.take<int64_t>(nn);
    w.totals = c.take<int64_t>(8);
    w.scan_ws = c.take<char>(scan_workspace_bytes(nn));
    w.bytes = c.used;
    return w;
}

// ---------------------------------------------------------------- reductions ----
constexpr int kRedFields = 16;  // must match HALMA_GROUP_FIELDS

// block-wide double-double sum (result valid in thread 0)
__device__ void block_dd_sum(double& hi, double& lo, double* sh) {
    for (int s = 32; s >= 1; s >>= 1) {
        const double h2 = __shfl_down(hi, s, kWave);
        const double l2 = __shfl_down(lo, s, kWave);
        dd_add_dd(hi, lo, h2, l2);
    }
    const int w = threadIdx.x / kWave;
    __syncthreads();
    if (__lane_id() == 0) {
        sh[2 * w] = hi;
        sh[2 * w + 1] = lo;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        hi = sh[0];
        lo = sh[1];
        for (int k = 1; k < static_cast<int>(blockDim.x) / kWave; ++k)
            dd_add_dd(hi, lo, sh[2 * k], sh[2 * k + 1]);
    }
    __syncthreads();
}

__device__ double block_max(double v, double* sh) {
    for (int s = 32; s >= 1; s >>= 1) v = fmax(v, __shfl_down(v, s, kWave));
    const int w = threadIdx.x / kWave;
    __syncthreads();
    if (__lane_id() == 0) sh[w] = v;
    __syncthreads();
    double r = sh[0];
    for (int k = 1; k < static_cast<int>(blockDim.x) / kWave; ++k) r = fmax(r, sh[k]);
    __syncthreads();
    return r;
}

__device__ double block_min(double v, double* sh) { return -block_max(-v, sh); }

__device__ int block_min_i(int v, int* sh) {
    for (int s = 32; s >= 1; s >>= 1) v = min(v, __shfl_down(v, s, kWave));
    const int w = threadIdx.x / kWave;
    __syncthreads();
    if (__lane_id() == 0) sh[w] = v;
    __syncthreads();
    int r = sh[0];
    for (int k = 1; k < static_cast<int>(blockDim.x) / kWave; ++k) r = min(r, sh[k]);
    __syncthreads();
    return r;
}

__device__ int block_sum_i(int v, int* sh) {
    for (int s = 32; s >= 1; s >>= 1) v += __shfl_down(v, s, kWave);
    const int w = threadIdx.x / kWave;
    __syncthreads();
    if (__lane_id() == 0) sh[w] = v;
    __syncthreads();
    int r = 0;
    for (int k = 0; k < static_cast<int>(blockDim.x) / kWave; ++k) r += sh[k];
    __syncthreads();
    return r;
}

// exponent of the smallest power of two dividing |m| (mass granularity); large if m == 0
__device__ __forceinline__ int mass_grain_exp(double m) {
    if (m == 0.0) return 4096;
    const uint64_t b = static_cast<uint64_t>(__double_as_longlong(m));
    const int e = static_cast<int>((b >> 52) & 0x7ff);
    const uint64_t mant = b & 0xfffffffffffffull;
    if (e == 0) return -1074 + __builtin_ctzll(mant);  // subnormal
    return e - 1075 + __builtin_ctzll(mant | (1ull << 52));
}

struct Src {
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

// All of a group's block reductions in one LDS round: every wave reduces its eight
// double-double sums (shuffle tree), three minima, three maxima and the grain exponent,
// writes them to LDS, and thread q then folds quantity q over the waves in wave order (the
// order of block_dd_sum / block_max), so the results are those of the separate
// reductions with two barriers instead of forty-five.
constexpr int kGrpQ = 8 * 2 + 6 + 1;

// Only the first nw wavefronts hold members (the others skip the trees and are left out
// of the fold), so a small group pays for the trees of the wavefronts it uses.
__device__ void block_group_reduce(double (&acc)[8][2], double (&lo)[3], double (&hi)[3],
                                   int& gexp, double* sh /* kGrpQ * kMaxWaves */, int nw) {
    const int w = threadIdx.x / kWave;
    if (w < nw) {
        for (int k = 0; k < 8; ++k)
            for (int s = 32; s >= 1; s >>= 1) {
                const double h2 = __shfl_down(acc[k][0], s, kWave);
                const double l2 = __shfl_down(acc[k][1], s, kWave);
                dd_add_dd(acc[k][0], acc[k][1], h2, l2);
            }
        for (int d = 0; d < 3; ++d)
            for (int s = 32; s >= 1; s >>= 1) {
                lo[d] = fmin(lo[d], __shfl_down(lo[d], s, kWave));
                hi[d] = fmax(hi[d], __shfl_down(hi[d], s, kWave));
            }
        for (int s = 32; s >= 1; s >>= 1) gexp = min(gexp, __shfl_down(gexp, s, kWave));
    }
    __syncthreads();
    if (w < nw && __lane_id() == 0) {
        double* o = sh + w * kGrpQ;
        for (int k = 0; k < 8; ++k) {
            o[2 * k] = acc[k][0];
            o[2 * k + 1] = acc[k][1];
        }
        for (int d = 0; d < 3; ++d) {
            o[16 + d] = lo[d];
            o[19 + d] = hi[d];
        }
        o[22] = static_cast<double>(gexp);
    }
    __syncthreads();
    const int q = threadIdx.x;
    if (q < 8) {
        double h = sh[2 * q], l = sh[2 * q + 1];
        for (int k = 1; k < nw; ++k) dd_add_dd(h, l, sh[k * kGrpQ + 2 * q], sh[k * kGrpQ + 2 * q + 1]);
        sh[nw * kGrpQ + 2 * q] = h;
        sh[nw * kGrpQ + 2 * q + 1] = l;
    } else if (q < 11) {
        double v = sh[16 + (q - 8)];
        for (int k = 1; k < nw; ++k) v = fmin(v, sh[k * kGrpQ + 16 + (q - 8)]);
        sh[nw * kGrpQ + 16 + (q - 8)] = v;
    } else if (q < 14) {
        double v = sh[19 + (q - 11)];
        for (int k = 1; k < nw; ++k) v = fmax(v, sh[k * kGrpQ + 19 + (q - 11)]);
        sh[nw * kGrpQ + 19 + (q - 11)] = v;
    } else if (q == 14) {
        double v = sh[22];
        for (int k = 1; k < nw; ++k) v = fmin(v, sh[k * kGrpQ + 22]);
        sh[nw * kGrpQ + 22] = v;
    }
    __syncthreads();
    const double* r = sh + nw * kGrpQ;
    for (int k = 0; k < 8; ++k) {
        acc[k][0] = r[2 * k];
        acc[k][1] = r[2 * k + 1];
    }
    for (int d = 0; d < 3; ++d) {
        lo[d] = r[16 + d];
        hi[d] = r[19 + d];
    }
    gexp = static_cast<int>(r[22]);
    __syncthreads();
}

template <int DT>
__global__ __launch_bounds__(kBig) void k_group_reduce(Src s, const int64_t* __restrict__ off,
                                                           const int32_t* __restrict__ mem,
                                                           int64_t G,
                                                           const double* __restrict__ centres,
                                                           double* __restrict__ out) {
    __shared__ double sh[kGrpQ * (kMaxWaves + 1)];
    __shared__ double cen[3];
    for (int64_t g = blockIdx.x; g < G; g += gridDim.x) {
        const int64_t b = off[g], e = off[g + 1];
        double acc[8][2];  // M, Mx, My, Mz, Mvx, Mvy, Mvz, sum |m|
        for (int k = 0; k < 8; ++k) acc[k][0] = acc[k][1] = 0.0;
        double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
        int gexp = 4096;
        // wavefronts that take members: enough for four or more members per lane
        const int64_t ng = e - b;
        const int nw = static_cast<int>(std::min<int64_t>(
            std::max<int64_t>((ng + 4 * kWave - 1) / (4 * kWave), 1), blockDim.x / kWave));
        const int nthr = nw * kWave;
        // four members per thread in flight: their indices, then mass and position, then
        // velocity; every accumulator takes the members in the same order as one member
        // at a time
        for (int64_t k0 = b + threadIdx.x; threadIdx.x < nthr && k0 < e; k0 += 4 * nthr) {
            int32_t iv[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int64_t k = k0 + u * nthr;
                iv[u] = k < e ? mem[k] : -1;
            }
            double mu[4], f[4][3];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                if (iv[u] < 0) continue;
                mu[u] = ldt<DT>(s.m, s.mass_t, iv[u]);
                f[u][0] = ldt<DT>(s.x, s.pos_t, iv[u]);
                f[u][1] = ldt<DT>(s.y, s.pos_t, iv[u]);
                f[u][2] = ldt<DT>(s.z, s.pos_t, iv[u]);
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                if (iv[u] < 0) continue;
                const double m = mu[u];
                dd_add(acc[0][0], acc[0][1], m);
#pragma unroll
                for (int d = 0; d < 3; ++d) {
                    dd_add(acc[1 + d][0], acc[1 + d][1], m * f[u][d]);
                    lo[d] = fmin(lo[d], f[u][d]);
                    hi[d] = fmax(hi[d], f[u][d]);
                }
                dd_add(acc[7][0], acc[7][1], fabs(m));
                gexp = min(gexp, mass_grain_exp(m));
            }
            if (!s.vx) continue;
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                if (iv[u] < 0) continue;
                f[u][0] = ldt<DT>(s.vx, s.vel_t, iv[u]);
                f[u][1] = ldt<DT>(s.vy, s.vel_t, iv[u]);
                f[u][2] = ldt<DT>(s.vz, s.vel_t, iv[u]);
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                if (iv[u] < 0) continue;
#pragma unroll
                for (int d = 0; d < 3; ++d) dd_add(acc[4 + d][0], acc[4 + d][1], mu[u] * f[u][d]);
            }
        }
        block_group_reduce(acc, lo, hi, gexp, sh, nw);
        const double* bl = lo;
        const double* bh = hi;
        if (threadIdx.x == 0) {
            double* o = out + g * kRedFields;
            const double M = acc[0][0] + acc[0][1];
            o[0] = M;
            for (int d = 0; d < 3; ++d) {
                const double num = acc[1 + d][0] + acc[1 + d][1];
                o[1 + d] = M > 0.0 ? num / M : 0.0;
                const double vn = acc[4 + d][0] + acc[4 + d][1];
                o[4 + d] = M > 0.0 ? vn / M : 0.0;
                o[8 + d] = bl[d];
                o[11 + d] = bh[d];
            }
            const double total_abs = (acc[7][0] + acc[7][1]) * (1.0 + 1e-12);
            o[14] = total_abs;
            // every partial sum of any subset is an integer multiple of 2^gexp below 2^53
            // of those units -> exact in FP64 whatever the order
            o[15] = (gexp < 4096 && total_abs < ldexp(1.0, 53 + gexp)) || total_abs == 0.0
                        ? 1.0 : 0.0;
            for (int d = 0; d < 3; ++d) cen[d] = centres ? centres[3 * g + d] : o[1 + d];
        }
        __syncthreads();
        double r2 = -1.0;
        for (int64_t k0 = b + threadIdx.x; k0 < e; k0 += 4 * static_cast<int64_t>(blockDim.x)) {
            int32_t iv[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int64_t k = k0 + u * static_cast<int64_t>(blockDim.x);
                iv[u] = k < e ? mem[k] : -1;
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                if (iv[u] < 0) continue;
                const double dx = cen[0] - ldt<DT>(s.x, s.pos_t, iv[u]);
                const double dy = cen[1] - ldt<DT>(s.y, s.pos_t, iv[u]);
                const double dz = cen[2] - ldt<DT>(s.z, s.pos_t, iv[u]);
                r2 = fmax(r2, (dx * dx + dy * dy) + dz * dz);
            }
        }
        r2 = block_max(r2, sh);
        if (threadIdx.x == 0) out[g * kRedFields + 7] = r2 > 0.0 ? sqrt(r2) : 0.0;
        __syncthreads();
    }
}

// ------------------------------------------------------------- density peak ----
struct Grid1 {
    double start, step, stop;
    __device__ __forceinline__ double node(int k) const {
        return k == kNodes - 1 ? stop : static_cast<double>(k) * step + start;
    }
};

// first k minimising |node(k) - v| (np.argmin over the 50 nodes)
__device__ int nearest_node(const Grid1& g, double v) {
    int best = 0;
    if (g.step > 0.0 && isfinite(v)) {
        const double t = (v - g.start) / g.step;
        int k0 = static_cast<int>(fmin(fmax(floor(t + 0.5), 0.0), double(kNodes - 1)));
        const int lo = max(k0 - 2, 0), hi = min(k0 + 2, kNodes - 1);
        best = lo;
        double bd = fabs(g.node(lo) - v);
        for (int k = lo + 1; k <= hi; ++k) {
            const double d = fabs(g.node(k) - v);
            if (d < bd) {
                bd = d;
                best = k;
            }
        }
        // the distances are unimodal in k; the window is trusted unless a neighbour
        // outside it ties or beats it
        const bool left_ok = best > lo || lo == 0 || fabs(g.node(lo - 1) - v) > bd;
        const bool right_ok = best < hi || hi == kNodes - 1 || fabs(g.node(hi + 1) - v) >= bd;
        if (left_ok && right_ok) return best;
    }
    double bd = fabs(g.node(0) - v);
    best = 0;
    for (int k = 1; k < kNodes; ++k) {
        const double d = fabs(g.node(k) - v);
        if (d < bd) {
            bd = d;
            best = k;
        }
    }
    return best;
}

struct PeakArgs {
    Src s;
    const int64_t* off;
    const int32_t* mem;
    int64_t G;
    const double* red;  // group_reduce output (bounding box, exactness flag)
    double ll;
    double* out;        // 3 per group
    int* over_list;     // groups that overflowed the LDS table
    int* over_count;
    double* dense;      // kDenseSlots x kCells
    int* err;
};

__device__ __forceinline__ void grids_of(const PeakArgs& a, int64_t g, Grid1* gr) {
    const double* r = a.red + g * kRedFields;
    for (int d = 0; d < 3; ++d) {
        const double start = r[8 + d] - a.ll, stop = r[11 + d] + a.ll;
        gr[d].start = start;
        gr[d].stop = stop;
        gr[d].step = (stop - start) / static_cast<double>(kNodes - 1);
    }
}

template <int DT = 0>
__device__ __forceinline__ int cell_of(const PeakArgs& a, const Grid1* gr, int64_t i) {
    const int ix = nearest_node(gr[0], ldt<DT>(a.s.x, a.s.pos_t, i));
    const int iy = nearest_node(gr[1], ldt<DT>(a.s.y, a.s.pos_t, i));
    const int iz = nearest_node(gr[2], ldt<DT>(a.s.z, a.s.pos_t, i));
    return (ix * kNodes + iy) * kNodes + iz;
}

// Exact per-node sum of one node in member order (the reference's serial loop order).
__device__ double exact_node_sum(const PeakArgs& a, const Grid1* gr, int64_t b, int64_t e,
                                 int node) {
    double s = 0.0;
    for (int64_t k = b; k < e; ++k) {
        const int64_t i = a.mem[k];
        if (cell_of(a, gr, i) == node) s += ldf(a.s.m, a.s.mass_t, i);
    }
    return s;
}

__device__ void write_peak(const PeakArgs& a, const Grid1* gr, int64_t g, int node) {
    int ix = kNodes / 2, iy = kNodes / 2, iz = kNodes / 2;
    if (node >= 0) {
        ix = node / (kNodes * kNodes);
        iy = (node / kNodes) % kNodes;
        iz = node % kNodes;
    }
    a.out[3 * g + 0] = gr[0].node(ix);
    a.out[3 * g + 1] = gr[1].node(iy);
    a.out[3 * g + 2] = gr[2].node(iz);
}

// Given (node, approximate mass) pairs held by the block (a thread-strided view through
// get()), picks the winner: exact sums -> first node with the maximum; otherwise nodes
// within the band of the maximum are re-summed in member order by one thread.
template <class Get>
__device__ int pick_peak(const PeakArgs& a, const Grid1* gr, int64_t g, int64_t b, int64_t e,
                         int nslots, Get get, double* shd, int* shi, int* cand) {
    double vmax = -INFINITY;
    for (int k = threadIdx.x; k < nslots; k += blockDim.x) {
        int node;
        double v;
        if (get(k, node, v)) vmax = fmax(vmax, v);
    }
    vmax = block_max(vmax, shd);
    if (!(vmax > 0.0)) return -1;
    const double* r = a.red + g * kRedFields;
    const bool exact = r[15] != 0.0;
    const double band = exact ? 0.0 : 2.0 * static_cast<double>(e - b) * 0x1p-53 * r[14];
    int first = INT32_MAX;
    int ncand = 0;
    for (int k = threadIdx.x; k < nslots; k += blockDim.x) {
        int node;
        double v;
        if (get(k, node, v) && v >= vmax - 2.0 * band) {
            first = min(first, node);
            if (!exact) {
                const int at = atomicAdd(&shi[kMaxWaves], 1);
                if (at < kMaxCandidates) cand[at] = node;
            }
        }
    }
    first = block_min_i(first, shi);
    if (exact) return first;
    __syncthreads();
    ncand = shi[kMaxWaves];
    if (ncand == 1) return first;
    __shared__ int winner;
    if (threadIdx.x == 0) {
        if (ncand > kMaxCandidates) atomicOr(a.err, 8);
        const int nc = min(ncand, kMaxCandidates);
        double best = 0.0;
        int bn = -1;
        for (int c = 0; c < nc; ++c) {
            const double s = exact_node_sum(a, gr, b, e, cand[c]);
            if (s > best || (s == best && bn >= 0 && cand[c] < bn)) {
                best = s;
                bn = cand[c];
            }
        }
        winner = bn;
    }
    __syncthreads();
    return winner;
}

template <int DT>
__global__ __launch_bounds__(kBig) void k_density_peak(PeakArgs a) {
    __shared__ int skey[kTable];
    __shared__ double sval[kTable];
    __shared__ double shd[kMaxWaves];
    __shared__ int shi[kMaxWaves + 1];
    __shared__ int cand[kMaxCandidates];
    __shared__ int used;
    __shared__ int overflow;
    for (int64_t g = blockIdx.x; g < a.G; g += gridDim.x) {
        for (int k = threadIdx.x; k < kTable; k += blockDim.x) {
            skey[k] = -1;
            sval[k] = 0.0;
        }
        if (threadIdx.x == 0) {
            used = 0;
            overflow = 0;
            shi[kMaxWaves] = 0;
        }
        __syncthreads();
        Grid1 gr[3];
        grids_of(a, g, gr);
        const int64_t b = a.off[g], e = a.off[g + 1];
        for (int64_t k0 = b + threadIdx.x; k0 < e; k0 += 4 * static_cast<int64_t>(blockDim.x)) {
            // a group with too many distinct nodes goes to the dense grid: stop at once
            // (probing a full table would cost up to kTable probes per member)
            if (__hip_atomic_load(&overflow, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP))
                break;
            // four members per thread in flight: their indices, then their fields
            int32_t iv[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int64_t k = k0 + u * static_cast<int64_t>(blockDim.x);
                iv[u] = k < e ? a.mem[k] : -1;
            }
            int nodes[4];
            double ms[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                if (iv[u] < 0) continue;
                nodes[u] = cell_of<DT>(a, gr, iv[u]);
                ms[u] = ldt<DT>(a.s.m, a.s.mass_t, iv[u]);
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                if (iv[u] < 0 || __hip_atomic_load(&overflow, __ATOMIC_RELAXED,
                                                   __HIP_MEMORY_SCOPE_WORKGROUP))
                    break;  // past the group's end, or the group goes to the dense grid
                const int node = nodes[u];
                uint32_t h = (static_cast<uint32_t>(node) * 2654435761u) & (kTable - 1);
                for (int probe = 0; probe < kTable; ++probe) {
                    const int old = atomicCAS(&skey[h], -1, node);
                    if (old == -1 || old == node) {
                        if (old == -1 && atomicAdd(&used, 1) >= kTableFill)
                            __hip_atomic_store(&overflow, 1, __ATOMIC_RELAXED,
                                               __HIP_MEMORY_SCOPE_WORKGROUP);
                        atomicAdd(&sval[h], ms[u]);
                        break;
                    }
                    h = (h + 1) & (kTable - 1);
                }
            }
        }
        __syncthreads();
        if (overflow) {
            if (threadIdx.x == 0) {
                const int at = atomicAdd(a.over_count, 1);
                a.over_list[at] = static_cast<int>(g);
            }
            __syncthreads();
            continue;
        }
        auto get = [&](int k, int& node, double& v) {
            node = skey[k];
            v = sval[k];
            return node >= 0;
        };
        const int node = pick_peak(a, gr, g, b, e, kTable, get, shd, shi, cand);
        if (threadIdx.x == 0) write_peak(a, gr, g, node);
        __syncthreads();
    }
}

// Dense-grid fallback for groups with more distinct nodes than the LDS table holds:
// one group per block, kDenseSlots groups per round, in three kernels so that the
// zeroing (plain stores), the accumulation (device atomics) and the read-back (plain
// loads) are separated by kernel boundaries.
__global__ __launch_bounds__(kThreads) void k_dense_zero(PeakArgs a, int first) {
    const int o = first + blockIdx.y;
    if (o >= *a.over_count) return;
    double* grid = a.dense + static_cast<int64_t>(blockIdx.y) * kCells;
    for (int k = blockIdx.x * kThreads + threadIdx.x; k < kCells; k += gridDim.x * kThreads)
        grid[k] = 0.0;
}

__global__ __launch_bounds__(kThreads) void k_dense_accum(PeakArgs a, int first) {
    const int o = first + blockIdx.y;
    if (o >= *a.over_count) return;
    const int64_t g = a.over_list[o];
    double* grid = a.dense + static_cast<int64_t>(blockIdx.y) * kCells;
    Grid1 gr[3];
    grids_of(a, g, gr);
    const int64_t b = a.off[g], e = a.off[g + 1];
    for (int64_t k = b + static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x; k < e;
         k += static_cast<int64_t>(gridDim.x) * kThreads) {
        const int64_t i = a.mem[k];
        atomicAdd(grid + cell_of(a, gr, i), ldf(a.s.m, a.s.mass_t, i));
    }
}

__global__ __launch_bounds__(kBig) void k_dense_pick(PeakArgs a, int first) {
    __shared__ double shd[kMaxWaves];
    __shared__ int shi[kMaxWaves + 1];
    __shared__ int cand[kMaxCandidates];
    const int o = first + blockIdx.x;
    if (o >= *a.over_count) return;
    const int64_t g = a.over_list[o];
    const double* grid = a.dense + static_cast<int64_t>(blockIdx.x) * kCells;
    if (threadIdx.x == 0) shi[kMaxWaves] = 0;
    __syncthreads();
    Grid1 gr[3];
    grids_of(a, g, gr);
    const int64_t b = a.off[g], e = a.off[g + 1];
    auto get = [&](int k, int& node, double& v) {
        node = k;
        v = grid[k];
        return v != 0.0;
    };
    const int node = pick_peak(a, gr, g, b, e, kCells, get, shd, shi, cand);
    if (threadIdx.x == 0) write_peak(a, gr, g, node);
}

// ------------------------------------------------------------- member gather ----
// Particle data arrive as separate arrays in particle order, and group members are
// particle indices in no spatial order, so gathering seven arrays member by member costs
// one memory line per value.  Packing each particle into one 32-byte (FP32) or 64-byte
// (FP64) record first makes the gather one line per member; the member-ordered arrays
// it writes are then read contiguously by every per-group kernel.
template <typename T, int DT>
__global__ void k_pack_records(const void* x, const void* y, const void* z, int pos_t,
                               const void* vx, const void* vy, const void* vz, int vel_t,
                               const void* m, int mass_t, int64_t n, T* __restrict__ rec) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    T* r = rec + 8 * i;
    r[0] = static_cast<T>(ldt<DT>(x, pos_t, i));
    r[1] = static_cast<T>(ldt<DT>(y, pos_t, i));
    r[2] = static_cast<T>(ldt<DT>(z, pos_t, i));
    r[3] = static_cast<T>(ldt<DT>(m, mass_t, i));
    r[4] = vx ? static_cast<T>(ldt<DT>(vx, vel_t, i)) : T(0);
    r[5] = vx ? static_cast<T>(ldt<DT>(vy, vel_t, i)) : T(0);
    r[6] = vx ? static_cast<T>(ldt<DT>(vz, vel_t, i)) : T(0);
    r[7] = T(0);
}

template <typename T>
__global__ void k_gather_records(const T* __restrict__ rec, const int32_t* __restrict__ mem,
                                 int64_t nm, T* __restrict__ ox, T* __restrict__ oy,
                                 T* __restrict__ oz, T* __restrict__ om, T* __restrict__ ovx,
                                 T* __restrict__ ovy, T* __restrict__ ovz) {
    const int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= nm) return;
    const T* r = rec + 8 * static_cast<int64_t>(mem[k]);
    ox[k] = r[0];
    oy[k] = r[1];
    oz[k] = r[2];
    om[k] = r[3];
    if (ovx) {
        ovx[k] = r[4];
        ovy[k] = r[5];
        ovz[k] = r[6];
    }
}

template <typename T>
void pack_records(const void* x, const void* y, const void* z, int pos_t, const void* vx,
                  const void* vy, const void* vz, int vel_t, const void* m, int mass_t,
                  int64_t n, T* rec, hipStream_t st) {
    const int dt = common_dtype(pos_t, mass_t, vx ? vel_t : mass_t);
    if (dt == 1)
        k_pack_records<T, 1><<<blocks_for(n), kThreads, 0, st>>>(x, y, z, pos_t, vx, vy, vz,
                                                                  vel_t, m, mass_t, n, rec);
    else if (dt == 2)
        k_pack_records<T, 2><<<blocks_for(n), kThreads, 0, st>>>(x, y, z, pos_t, vx, vy, vz,
                                                                  vel_t, m, mass_t, n, rec);
    else
        k_pack_records<T, 0><<<blocks_for(n), kThreads, 0, st>>>(x, y, z, pos_t, vx, vy, vz,
                                                                  vel_t, m, mass_t, n, rec);
}

}  // namespace

size_t groups_csr_workspace_bytes(int64_t n) { return carve_csr(nullptr, n).bytes; }

int groups_csr_run(const int32_t* labels, int64_t n, int min_size, int32_t* members,
                   int64_t* offsets, int64_t* n_groups, int64_t* n_members, void* ws,
                   hipStream_t st) {
    CsrWs w = carve_csr(ws, n);
    int bits = 1;
    while ((int64_t(1) << bits) < n) ++bits;
    int rc = radix_sort_pairs(reinterpret_cast<const uint32_t*>(labels), nullptr, w.sk, w.sv, n,
                              bits, w.radix_ws, st);
    if (rc) return rc;
    k_heads<<<blocks_for(n), kThreads, 0, st>>>(w.sk, n, w.head);
    HALMA_LAUNCH_CHECK();
    rc = exclusive_scan_i32_to_i64(w.head, w.segid, n, w.totals + 0, w.scan_ws, st);
    if (rc) return rc;
    k_seg_begin<<<blocks_for(n), kThreads, 0, st>>>(w.head, w.segid, n, w.seg_begin);
    HALMA_LAUNCH_CHECK();
    k_seg_keep<<<blocks_for(n), kThreads, 0, st>>>(w.seg_begin, w.totals + 0, n, min_size,
                                                   w.keep, w.ksize);
    HALMA_LAUNCH_CHECK();
    rc = exclusive_scan_i32_to_i64(w.keep, w.gidx, n, w.totals + 1, w.scan_ws, st);
    if (rc) return rc;
    rc = exclusive_scan_i64(w.ksize, w.moff, n, w.totals + 2, w.scan_ws, st);
    if (rc) return rc;
    k_emit<<<blocks_for(n), kThreads, 0, st>>>(w.sv, w.head, w.segid, w.seg_begin, w.keep,
                                               w.moff, n, members);
    HALMA_LAUNCH_CHECK();
    k_offsets<<<blocks_for(n), kThreads, 0, st>>>(w.keep, w.gidx, w.moff, w.totals, n, offsets);
    HALMA_LAUNCH_CHECK();
    long long* h = pinned_words(4);
    if (!h) return set_error(HALMA_E_STATE, "pinned host buffer unavailable");
    HALMA_HIP(hipMemcpyAsync(h, w.totals, 3 * sizeof(int64_t), hipMemcpyDeviceToHost, st));
    HALMA_HIP(hipStreamSynchronize(st));
    *n_groups = h[1];
    *n_members = h[2];
    return 0;
}

struct PeakWs {
    int* over_list;
    int* over_count;
    double* dense;
    int* err;
    size_t bytes;
};

PeakWs carve_peak(void* base, int64_t G) {
    Carver c(base);
    PeakWs w;
    w.over_list = c.take<int>(G > 0 ? G : 1);
    w.over_count = c.take<int>(64);
    w.err = c.take<int>(64);
    w.dense = c.take<double>(static_cast<size_t>(kDenseSlots) * kCells);
    w.bytes = c.used;
    return w;
}

}  // namespace halma

using namespace halma;

namespace {
Src make_src(const void* x, const void* y, const void* z, int pos_t, const void* vx,
             const void* vy, const void* vz, int vel_t, const void* m, int mass_t) {
    Src s;
    s.x = x;
    s.y = y;
    s.z = z;
    s.pos_t = pos_t;
    s.vx = vx;
    s.vy = vy;
    s.vz = vz;
    s.vel_t = vel_t;
    s.m = m;
    s.mass_t = mass_t;
    return s;
}

bool bad_dtype(int t) { return t != HALMA_F32 && t != HALMA_F64; }
}  // namespace

extern "C" {

size_t halma_groups_workspace_size(int64_t n) {
    if (n < 0) return 0;
    return groups_csr_workspace_bytes(n);
}

int halma_groups_from_labels(const int32_t* labels, int64_t n, int32_t min_size,
                             int32_t* members, int64_t* offsets, int64_t* n_groups,
                             int64_t* n_members, void* ws, size_t ws_bytes, void* stream) {
    if (n < 0 || !n_groups || !n_members) return set_error(HALMA_E_ARG, "bad argument");
    *n_groups = 0;
    *n_members = 0;
    if (n == 0) return 0;
    if (!labels || !members || !offsets) return set_error(HALMA_E_ARG, "null pointer");
    if (n > INT32_MAX) return set_error(HALMA_E_LIMIT, "at most 2^31-1 particles");
    if (ws_bytes < groups_csr_workspace_bytes(n))
        return set_error(HALMA_E_WORKSPACE, "workspace too small");
    return groups_csr_run(labels, n, min_size < 1 ? 1 : min_size, members, offsets, n_groups,
                          n_members, ws, static_cast<hipStream_t>(stream));
}

int halma_group_reduce(int64_t n_groups, const int64_t* offsets, const int32_t* members,
                       const void* x, const void* y, const void* z, int pos_dtype,
                       const void* vx, const void* vy, const void* vz, int vel_dtype,
                       const void* m, int mass_dtype, const double* centres, double* out,
                       void* stream) {
    if (n_groups < 0) return set_error(HALMA_E_ARG, "negative group count");
    if (n_groups == 0) return 0;
    if (!offsets || !members || !x || !y || !z || !m || !out)
        return set_error(HALMA_E_ARG, "null pointer");
    if (bad_dtype(pos_dtype) || bad_dtype(mass_dtype) || (vx && bad_dtype(vel_dtype)))
        return set_error(HALMA_E_ARG, "bad dtype");
    if (vx && (!vy || !vz)) return set_error(HALMA_E_ARG, "velocities need vx, vy and vz");
    Src s = make_src(x, y, z, pos_dtype, vx, vy, vz, vel_dtype, m, mass_dtype);
    const unsigned grid = static_cast<unsigned>(n_groups < 65536 ? n_groups : 65536);
    const int dt = common_dtype(pos_dtype, mass_dtype, vx ? vel_dtype : mass_dtype);
    hipStream_t st = static_cast<hipStream_t>(stream);
    if (dt == 1)
        k_group_reduce<1><<<grid, kBig, 0, st>>>(s, offsets, members, n_groups, centres, out);
    else if (dt == 2)
        k_group_reduce<2><<<grid, kBig, 0, st>>>(s, offsets, members, n_groups, centres, out);
    else
        k_group_reduce<0><<<grid, kBig, 0, st>>>(s, offsets, members, n_groups, centres, out);
    HALMA_LAUNCH_CHECK();
    return 0;
}

size_t halma_density_peak_workspace_size(int64_t n_groups) {
    if (n_groups < 0) return 0;
    return carve_peak(nullptr, n_groups).bytes;
}

int halma_density_peak(int64_t n_groups, const int64_t* offsets, const int32_t* members,
                       const void* x, const void* y, const void* z, int pos_dtype,
                       const void* m, int mass_dtype, double ll, const double* reduced,
                       double* out, void* ws, size_t ws_bytes, void* stream) {
    if (n_groups < 0) return set_error(HALMA_E_ARG, "negative group count");
    if (n_groups == 0) return 0;
    if (!offsets || !members || !x || !y || !z || !m || !reduced || !out)
        return set_error(HALMA_E_ARG, "null pointer");
    if (bad_dtype(pos_dtype) || bad_dtype(mass_dtype)) return set_error(HALMA_E_ARG, "bad dtype");
    PeakWs w = carve_peak(ws, n_groups);
    if (ws_bytes < w.bytes) return set_error(HALMA_E_WORKSPACE, "workspace too small");
    hipStream_t st = static_cast<hipStream_t>(stream);
    PeakArgs a;
    a.s = make_src(x, y, z, pos_dtype, nullptr, nullptr, nullptr, HALMA_F64, m, mass_dtype);
    a.off = offsets;
    a.mem = members;
    a.G = n_groups;
    a.red = reduced;
    a.ll = ll;
    a.out = out;
    a.over_list = w.over_list;
    a.over_count = w.over_count;
    a.dense = w.dense;
    a.err = w.err;
    k_fill_ints<<<1, 64, 0, st>>>(w.over_count, 1);
    HALMA_LAUNCH_CHECK();
    k_fill_ints<<<1, 64, 0, st>>>(w.err, 1);
    HALMA_LAUNCH_CHECK();
    const unsigned grid = static_cast<unsigned>(n_groups < 65536 ? n_groups : 65536);
    const int dt = common_dtype(pos_dtype, mass_dtype, mass_dtype);
    if (dt == 1)
        k_density_peak<1><<<grid, kBig, 0, st>>>(a);
    else if (dt == 2)
        k_density_peak<2><<<grid, kBig, 0, st>>>(a);
    else
        k_density_peak<0><<<grid, kBig, 0, st>>>(a);
    HALMA_LAUNCH_CHECK();
    // the first round of the dense fallback is queued whatever the overflow count (its
    // blocks return at once when no group overflowed), so that one read-back gives both
    // the count and the error flags; further rounds, when needed, cost a second one
    auto dense_round = [&](int first) -> int {
        k_dense_zero<<<dim3(64, kDenseSlots), kThreads, 0, st>>>(a, first);
        HALMA_LAUNCH_CHECK();
        k_dense_accum<<<dim3(64, kDenseSlots), kThreads, 0, st>>>(a, first);
        HALMA_LAUNCH_CHECK();
        k_dense_pick<<<kDenseSlots, kBig, 0, st>>>(a, first);
        HALMA_LAUNCH_CHECK();
        return 0;
    };
    int rc = dense_round(0);
    if (rc) return rc;
    long long* h = pinned_words(2);
    if (!h) return set_error(HALMA_E_STATE, "pinned host buffer unavailable");
    h[0] = h[1] = 0;
    HALMA_HIP(hipMemcpyAsync(h, w.over_count, sizeof(int), hipMemcpyDeviceToHost, st));
    HALMA_HIP(hipMemcpyAsync(h + 1, w.err, sizeof(int), hipMemcpyDeviceToHost, st));
    HALMA_HIP(hipStreamSynchronize(st));
    const int nover = static_cast<int>(h[0] & 0xffffffff);
    if (nover > kDenseSlots) {
        for (int first = kDenseSlots; first < nover; first += kDenseSlots) {
            rc = dense_round(first);
            if (rc) return rc;
        }
        h[1] = 0;
        HALMA_HIP(hipMemcpyAsync(h + 1, w.err, sizeof(int), hipMemcpyDeviceToHost, st));
        HALMA_HIP(hipStreamSynchronize(st));
    }
    if (h[1] & 8)
        return set_error(HALMA_E_LIMIT, "density peak: more than %d near-tied nodes",
                         kMaxCandidates);
    return 0;
}

size_t halma_gather_workspace_size(int64_t n_particles, int dtype) {
    if (n_particles < 0) return 0;
    const size_t elt = dtype == HALMA_F64 ? sizeof(double) : sizeof(float);
    return align256(static_cast<size_t>(n_particles > 0 ? n_particles : 1) * 8 * elt);
}

int halma_gather_members(const int32_t* members, int64_t n_members, int64_t n_particles,
                         const void* x, const void* y, const void* z, int pos_dtype,
                         const void* vx, const void* vy, const void* vz, int vel_dtype,
                         const void* m, int mass_dtype, int out_dtype, void* ox, void* oy,
                         void* oz, void* om, void* ovx, void* ovy, void* ovz, void* ws,
                         size_t ws_bytes, void* stream) {
    if (n_members < 0 || n_particles < 0) return set_error(HALMA_E_ARG, "negative size");
    if (n_members == 0) return 0;
    if (!members || !x || !y || !z || !m || !ox || !oy || !oz || !om)
        return set_error(HALMA_E_ARG, "null pointer");
    if (bad_dtype(pos_dtype) || bad_dtype(mass_dtype) || bad_dtype(out_dtype) ||
        (vx && bad_dtype(vel_dtype)))
        return set_error(HALMA_E_ARG, "bad dtype");
    if (vx && (!vy || !vz || !ovx || !ovy || !ovz))
        return set_error(HALMA_E_ARG, "velocities need all three input and output arrays");
    if (ws_bytes < halma_gather_workspace_size(n_particles, out_dtype))
        return set_error(HALMA_E_WORKSPACE, "workspace too small");
    hipStream_t st = static_cast<hipStream_t>(stream);
    if (out_dtype == HALMA_F64) {
        double* rec = static_cast<double*>(ws);
        pack_records(x, y, z, pos_dtype, vx, vy, vz, vel_dtype, m, mass_dtype, n_particles, rec,
                     st);
        HALMA_LAUNCH_CHECK();
        k_gather_records<double><<<blocks_for(n_members), kThreads, 0, st>>>(
            rec, members, n_members, static_cast<double*>(ox), static_cast<double*>(oy),
            static_cast<double*>(oz), static_cast<double*>(om), static_cast<double*>(vx ? ovx : nullptr),
            static_cast<double*>(ovy), static_cast<double*>(ovz));
    } else {
        float* rec = static_cast<float*>(ws);
        pack_records(x, y, z, pos_dtype, vx, vy, vz, vel_dtype, m, mass_dtype, n_particles, rec,
                     st);
        HALMA_LAUNCH_CHECK();
        k_gather_records<float><<<blocks_for(n_members), kThreads, 0, st>>>(
            rec, members, n_members, static_cast<float*>(ox), static_cast<float*>(oy),
            static_cast<float*>(oz), static_cast<float*>(om), static_cast<float*>(vx ? ovx : nullptr),
            static_cast<float*>(ovy), static_cast<float*>(ovz));
    }
    HALMA_LAUNCH_CHECK();
    return 0;
}

}  // extern "C"
