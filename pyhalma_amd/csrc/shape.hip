// halo_shape and sigma_projections of the particle module (particle_subroutines.f90:160-461),
// the two halo-property routines its callers reach through the same module as the pair
// sums (halo_properties.py:805, :861).
//
// halo_shape: one block per group sums the FP32 terms (m x_i) x_j and m in double-double,
// rounds the tensor to FP32 and divides by the FP32 mass as the routine does; one thread
// then runs the file's Jacobi variant in FP64 (the cyclic method with its stopping rule
// relative to the initial element sum: its early stop decides the eigenvalues' last
// bits), sorts the eigenvalues into decreasing order and takes FP32 square roots.
//
// sigma_projections (one halo): the nearest grid node per axis (first of equal
// distances), FP64 atomic sums of the FP32 terms into the three projected 2D grids,
// FP32 normalisation per cell, the dispersion pass, then one block for the mean
// dispersion inside R05 per projection and the sequential FP32 sums over cells in the
// reference's loop order.  The reference's per-particle sums are FP32 OpenMP reductions
// in a schedule-dependent order; these are FP64 sums rounded to FP32.
#include <math.h>

#include <algorithm>

#include "halma_common.h"

namespace halma {

namespace {

constexpr int kThreads = 256;

inline unsigned blocks_for(int64_t n) {
    return static_cast<unsigned>((n + kThreads - 1) / kThreads);
}

// eigenvalues of a symmetric 3x3 matrix (upper triangle of a) by the Jacobi variant
__device__ void jacobi3(double a[3][3], double d[3]) {
    double b[3], z[3];
    for (int p = 0; p < 3; ++p) {
        b[p] = a[p][p];
        d[p] = b[p];
        z[p] = 0.0;
    }
    double total = 0.0;
    for (int p = 0; p < 3; ++p)
        for (int q = p; q < 3; ++q) total += fabs(a[p][q]);
    const double stop = static_cast<double>(1e-4f) * total;  // a single-precision literal there
    for (int sweep = 1; sweep <= 100; ++sweep) {
        double sm = 0.0;
        for (int p = 0; p < 2; ++p)
            for (int q = p + 1; q < 3; ++q) sm += fabs(a[p][q]);
        if (sm < stop) break;
        const double tresh = sweep < 4 ? 0.2 * sm * sm : 0.0;
        for (int p = 0; p < 2; ++p)
            for (int q = p + 1; q < 3; ++q) {
                const double g = 100.0 * fabs(a[p][q]);
                if (sweep > 4 && fabs(d[p]) + g == fabs(d[p]) && fabs(d[q]) + g == fabs(d[q])) {
                    a[p][q] = 0.0;
                } else if (fabs(a[p][q]) > tresh) {
                    double h = d[q] - d[p];
                    double t;
                    if (fabs(h) + g == fabs(h)) {
                        t = a[p][q] / h;
                    } else {
                        const double theta = 0.5 * h / a[p][q];
                        t = 1.0 / (fabs(theta) + sqrt(1.0 + theta * theta));
                        if (theta < 0.0) t = -t;
                    }
                    const double c = 1.0 / sqrt(1.0 + t * t);
                    const double s = t * c;
                    const double tau = s / (1.0 + c);
                    h = t * a[p][q];
                    z[p] -= h;
                    z[q] += h;
                    d[p] -= h;
                    d[q] += h;
                    a[p][q] = 0.0;
                    for (int j = 0; j < p; ++j) {
                        const double g1 = a[j][p], h1 = a[j][q];
                        a[j][p] = g1 - s * (h1 + g1 * tau);
                        a[j][q] = h1 + s * (g1 - h1 * tau);
                    }
                    for (int j = p + 1; j < q; ++j) {
                        const double g1 = a[p][j], h1 = a[j][q];
                        a[p][j] = g1 - s * (h1 + g1 * tau);
                        a[j][q] = h1 + s * (g1 - h1 * tau);
                    }
                    for (int j = q + 1; j < 3; ++j) {
                        const double g1 = a[p][j], h1 = a[q][j];
                        a[p][j] = g1 - s * (h1 + g1 * tau);
                        a[q][j] = h1 + s * (g1 - h1 * tau);
                    }
                }
            }
        for (int p = 0; p < 3; ++p) {
            b[p] += z[p];
            d[p] = b[p];
            z[p] = 0.0;
        }
    }
}

__device__ double block_dd_total(double hi, double lo, double* sh /* 2 * waves */) {
    for (int s = kWave / 2; s >= 1; s >>= 1) {
        const double h2 = __shfl_down(hi, s, kWave), l2 = __shfl_down(lo, s, kWave);
        dd_add_dd(hi, lo, h2, l2);
    }
    const int w = threadIdx.x / kWave, nw = blockDim.x / kWave;
    __syncthreads();
    if (__lane_id() == 0) {
        sh[2 * w] = hi;
        sh[2 * w + 1] = lo;
    }
    __syncthreads();
    double th = sh[0], tl = sh[1];
    for (int v = 1; v < nw; ++v) dd_add_dd(th, tl, sh[2 * v], sh[2 * v + 1]);
    return th + tl;
}

__global__ __launch_bounds__(kThreads) void k_halo_shape(int64_t G, const int64_t* __restrict__ off,
                                                         const float* __restrict__ x,
                                                         const float* __restrict__ y,
                                                         const float* __restrict__ z,
                                                         const float* __restrict__ m,
                                                         float* __restrict__ out) {
    __shared__ double sh[2 * (kThreads / kWave)];
    __shared__ float tensor[7];
    for (int64_t g = blockIdx.x; g < G; g += gridDim.x) {
        double acc[7][2];
        for (int q = 0; q < 7; ++q) acc[q][0] = acc[q][1] = 0.0;
        for (int64_t k = off[g] + threadIdx.x; k < off[g + 1]; k += blockDim.x) {
            const float r[3] = {x[k], y[k], z[k]};
            const float mk = m[k];
            int q = 0;
            for (int i = 0; i < 3; ++i)
                for (int j = i; j < 3; ++j, ++q)
                    dd_add(acc[q][0], acc[q][1], static_cast<double>((mk * r[i]) * r[j]));
            dd_add(acc[6][0], acc[6][1], static_cast<double>(mk));
        }
        for (int q = 0; q < 7; ++q) {
            const double t = block_dd_total(acc[q][0], acc[q][1], sh);
            if (threadIdx.x == 0) tensor[q] = static_cast<float>(t);
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            const float mt = tensor[6];
            double a[3][3];
            int q = 0;
            for (int i = 0; i < 3; ++i)
                for (int j = i; j < 3; ++j, ++q) {
                    a[i][j] = static_cast<double>(tensor[q] / mt);
                    a[j][i] = a[i][j];
                }
            double d[3];
            jacobi3(a, d);
            float e[3] = {static_cast<float>(d[0]), static_cast<float>(d[1]),
                          static_cast<float>(d[2])};
            for (int i = 0; i < 2; ++i)  // decreasing order
                for (int j = i + 1; j < 3; ++j)
                    if (e[j] >= e[i]) {
                        const float t = e[i];
                        e[i] = e[j];
                        e[j] = t;
                    }
            for (int i = 0; i < 3; ++i) out[3 * g + i] = sqrtf(e[i]);
        }
        __syncthreads();
    }
}

// --------------------------------------------------------------- projections --
struct SpIn {
    int64_t n;
    const float* x;
    const float* y;
    const float* z;
    const float* v[3];  // vx, vy, vz (relative to the halo)
    const float* m;
    const float* grid;
    int ng;
    float c[3];
};

struct SpWs {
    double* vm;    // 3 ng^2 each
    double* sd;
    double* ss;
    int32_t* q;
    float* vcm;
    float* sdf;
    float* sig;
    int32_t* cell;  // 3 per particle: the cell of each projection
};

__device__ __forceinline__ int nearest_node(const float* grid, int ng, float v) {
    int best = 0;
    float bd = fabsf(grid[0] - v);
    for (int k = 1; k < ng; ++k) {
        const float d = fabsf(grid[k] - v);
        if (d < bd) {
            bd = d;
            best = k;
        }
    }
    return best;
}

__global__ void k_sp_zero(SpWs w, int64_t cells) {
    const int64_t c = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (c >= cells) return;
    w.vm[c] = 0.0;
    w.sd[c] = 0.0;
    w.ss[c] = 0.0;
    w.q[c] = 0;
}

__global__ void k_sp_cells(SpIn in, SpWs w) {
    const int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= in.n) return;
    const int ix = nearest_node(in.grid, in.ng, in.x[k] - in.c[0]);
    const int iy = nearest_node(in.grid, in.ng, in.y[k] - in.c[1]);
    const int iz = nearest_node(in.grid, in.ng, in.z[k] - in.c[2]);
    const int64_t n2 = static_cast<int64_t>(in.ng) * in.ng;
    // projection a: cells (iy, iz), (ix, iz), (ix, iy) as the routine indexes them
    const int64_t cell[3] = {static_cast<int64_t>(iy) * in.ng + iz,
                             static_cast<int64_t>(ix) * in.ng + iz,
                             static_cast<int64_t>(ix) * in.ng + iy};
    const float mk = in.m[k];
    for (int a = 0; a < 3; ++a) {
        const int64_t c = a * n2 + cell[a];
        w.cell[3 * k + a] = static_cast<int32_t>(c);
        atomicAdd(w.vm + c, static_cast<double>(in.v[a][k] * mk));
        atomicAdd(w.sd + c, static_cast<double>(mk));
        atomicAdd(w.q + c, 1);
    }
}

__global__ void k_sp_vcm(SpWs w, int64_t cells) {
    const int64_t c = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (c >= cells) return;
    const float sd = static_cast<float>(w.sd[c]);
    const float vm = static_cast<float>(w.vm[c]);
    w.vcm[c] = sd != 0.0f ? vm / sd : vm;
    w.sdf[c] = sd;
}

__global__ void k_sp_disp(SpIn in, SpWs w) {
    const int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= in.n) return;
    for (int a = 0; a < 3; ++a) {
        const int32_t c = w.cell[3 * k + a];
        const float dv = in.v[a][k] - w.vcm[c];
        atomicAdd(w.ss + c, static_cast<double>(dv * dv));
    }
}

__global__ void k_sp_sig(SpWs w, int64_t cells) {
    const int64_t c = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (c >= cells) return;
    const float ss = static_cast<float>(w.ss[c]);
    w.sig[c] = w.q[c] != 0 ? sqrtf(ss / static_cast<float>(w.q[c])) : ss;
}

// one block: mean dispersion inside R05 per projection, then the plane sums (thread 0)
__global__ __launch_bounds__(kThreads) void k_sp_final(SpIn in, SpWs w, float R05x, float R05y,
                                                       float R05z, float ll, float* out) {
    __shared__ double sh[2 * (kThreads / kWave)];
    __shared__ double tot[6];
    const float R05[3] = {R05x, R05y, R05z};
    double acc[3][2] = {{0, 0}, {0, 0}, {0, 0}};
    double cnt[3] = {0, 0, 0};
    for (int64_t k = threadIdx.x; k < in.n; k += blockDim.x) {
        const float dx = in.x[k] - in.c[0], dy = in.y[k] - in.c[1], dz = in.z[k] - in.c[2];
        const float dist[3] = {sqrtf(dy * dy + dz * dz), sqrtf(dx * dx + dz * dz),
                               sqrtf(dx * dx + dy * dy)};
        for (int a = 0; a < 3; ++a)
            if (dist[a] < R05[a]) {
                dd_add(acc[a][0], acc[a][1], static_cast<double>(w.sig[w.cell[3 * k + a]]));
                cnt[a] += 1.0;
            }
    }
    for (int a = 0; a < 3; ++a) {
        const double s = block_dd_total(acc[a][0], acc[a][1], sh);
        const double c = block_dd_total(cnt[a], 0.0, sh);
        if (threadIdx.x == 0) {
            tot[a] = s;
            tot[3 + a] = c;
        }
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    float sig05[3], vs[3], lam[3];
    for (int a = 0; a < 3; ++a) {
        const int c = static_cast<int>(tot[3 + a]);
        sig05[a] = c > 0 ? static_cast<float>(tot[a]) / static_cast<float>(c) : 0.0f;
    }
    const int ng = in.ng;
    const int64_t n2 = static_cast<int64_t>(ng) * ng;
    // planes in the routine's order: XY (projection z), XZ (y), YZ (x); the outer loop runs
    // over the second cell index, the inner over the first
    const int order[3] = {2, 1, 0};
    for (int t = 0; t < 3; ++t) {
        const int a = order[t];
        float sumV = 0.0f, sumS = 0.0f, sumU = 0.0f, sumD = 0.0f;
        const float lim = R05[a] + 2.0f * ll;
        for (int o = 0; o < ng; ++o)
            for (int i = 0; i < ng; ++i) {
                const float gi = in.grid[i], go = in.grid[o];
                const float rbin = sqrtf(gi * gi + go * go);
                if (rbin < lim) {
                    const int64_t c = a * n2 + static_cast<int64_t>(i) * ng + o;
                    const float v = w.vcm[c], s = w.sig[c], m = w.sdf[c];
                    sumV = sumV + (v * v) * m;
                    sumS = sumS + (s * s) * m;
                    sumU = sumU + (m * rbin) * fabsf(v);
                    sumD = sumD + (m * rbin) * sqrtf(v * v + s * s);
                }
            }
        vs[a] = sumS > 0.0f ? sqrtf(sumV / sumS) : 0.0f;
        lam[a] = sumD > 0.0f ? sumU / sumD : 0.0f;
    }
    out[0] = sig05[0];
    out[1] = sig05[1];
    out[2] = sig05[2];
    out[3] = (vs[0] + vs[1] + vs[2]) / 3.0f;
    out[4] = (lam[0] + lam[1] + lam[2]) / 3.0f;
}

SpWs carve_sp(void* base, int64_t n, int ng) {
    Carver c(base);
    const int64_t cells = 3 * static_cast<int64_t>(std::max(ng, 1)) * std::max(ng, 1);
    SpWs w;
    w.vm = c.take<double>(cells);
    w.sd = c.take<double>(cells);
    w.ss = c.take<double>(cells);
    w.q = c.take<int32_t>(cells);
    w.vcm = c.take<float>(cells);
    w.sdf = c.take<float>(cells);
    w.sig = c.take<float>(cells);
    w.cell = c.take<int32_t>(3 * std::max<int64_t>(n, 1));
    return w;
}

size_t carve_sp_bytes(int64_t n, int ng) {
    Carver c(nullptr);
    const int64_t cells = 3 * static_cast<int64_t>(std::max(ng, 1)) * std::max(ng, 1);
    c.take<double>(cells);
    c.take<double>(cells);
    c.take<double>(cells);
    c.take<int32_t>(cells);
    c.take<float>(cells);
    c.take<float>(cells);
    c.take<float>(cells);
    c.take<int32_t>(3 * std::max<int64_t>(n, 1));
    return c.used;
}

}  // namespace

}  // namespace halma

using namespace halma;

extern "C" {

int halma_halo_shape(int64_t n_groups, const int64_t* offsets, const float* x, const float* y,
                     const float* z, const float* mass, float* out, void* stream) {
    if (n_groups < 0) return set_error(HALMA_E_ARG, "negative count");
    if (n_groups == 0) return 0;
    if (!offsets || !x || !y || !z || !mass || !out) return set_error(HALMA_E_ARG, "null pointer");
    const unsigned gb = static_cast<unsigned>(std::min<int64_t>(n_groups, 65536));
    k_halo_shape<<<gb, kThreads, 0, static_cast<hipStream_t>(stream)>>>(n_groups, offsets, x, y, z,
                                                                        mass, out);
    HALMA_LAUNCH_CHECK();
    return 0;
}

size_t halma_sigma_projections_workspace_size(int64_t n, int n_cell) {
    if (n < 0 || n_cell < 0) return 0;
    return carve_sp_bytes(n, n_cell);
}

int halma_sigma_projections(int64_t n, const float* x, const float* y, const float* z,
                            const float* vx, const float* vy, const float* vz, const float* mass,
                            const float* grid, int n_cell, float cx, float cy, float cz,
                            float R05x, float R05y, float R05z, float ll, float* out, void* ws,
                            size_t ws_bytes, void* stream) {
    if (n < 0 || n_cell < 1) return set_error(HALMA_E_ARG, "bad argument (n_cell >= 1)");
    if (!grid || !out || (n > 0 && (!x || !y || !z || !vx || !vy || !vz || !mass)))
        return set_error(HALMA_E_ARG, "null pointer");
    if (n > INT32_MAX / 3 || static_cast<int64_t>(n_cell) * n_cell * 3 > INT32_MAX)
        return set_error(HALMA_E_LIMIT, "too many particles or cells");
    if (ws_bytes < carve_sp_bytes(n, n_cell))
        return set_error(HALMA_E_WORKSPACE, "workspace %zu < required %zu", ws_bytes,
                         carve_sp_bytes(n, n_cell));
    hipStream_t st = static_cast<hipStream_t>(stream);
    SpWs w = carve_sp(ws, n, n_cell);
    SpIn in{n, x, y, z, {vx, vy, vz}, mass, grid, n_cell, {cx, cy, cz}};
    const int64_t cells = 3 * static_cast<int64_t>(n_cell) * n_cell;
    k_sp_zero<<<blocks_for(cells), kThreads, 0, st>>>(w, cells);
    HALMA_LAUNCH_CHECK();
    if (n > 0) {
        k_sp_cells<<<blocks_for(n), kThreads, 0, st>>>(in, w);
        HALMA_LAUNCH_CHECK();
    }
    k_sp_vcm<<<blocks_for(cells), kThreads, 0, st>>>(w, cells);
    HALMA_LAUNCH_CHECK();
    if (n > 0) {
        k_sp_disp<<<blocks_for(n), kThreads, 0, st>>>(in, w);
        HALMA_LAUNCH_CHECK();
    }
    k_sp_sig<<<blocks_for(cells), kThreads, 0, st>>>(w, cells);
    HALMA_LAUNCH_CHECK();
    k_sp_final<<<1, kThreads, 0, st>>>(in, w, R05x, R05y, R05z, ll, out);
    HALMA_LAUNCH_CHECK();
    return 0;
}

}  // extern "C"
