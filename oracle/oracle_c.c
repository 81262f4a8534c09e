/*
 * CPU oracle for the halo-finding hot path -- TEST INFRASTRUCTURE ONLY.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
 * this library, and only as the checker or the timed CPU baseline.  The product
 * (pyhalma_amd, libhalma.so) never calls it.
 *
 * Contents (every function cites the reference lines it restates):
 *   orc_pot_sum_f32      FP32 restatement of particle_subroutines.f90:466-514
 *                        (brute_force_binding_energy), OpenMP over test particles,
 *                        sequential inner sum, the same three-coordinate skip rule.
 *   orc_pot_sum_f64      the FP64 direct-sum oracle of the same sum on FP32-cast
 *                        inputs (halo_gas.py:172-178 casts every input to float32).
 *                        Fixed summation order (256 strided partials + pairwise tree),
 *                        no contraction, IEEE sqrt and divide: its bits are the
 *                        contract the GPU FP64 band re-test reproduces.
 *   orc_pot_soft_f64 /   softened potential and acceleration for the centring
 *   orc_acc_soft_f64     workload (an extension; no reference counterpart).
 *   orc_fof_labels       friends-of-friends labels (connected components of the
 *                        d^2 <= ll^2 graph of fof.py:18-79 with scipy's FP64
 *                        predicate ((dx*dx+dy*dy)+dz*dz) <= ll*ll), canonical
 *                        label = minimum particle index of the component.
 *
 * Build: oracle/Makefile (gcc -O3 -fopenmp -ffp-contract=off).  The file must be
 * compiled without FMA contraction: the FP64 sums are compared bit-for-bit.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* particle_subroutines.f90:466-514.  Loop nest, predicate (:499-501), distance
 * (:502-504) and accumulation (:505) in FP32, as the Fortran source states. */
void orc_pot_sum_f32(int nthreads, int64_t ntotal, const float *m, const float *x,
                     const float *y, const float *z, int64_t ntest, const float *tx,
                     const float *ty, const float *tz, float *out) {
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(static)
#endif
    for (int64_t i = 0; i < ntest; ++i) {
        float acc = 0.0f;
        const float xi = tx[i], yi = ty[i], zi = tz[i];
        for (int64_t j = 0; j < ntotal; ++j) {
            if (x[j] != xi && y[j] != yi && z[j] != zi) {
                const float dx = x[j] - xi, dy = y[j] - yi, dz = z[j] - zi;
                const float r = sqrtf(dx * dx + dy * dy + dz * dz);
                acc = acc + m[j] / r;
            }
        }
        out[i] = acc;
    }
}

/* FP64 direct sum on FP32 inputs.  Same predicate as particle_subroutines.f90:499-501;
 * per included pair d2 = (dx*dx + dy*dy) + dz*dz, t = m / sqrt(d2).  Summation order
 * (fixed, so the GPU FP64 re-test can reproduce the bits with one 256-thread block per
 * test particle): ORD_W = 256 strided partial sums L[k] = sum over j = k, k+256, ...
 * ascending, each starting from 0.0, then the pairwise tree L[k] += L[k+s] for
 * s = 128, 64, ..., 1 (k < s); the result is L[0].  sqrt and divide are IEEE correctly
 * rounded. */
#define ORD_W 256
static inline double pot_one_f64(int64_t ntotal, const float *m, const float *x,
                                 const float *y, const float *z, float xi, float yi,
                                 float zi) {
    double lane[ORD_W];
    const double xd = xi, yd = yi, zd = zi;
    for (int k = 0; k < ORD_W; ++k) {
        double acc = 0.0;
        for (int64_t j = k; j < ntotal; j += ORD_W) {
            if (x[j] != xi && y[j] != yi && z[j] != zi) {
                const double dx = (double)x[j] - xd;
                const double dy = (double)y[j] - yd;
                const double dz = (double)z[j] - zd;
                const double d2 = (dx * dx + dy * dy) + dz * dz;
                acc = acc + (double)m[j] / sqrt(d2);
            }
        }
        lane[k] = acc;
    }
    for (int s = ORD_W / 2; s >= 1; s >>= 1)
        for (int k = 0; k < s; ++k) lane[k] = lane[k] + lane[k + s];
    return lane[0];
}

void orc_pot_sum_f64(int nthreads, int64_t ntotal, const float *m, const float *x,
                     const float *y, const float *z, int64_t ntest, const float *tx,
                     const float *ty, const float *tz, double *out) {
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 64)
#endif
    for (int64_t i = 0; i < ntest; ++i)
        out[i] = pot_one_f64(ntotal, m, x, y, z, tx[i], ty[i], tz[i]);
}

/* Softened potential sum_j m_j / sqrt(r^2 + eps^2) (all pairs with r > 0 or eps > 0,
 * using the same three-coordinate skip rule as the unsoftened sum), FP64. */
void orc_pot_soft_f64(int nthreads, int64_t ntotal, const float *m, const float *x,
                      const float *y, const float *z, int64_t ntest, const double *tx,
                      const double *ty, const double *tz, double eps, double *out) {
    const double e2 = eps * eps;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 16)
#endif
    for (int64_t i = 0; i < ntest; ++i) {
        double acc = 0.0;
        for (int64_t j = 0; j < ntotal; ++j) {
            const double dx = (double)x[j] - tx[i];
            const double dy = (double)y[j] - ty[i];
            const double dz = (double)z[j] - tz[i];
            if (dx != 0.0 && dy != 0.0 && dz != 0.0) {
                const double d2 = (dx * dx + dy * dy) + dz * dz + e2;
                acc += (double)m[j] / sqrt(d2);
            }
        }
        out[i] = acc;
    }
}

/* Gradient of S_i = sum_j m_j (r_ij^2 + eps^2)^(-1/2) with respect to the test position:
 * dS_i/dx_i = sum_j m_j (x_j - x_i) (r^2 + eps^2)^(-3/2).  FP64. */
void orc_acc_soft_f64(int nthreads, int64_t ntotal, const float *m, const float *x,
                      const float *y, const float *z, int64_t ntest, const double *tx,
                      const double *ty, const double *tz, double eps, double *gx,
                      double *gy, double *gz) {
    const double e2 = eps * eps;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 16)
#endif
    for (int64_t i = 0; i < ntest; ++i) {
        double ax = 0.0, ay = 0.0, az = 0.0;
        for (int64_t j = 0; j < ntotal; ++j) {
            const double dx = (double)x[j] - tx[i];
            const double dy = (double)y[j] - ty[i];
            const double dz = (double)z[j] - tz[i];
            if (dx != 0.0 && dy != 0.0 && dz != 0.0) {
                const double d2 = (dx * dx + dy * dy) + dz * dz + e2;
                const double inv = 1.0 / sqrt(d2);
                const double w = (double)m[j] * inv * inv * inv;
                ax += w * dx;
                ay += w * dy;
                az += w * dz;
            }
        }
        gx[i] = ax;
        gy[i] = ay;
        gz[i] = az;
    }
}

/* ---------------------------------------------------------------- FoF ---- */

static int64_t uf_find(int64_t *parent, int64_t a) {
    while (parent[a] != a) {
        parent[a] = parent[parent[a]];
        a = parent[a];
    }
    return a;
}

static void uf_union(int64_t *parent, int64_t a, int64_t b) {
    a = uf_find(parent, a);
    b = uf_find(parent, b);
    if (a == b) return;
    if (a < b) parent[b] = a;
    else parent[a] = b;
}

typedef struct {
    int64_t key;
    int64_t idx;
} key_idx;

static int cmp_key_idx(const void *pa, const void *pb) {
    const key_idx *a = (const key_idx *)pa, *b = (const key_idx *)pb;
    if (a->key != b->key) return a->key < b->key ? -1 : 1;
    return a->idx < b->idx ? -1 : (a->idx > b->idx);
}

/* Connected components of the graph with an edge wherever
 * ((dx*dx + dy*dy) + dz*dz) <= ll*ll in FP64 (fof.py:20,40,61 through scipy's
 * cKDTree; non-periodic, inclusive).  Plain cell list with cells of side >= ll so
 * that every link lies in the 27-cell neighbourhood; every candidate pair is tested
 * unless the two particles already share a root.  labels[i] = min index of i's
 * component (singletons label themselves).  Coordinates are FP64 (callers upcast
 * FP32 exactly).  Returns 0, or -1 on allocation failure / bad input. */
int orc_fof_labels(int64_t n, const double *x, const double *y, const double *z,
                   double ll, int64_t *labels) {
    if (n <= 0) return 0;
    if (!(ll > 0.0)) {
        /* ll <= 0: only coincident points link when ll == 0 (d2 <= 0) */
        for (int64_t i = 0; i < n; ++i) labels[i] = i;
        if (ll < 0.0) return 0;
    }
    const double ll2 = ll * ll;
    double lo[3] = {x[0], y[0], z[0]}, hi[3] = {x[0], y[0], z[0]};
    for (int64_t i = 1; i < n; ++i) {
        if (x[i] < lo[0]) lo[0] = x[i];
        if (x[i] > hi[0]) hi[0] = x[i];
        if (y[i] < lo[1]) lo[1] = y[i];
        if (y[i] > hi[1]) hi[1] = y[i];
        if (z[i] < lo[2]) lo[2] = z[i];
        if (z[i] > hi[2]) hi[2] = z[i];
    }
    /* cell side: at least ll (plus a relative margin so rounding never shrinks it) */
    double side = ll > 0.0 ? ll * (1.0 + 1e-9) : 1.0;
    int64_t nc[3];
    for (int d = 0; d < 3; ++d) { /* keep the key within 64 bits; coarsen cells */
        const double need = (hi[d] - lo[d]) / 1.5e6;
        if (need > side) side = need;
    }
    for (int d = 0; d < 3; ++d) nc[d] = (int64_t)floor((hi[d] - lo[d]) / side) + 1;
    key_idx *ki = (key_idx *)malloc(sizeof(key_idx) * (size_t)n);
    int64_t *parent = (int64_t *)malloc(sizeof(int64_t) * (size_t)n);
    int64_t *cx = (int64_t *)malloc(sizeof(int64_t) * 3 * (size_t)n);
    if (!ki || !parent || !cx) {
        free(ki);
        free(parent);
        free(cx);
        return -1;
    }
    for (int64_t i = 0; i < n; ++i) {
        int64_t c0 = (int64_t)floor((x[i] - lo[0]) / side);
        int64_t c1 = (int64_t)floor((y[i] - lo[1]) / side);
        int64_t c2 = (int64_t)floor((z[i] - lo[2]) / side);
        cx[3 * i] = c0;
        cx[3 * i + 1] = c1;
        cx[3 * i + 2] = c2;
        ki[i].key = (c0 * nc[1] + c1) * nc[2] + c2;
        ki[i].idx = i;
        parent[i] = i;
    }
    qsort(ki, (size_t)n, sizeof(key_idx), cmp_key_idx);
    /* cell start table: binary search over the sorted keys */
    for (int64_t s = 0; s < n; ++s) {
        const int64_t i = ki[s].idx;
        const int64_t c0 = cx[3 * i], c1 = cx[3 * i + 1], c2 = cx[3 * i + 2];
        for (int64_t a = c0 - 1; a <= c0 + 1; ++a) {
            if (a < 0 || a >= nc[0]) continue;
            for (int64_t b = c1 - 1; b <= c1 + 1; ++b) {
                if (b < 0 || b >= nc[1]) continue;
                for (int64_t c = c2 - 1; c <= c2 + 1; ++c) {
                    if (c < 0 || c >= nc[2]) continue;
                    const int64_t key = (a * nc[1] + b) * nc[2] + c;
                    /* lower_bound of key */
                    int64_t l = 0, h = n;
                    while (l < h) {
                        int64_t mid = (l + h) >> 1;
                        if (ki[mid].key < key) l = mid + 1;
                        else h = mid;
                    }
                    for (int64_t t = l; t < n && ki[t].key == key; ++t) {
                        const int64_t j = ki[t].idx;
                        if (j <= i) continue;
                        if (uf_find(parent, i) == uf_find(parent, j)) continue;
                        const double dx = x[i] - x[j], dy = y[i] - y[j], dz = z[i] - z[j];
                        const double d2 = (dx * dx + dy * dy) + dz * dz;
                        if (d2 <= ll2) uf_union(parent, i, j);
                    }
                }
            }
        }
    }
    /* with hooks from larger to smaller root, every root is its component's minimum */
    for (int64_t i = 0; i < n; ++i) labels[i] = uf_find(parent, i);
    free(ki);
    free(parent);
    free(cx);
    return 0;
}

/* The same labels as orc_fof_labels, computed for catalogues with dense clumps (a
 * checker fast enough for millions of particles).  Particles are binned in sub-cells of
 * side ll / sqrt(3) * (1 - 1e-6): any two particles of one sub-cell are closer than ll
 * even after the FP64 rounding of the cell index and of d^2, so a sub-cell is one
 * connected piece.  Two sub-cells within Chebyshev distance 2 (the farthest a link of
 * length ll can reach) are joined as soon as one of their pairs passes the same FP64
 * predicate ((dx*dx + dy*dy) + dz*dz) <= ll*ll; pairs of sub-cells already in one
 * component are not tested.  labels[i] = min index of i's component.  Falls back to
 * orc_fof_labels for ll <= 0 or for a grid of more than 2^21 sub-cells per axis.
 * Returns 0, or -1 on allocation failure. */
int orc_fof_labels_cells(int64_t n, const double *x, const double *y, const double *z,
                         double ll, int64_t *labels) {
    if (n <= 0) return 0;
    if (!(ll > 0.0)) return orc_fof_labels(n, x, y, z, ll, labels);
    const double ll2 = ll * ll;
    double lo[3] = {x[0], y[0], z[0]}, hi[3] = {x[0], y[0], z[0]};
    for (int64_t i = 1; i < n; ++i) {
        if (x[i] < lo[0]) lo[0] = x[i];
        if (x[i] > hi[0]) hi[0] = x[i];
        if (y[i] < lo[1]) lo[1] = y[i];
        if (y[i] > hi[1]) hi[1] = y[i];
        if (z[i] < lo[2]) lo[2] = z[i];
        if (z[i] > hi[2]) hi[2] = z[i];
    }
    const double side = ll / sqrt(3.0) * (1.0 - 1e-6);
    int64_t nc[3];
    for (int d = 0; d < 3; ++d) {
        const double cells = floor((hi[d] - lo[d]) / side) + 1.0;
        if (!(cells < 2097152.0)) return orc_fof_labels(n, x, y, z, ll, labels);
        nc[d] = (int64_t)cells;
    }
    key_idx *ki = (key_idx *)malloc(sizeof(key_idx) * (size_t)n);
    if (!ki) return -1;
    for (int64_t i = 0; i < n; ++i) {
        const int64_t c0 = (int64_t)floor((x[i] - lo[0]) / side);
        const int64_t c1 = (int64_t)floor((y[i] - lo[1]) / side);
        const int64_t c2 = (int64_t)floor((z[i] - lo[2]) / side);
        ki[i].key = (c0 * nc[1] + c1) * nc[2] + c2;
        ki[i].idx = i;
    }
    qsort(ki, (size_t)n, sizeof(key_idx), cmp_key_idx);
    /* cells: runs of one key in the sorted order */
    int64_t ncell = 0;
    for (int64_t s = 0; s < n; ++s)
        if (s == 0 || ki[s].key != ki[s - 1].key) ++ncell;
    int64_t *start = (int64_t *)malloc(sizeof(int64_t) * (size_t)(ncell + 1));
    int64_t *ckey = (int64_t *)malloc(sizeof(int64_t) * (size_t)ncell);
    int64_t *parent = (int64_t *)malloc(sizeof(int64_t) * (size_t)ncell);
    int64_t *cmin = (int64_t *)malloc(sizeof(int64_t) * (size_t)ncell);
    if (!start || !ckey || !parent || !cmin) {
        free(ki);
        free(start);
        free(ckey);
        free(parent);
        free(cmin);
        return -1;
    }
    int64_t c = -1;
    for (int64_t s = 0; s < n; ++s) {
        if (s == 0 || ki[s].key != ki[s - 1].key) {
            ++c;
            start[c] = s;
            ckey[c] = ki[s].key;
            parent[c] = c;
            cmin[c] = ki[s].idx;
        } else if (ki[s].idx < cmin[c]) {
            cmin[c] = ki[s].idx;
        }
    }
    start[ncell] = n;
    for (int64_t a = 0; a < ncell; ++a) {
        const int64_t k = ckey[a];
        const int64_t a2 = k % nc[2], a1 = (k / nc[2]) % nc[1], a0 = k / (nc[1] * nc[2]);
        for (int64_t d0 = -2; d0 <= 2; ++d0) {
            const int64_t b0 = a0 + d0;
            if (b0 < 0 || b0 >= nc[0]) continue;
            for (int64_t d1 = -2; d1 <= 2; ++d1) {
                const int64_t b1 = a1 + d1;
                if (b1 < 0 || b1 >= nc[1]) continue;
                for (int64_t d2 = -2; d2 <= 2; ++d2) {
                    const int64_t b2 = a2 + d2;
                    if (b2 < 0 || b2 >= nc[2]) continue;
                    const int64_t kb = (b0 * nc[1] + b1) * nc[2] + b2;
                    if (kb <= k) continue; /* each unordered pair of cells once */
                    int64_t l = a + 1, h = ncell;
                    while (l < h) {
                        const int64_t mid = (l + h) >> 1;
                        if (ckey[mid] < kb) l = mid + 1;
                        else h = mid;
                    }
                    if (l >= ncell || ckey[l] != kb) continue;
                    if (uf_find(parent, a) == uf_find(parent, l)) continue;
                    int linked = 0;
                    for (int64_t s = start[a]; s < start[a + 1] && !linked; ++s) {
                        const int64_t i = ki[s].idx;
                        for (int64_t t = start[l]; t < start[l + 1]; ++t) {
                            const int64_t j = ki[t].idx;
                            const double dx = x[i] - x[j], dy = y[i] - y[j], dz = z[i] - z[j];
                            if ((dx * dx + dy * dy) + dz * dz <= ll2) {
                                linked = 1;
                                break;
                            }
                        }
                    }
                    if (linked) uf_union(parent, a, l);
                }
            }
        }
    }
    /* component minimum particle index at each root, then the labels */
    for (int64_t a = 0; a < ncell; ++a) {
        const int64_t r = uf_find(parent, a);
        if (cmin[a] < cmin[r]) cmin[r] = cmin[a];
    }
    for (int64_t a = 0; a < ncell; ++a) {
        const int64_t m = cmin[uf_find(parent, a)];
        for (int64_t s = start[a]; s < start[a + 1]; ++s) labels[ki[s].idx] = m;
    }
    free(ki);
    free(start);
    free(ckey);
    free(parent);
    free(cmin);
    return 0;
}
