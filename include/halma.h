/*
 * libhalma -- MI355X (gfx950) halo-finding hot path: friends-of-friends grouping,
 * per-group reductions and O(N^2) pairwise-potential unbinding.
 *
 * C ABI.  Every entry point:
 *   - takes device pointers for particle data (caller-allocated; nothing is mutated
 *     except the declared outputs and the workspace), plain sizes, and a HIP stream
 *     passed as `void *` (NULL = the default stream);
 *   - is stream-ordered unless stated otherwise (a few entry points synchronise the
 *     stream because they return sizes or iteration counts to the host; they say so);
 *   - uses only the caller's workspace for scratch device memory (query the size with
 *     the matching *_workspace_size function) -- the library calls no hipMalloc;
 *   - returns 0 on success, a positive hipError_t value on a HIP failure, or a negative
 *     HALMA_E_* code on an argument error; halma_last_error() then describes it.
 *
 * Particle coordinate arrays are structure-of-arrays.  Where an entry point accepts
 * either precision, a dtype flag says which (HALMA_F32 or HALMA_F64).
 */
#ifndef HALMA_H
#define HALMA_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HALMA_F32 0
#define HALMA_F64 1

#define HALMA_OK 0
#define HALMA_E_ARG (-1)        /* invalid argument (null pointer, negative size, ...) */
#define HALMA_E_WORKSPACE (-2)  /* workspace smaller than the *_workspace_size query */
#define HALMA_E_LIMIT (-3)      /* a size beyond a documented limit */
#define HALMA_E_STATE (-4)      /* internal consistency check failed */

/* Message describing the last non-zero return on this host thread ("" if none). */
const char *halma_last_error(void);

/* Library ABI version (major*10000 + minor*100 + patch). */
int halma_version(void);

/* Compute units of the current device (0 if no device). */
int halma_device_cu_count(void);

/* ------------------------------------------------------------------------------------
 * Pair potential.  Replaces particle.particle.brute_force_binding_energy
 * (fortran_modules/particle_subroutines.f90:466-514, f2py binding called at
 * python_scripts/halo_gas.py:182) and its serial twin (:517-556).
 *
 * out[i] = sum_j m[j] / |x_j - x_i| over the sources j whose three coordinates all
 * differ from test i's (particle_subroutines.f90:499-501), inputs FP32 (the caller
 * casts, as halo_gas.py:172-178 does).  Summed in FP32 on the vector ALUs with
 * Kahan-compensated tiles; the FP64 value of the compensated sum is written to
 * out_f64 when that pointer is non-NULL, and out (float32, the reference's output
 * type) may be NULL when only out_f64 is wanted.  The relative error against the
 * FP64 direct sum is below halma_pot_error_bound() (a rigorous bound).
 * ---------------------------------------------------------------------------------- */
size_t halma_pot_sum_workspace_size(int64_t n_src, int64_t n_test);
int halma_pot_sum_f32(const float *m, const float *x, const float *y, const float *z,
                      int64_t n_src, const float *tx, const float *ty, const float *tz,
                      int64_t n_test, float *out, double *out_f64, void *workspace,
                      size_t workspace_bytes, void *stream);

/* Relative error bound B (|S_f32 - S_exact| <= B * S_f32) of halma_pot_sum_f32 and the
 * batched kernels for the current tuning. */
double halma_pot_error_bound(void);

/* Batched form: H independent haloes in CSR layout.  Halo h's sources are
 * [src_off[h], src_off[h+1]) of (sm, sx, sy, sz) and its tests are
 * [tst_off[h], tst_off[h+1]) of (tx, ty, tz); out_f64 is indexed like the tests.
 * The offsets are HOST arrays of H+1 entries (the work queue is planned on the host,
 * sorted by pair count, largest first, and executed by one persistent launch).
 * Replaces the per-halo loop of calls at pyHALMA.py:930-1067 / halo_gas.py:285-634. */
size_t halma_pot_sum_batched_workspace_size(int64_t n_halo, const int64_t *src_off,
                                            const int64_t *tst_off);
int halma_pot_sum_batched_f32(int64_t n_halo, const int64_t *src_off, const float *sm,
                              const float *sx, const float *sy, const float *sz,
                              const int64_t *tst_off, const float *tx, const float *ty,
                              const float *tz, double *out_f64, void *workspace,
                              size_t workspace_bytes, void *stream);

/* The same sum in FP64 with the oracle's fixed summation order (64 strided partials
 * of ascending j, then a pairwise tree), one wavefront per test particle.  Bitwise
 * reproducible; used for the band re-test and exposed for verification. */
int halma_pot_sum_f64_ordered(const float *m, const float *x, const float *y,
                              const float *z, int64_t n_src, const float *tx,
                              const float *ty, const float *tz, int64_t n_test,
                              double *out, void *stream);

/* Plain FP64 comparator: one thread per test particle, sequential FP64 sum with FP64
 * sqrt and divide, FP64 inputs, one launch per call (no batching, no graphs). */
int halma_pot_sum_f64_plain(const double *m, const double *x, const double *y,
                            const double *z, int64_t n_src, const double *tx,
                            const double *ty, const double *tz, int64_t n_test,
                            double *out, void *stream);

/* ------------------------------------------------------------------------------------
 * Friends-of-friends.  Replaces fof.friends_of_friends (python_scripts/fof.py:18-79, the
 * scipy-KDTree breadth-first search) and, for the same groups, the third-party pyfof
 * call of fof.py:184-187.
 *
 * labels[i] = smallest particle index of i's connected component in the graph with an
 * edge wherever ((dx*dx + dy*dy) + dz*dz) <= ll*ll in FP64 (FP32 coordinates are
 * widened exactly), non-periodic.  Singletons label themselves.  Coordinates may span
 * at most 2^21 sub-cells of side ll/sqrt(3) per axis (HALMA_E_LIMIT otherwise).
 * Synchronises the stream (it checks the kernels' error flag before returning).
 * ---------------------------------------------------------------------------------- */
size_t halma_fof_workspace_size(int64_t n, int dtype);
int halma_fof_labels(const void *x, const void *y, const void *z, int dtype, int64_t n,
                     double ll, int32_t *labels, void *workspace, size_t workspace_bytes,
                     void *stream);

/* Tuning knobs of the FP32 pair kernels (set by the autotuner; read back with get).
 *   HALMA_TUNE_TESTS_PER_LANE : test particles per lane (1, 2, 4)
 *   HALMA_TUNE_SUBTILE        : sources per plain FP32 sub-tile before the Kahan add
 *                               (4 or 8)
 *   HALMA_TUNE_MIN_CHUNK      : smallest source chunk when a halo's source range is
 *                               split across waves (multiple of 64, >= 64)
 *   HALMA_TUNE_ITEMS_PER_WAVE : target queue items per resident wave
 *   HALMA_TUNE_WAVES_PER_CU   : persistent waves launched per compute unit (4..32) */
#define HALMA_TUNE_TESTS_PER_LANE 1
#define HALMA_TUNE_SUBTILE 2
#define HALMA_TUNE_MIN_CHUNK 3
#define HALMA_TUNE_ITEMS_PER_WAVE 4
#define HALMA_TUNE_WAVES_PER_CU 5
int halma_tune_set(int key, int value);
int halma_tune_get(int key);

#ifdef __cplusplus
}
#endif
#endif /* HALMA_H */
