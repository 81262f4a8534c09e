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
 * FP64 direct sum is below halma_pot_error_bound() (a rigorous bound).  The rule is
 * evaluated as "the smallest |coordinate difference| is not zero", which is the
 * reference's for finite and NaN coordinates; two equal infinite coordinates give
 * inf - inf = NaN, so their pair is counted (the reference drops it): infinite
 * coordinates are outside parity.
 * ---------------------------------------------------------------------------------- */
size_t halma_pot_sum_workspace_size(int64_t n_src, int64_t n_test);
int halma_pot_sum_f32(const float *m, const float *x, const float *y, const float *z,
                      int64_t n_src, const float *tx, const float *ty, const float *tz,
                      int64_t n_test, float *out, double *out_f64, void *workspace,
                      size_t workspace_bytes, void *stream);

/* halma_pot_sum_f32 without a workspace argument and without any allocation.  No
 * counterpart in the reference: for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * scratch memory is a static device array of halma_pot_sum_noalloc_max_bytes() bytes
 * (64 MiB) per GPU, held until the process ends; sizes whose plan needs more return
 * HALMA_E_LIMIT before any GPU work.  Same plan and kernels as halma_pot_sum_f32, so the
 * same bits for the same inputs and tuning.  Every call on one GPU shares that GPU's
 * array: calls that overlap on one GPU (other streams, other host threads) race, so
 * callers must order them.  The array keeps the last call's contents; nothing in it is
 * assumed to start at zero (the queue counter is cleared by a kernel each call).
 * Stream-ordered. */
size_t halma_pot_sum_noalloc_max_bytes(void);
int halma_pot_sum_f32_noalloc(const float *m, const float *x, const float *y, const float *z,
                              int64_t n_src, const float *tx, const float *ty, const float *tz,
                              int64_t n_test, float *out, double *out_f64, void *stream);

/* Relative error bound B (|S_f32 - S_exact| <= B * S_f32) of halma_pot_sum_f32 and the
 * batched kernels for the current tuning. */
double halma_pot_error_bound(void);

/* Self-potential of n_halo haloes with tests = sources (the unbinding's full passes):
 * out[k] = sum over the other members j of k's halo of m_j / r_kj, with the same
 * three-coordinate rule (which also drops j = k).  Members are the CSR ranges
 * [offsets[h], offsets[h+1]) (HOST offsets) of the FP32 arrays m, x, y, z; out is FP64
 * and indexed like them.  Every pair distance is computed once for both of its terms;
 * the relative error is below halma_pot_sym_error_bound().  Haloes of more than 64
 * blocks of 512 members run their block pairs in tiles of 64 x 64 blocks, one launch per
 * tile, each tile's partial sums added (FP64, fixed order) before the next, so the
 * workspace stays linear in the member count (at most 65 partial slots per member).
 * Stream-ordered. */
size_t halma_pot_sum_self_workspace_size(int64_t n_halo, const int64_t *offsets);
int halma_pot_sum_self_batched(int64_t n_halo, const int64_t *offsets, const float *m,
                               const float *x, const float *y, const float *z, double *out,
                               void *workspace, size_t workspace_bytes, void *stream);
double halma_pot_sym_error_bound(void);

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

/* The same sum in FP64 with the oracle's fixed summation order (256 strided partials
 * of ascending j, then a pairwise tree), one 256-thread block per test particle.
 * Bitwise reproducible; used for the band re-test and exposed for verification. */
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
 * at most 2^21 sub-cells of side ll/sqrt(3) per axis and n <= 2^30 (HALMA_E_LIMIT
 * otherwise).
 * Synchronises the stream (it checks the kernels' error flag before returning).
 * ---------------------------------------------------------------------------------- */
size_t halma_fof_workspace_size(int64_t n, int dtype);
int halma_fof_labels(const void *x, const void *y, const void *z, int dtype, int64_t n,
                     double ll, int32_t *labels, void *workspace, size_t workspace_bytes,
                     void *stream);

/* ------------------------------------------------------------------------------------
 * Groups.  From FoF labels to the groups of at least min_size members (the size filter
 * of good_groups, pyHALMA.py:148-149,592), in CSR form: groups ordered by label (their
 * minimum member index, the order fof.py:34-77 emits them), members ascending (fof.py:76).
 * members needs n entries, offsets n+1.  Synchronises the stream (returns the counts).
 * ---------------------------------------------------------------------------------- */
size_t halma_groups_workspace_size(int64_t n);
int halma_groups_from_labels(const int32_t *labels, int64_t n, int32_t min_size,
                             int32_t *members, int64_t *offsets, int64_t *n_groups,
                             int64_t *n_members, void *workspace, size_t workspace_bytes,
                             void *stream);

/* Per-group reductions (halo_properties.py:16-24 total_mass, :26-43 center_of_mass,
 * :45-60 CM_velocity, :98-111 furthest_particle), FP64 with double-double accumulation.
 * out has HALMA_GROUP_FIELDS doubles per group:
 *   [0] M  [1..3] centre of mass (0 if M <= 0)  [4..6] mass-weighted velocity
 *   (0 if M <= 0 or no velocities)  [7] distance of the furthest member from the centre
 *   (the given centres when non-NULL, 3 per group, else the centre of mass)
 *   [8..10] min x,y,z  [11..13] max x,y,z  [14] sum |m| (upper bound)
 *   [15] 1 if every partial sum of the group's masses is exact in FP64, else 0.
 * Velocity pointers may be NULL.  Stream-ordered. */
#define HALMA_GROUP_FIELDS 16
int halma_group_reduce(int64_t n_groups, const int64_t *offsets, const int32_t *members,
                       const void *x, const void *y, const void *z, int pos_dtype,
                       const void *vx, const void *vy, const void *vz, int vel_dtype,
                       const void *m, int mass_dtype, const double *centres, double *out,
                       void *stream);

/* Member-ordered copies of the particle arrays: out_*[k] = in_*[members[k]] for the
 * n_members members (x, y, z, m, and vx, vy, vz when given), in out_dtype.  Particles
 * are first packed into 32/64-byte records so the gather costs one memory line per
 * member instead of one per value.  Reductions, density peaks and unbinding can then run
 * on the copies with members = 0..n_members-1.  The workspace holds n_particles records.
 * Stream-ordered. */
size_t halma_gather_workspace_size(int64_t n_particles, int dtype);
int halma_gather_members(const int32_t *members, int64_t n_members, int64_t n_particles,
                         const void *x, const void *y, const void *z, int pos_dtype,
                         const void *vx, const void *vy, const void *vz, int vel_dtype,
                         const void *m, int mass_dtype, int out_dtype, void *out_x,
                         void *out_y, void *out_z, void *out_m, void *out_vx, void *out_vy,
                         void *out_vz, void *workspace, size_t workspace_bytes,
                         void *stream);

/* density_peak (halo_properties.py:518-556) for every group: node coordinates (3 doubles
 * per group) of the heaviest cell of the 50^3 nearest-node grid spanning
 * [min - ll, max + ll] per axis.  `reduced` is halma_group_reduce's output for the same
 * groups.  Synchronises the stream. */
size_t halma_density_peak_workspace_size(int64_t n_groups);
int halma_density_peak(int64_t n_groups, const int64_t *offsets, const int32_t *members,
                       const void *x, const void *y, const void *z, int pos_dtype,
                       const void *m, int mass_dtype, double ll, const double *reduced,
                       double *out, void *workspace, size_t workspace_bytes, void *stream);

/* half_mass_radius (halo_properties.py:366-392) for every group: the distance to the
 * centre of the last member, members by increasing distance, before the running mass
 * first exceeds M / 2 (the smallest distance if the first member already exceeds it,
 * the largest if none does; 0 for an empty group).  Groups are the CSR ranges
 * offsets[g]..offsets[g+1] (device, offsets[0] = 0, n_members = offsets[G]) of
 * `members`; `reduced` is halma_group_reduce's output for the same groups (M and the
 * centre).  Equal distances keep member order (the reference's np.argsort leaves them
 * unordered).  Used by the substructure splitting of pyHALMA.py:703-766. */
size_t halma_half_mass_radius_workspace_size(int64_t n_members);
int halma_half_mass_radius(int64_t n_groups, int64_t n_members, const int64_t *offsets,
                           const int32_t *members, const void *x, const void *y,
                           const void *z, int pos_dtype, const void *m, int mass_dtype,
                           const double *reduced, double *out, void *workspace,
                           size_t workspace_bytes, void *stream);

/* Phase-space cleaning (halo_properties.py:113-278: calc_cell / clean_cell, driver
 * pyHALMA.py:789-843), one round for many groups.  The set S of the round is given as
 * n_set members (particle indices `set_members`, group `set_groups`, groups' members
 * contiguous and in ascending position).  halma_clean_cells puts every member of S in the
 * nearest node of its group's grid per axis (node k of group g: grid_start[g] for k = 0,
 * grid_g1[g] for k = 1, grid_start[g] + k * grid_d[g] after, grid_n[g] nodes: the values of
 * np.arange(-(R + ll), R + ll, 2 ll)) about centre[3 g..] (first node among equal
 * distances; max_grid_n (host) >= every grid_n of the set, at most 2^21), writes the three
 * node indices to ijk (3 per member of S, device, S order), and the per-cell
 * statistics to cell_stats (6 doubles per cell, cells in (group, cell) order: mass-
 * weighted velocity (3), mass, member count, dispersion sqrt(sum |v_cell - v|^2 / 3 / q));
 * *n_cells (HOST) their number.  Sums are double-double (the reference's member-order FP64
 * sums agree to a few ulp); v * m is formed in FP32 when both arrays are FP32, as numba
 * does.  Synchronises the stream.  halma_clean_decide, called next with the same set and
 * workspace, tests every member of S whose group has phase 0 (first round: |v_cell - v| /
 * sigma counts only in cells of more than two members) or 1 (later rounds: two or more):
 * a member of a cell with at most q_fil members, or with the scaled deviation >= fac[g] *
 * sig_fil, is removed: control[set_target[k]] = 0 and cleaned[g] += 1. */
size_t halma_clean_workspace_size(int64_t n_set);
/* The set S of a round (clean_cell's clean_particles = part_list[control != 0] of every
 * live group, halo_properties.py:209 and :263): over the CSR groups offsets (device, G+1,
 * offsets[0] = 0, n_members = offsets[G]), member position k of group g is in S when
 * live[g] and (control[k] or force[g]) (device uint8 arrays; force may be NULL).  S in
 * position order: set_members = members[k] (the position k itself when members is NULL),
 * set_groups = g, set_target = offsets[g] + the member's rank in its group's part of S
 * (the control entry that clean_cell clears for it); set_offsets (device, G+1) the CSR
 * offsets of S per group; *n_set (HOST) its size.  Synchronises the stream. */
size_t halma_clean_select_workspace_size(int64_t n_members);
int halma_clean_select(int64_t n_groups, int64_t n_members, const int64_t *offsets,
                       const int32_t *members, const uint8_t *control, const uint8_t *live,
                       const uint8_t *force, int32_t *set_members, int32_t *set_groups,
                       int64_t *set_target, int64_t *set_offsets, int64_t *n_set,
                       void *workspace, size_t workspace_bytes, void *stream);
int halma_clean_cells(int64_t n_set, int64_t n_groups, const int32_t *set_members,
                      const int32_t *set_groups, const void *x, const void *y, const void *z,
                      int pos_dtype, const void *vx, const void *vy, const void *vz,
                      int vel_dtype, const void *m, int mass_dtype, const double *centre,
                      const double *grid_start, const double *grid_g1, const double *grid_d,
                      const int32_t *grid_n, int32_t max_grid_n, int32_t *ijk,
                      double *cell_stats, int64_t *n_cells, void *workspace,
                      size_t workspace_bytes, void *stream);
int halma_clean_decide(int64_t n_set, const int32_t *set_members, const int32_t *set_groups,
                       const int64_t *set_target, const void *vx, const void *vy,
                       const void *vz, int vel_dtype, const double *cell_stats,
                       const int32_t *phase, const double *fac, double q_fil, double sig_fil,
                       uint8_t *control, int32_t *cleaned, void *workspace,
                       size_t workspace_bytes, void *stream);

/* Particles within a sphere, many spheres at once: the DM and star KD-tree queries of
 * st_gas_dm_particles_inside (halo_gas.py:255, :269, query_ball_point: distance <= R,
 * strict = 0) or parallel_inside (halo_gas.py:218-220, sqrt(d2) < R, strict = 1), and
 * the dark-matter mass inside a sphere (pyHALMA.py:1172-1185).  Points x, y, z (device,
 * pos_dtype) are binned in cells of side `cell` (<= 0: chosen from the bounding box;
 * coarsened when the grid would exceed 4 cells per point); centres (device, 3 per sphere,
 * FP64), radii (device FP64).  Outputs: offsets (device, n_centres+1) and indices (device;
 * NULL to count only, max_found is then not checked): the indices of the points of sphere h at offsets[h]..,
 * ascending when `sorted`; mass_sums (device FP64 per sphere, may be NULL): the
 * double-double sum of m over each sphere's points; *n_found (HOST) the total.  Returns
 * HALMA_E_LIMIT (with *n_found set) when the total exceeds max_found, so a caller can
 * size the workspace again.  FP64 predicate (dx^2 + dy^2) + dz^2 against R^2.
 * Synchronises the stream. */
size_t halma_ball_query_workspace_size(int64_t n_points, int64_t n_centres, int64_t max_found);
int halma_ball_query(int64_t n_points, const void *x, const void *y, const void *z,
                     int pos_dtype, const void *m, int mass_dtype, int64_t n_centres,
                     const double *centres, const double *radii, int strict, double cell,
                     int sorted, int64_t max_found, int64_t *offsets, int32_t *indices,
                     double *mass_sums, int64_t *n_found, void *workspace,
                     size_t workspace_bytes, void *stream);

/* Gas particles from AMR cells for many queries (halo_gas.py:9-52 patch_to_particles,
 * :56-141 AMRgrid_to_particles, and the sphere test and mass scaling of :236-246).
 * Patches (device arrays, n_patches each): level, nx, ny, nz, rx, ry, rz (the centre of the
 * patch's first parent cell), cell_offsets (n_patches+1: patch p's cells at
 * cell_offsets[p] + (ix ny + iy) nz + iz); per cell (device): delta, vx, vy, vz, temp
 * (gas_dtype) and the cr0amr / solapst flags (uint8).  res / res3 (HOST, n_levels): the
 * cell side (L/ncoarse)/2^l of each level and its cube as the caller computes it.  For
 * query h (boxes: device FP64, 6 each: x lo, x hi, y lo, y hi, z lo, z hi) every cell of a
 * level >= 1 patch whose centre x0 + ix res (x0 = rx - res/2) is strictly inside the box
 * per axis, with cr0amr and solapst set -- and, when centres (device FP64, 3 each) is not
 * NULL, sqrt((dx^2 + dy^2) + dz^2) < radii[h] -- gives a particle: x, y, z,
 * v * vel_scale (3), (1 + delta) rho_B res3 mass_scale, temp, written to out (device FP64,
 * 8 arrays of max_found: quantity q of particle k at out[q max_found + k]) in the
 * reference's order (patches ascending, then ix, iy, iz) at offsets[h] (device,
 * n_queries+1).  *n_found (HOST) the total; HALMA_E_LIMIT when it exceeds max_found (out
 * NULL: count only).  Synchronises the stream. */
size_t halma_amr_particles_workspace_size(int64_t n_queries);
int halma_amr_particles(int64_t n_patches, const int32_t *level, const int32_t *nx,
                        const int32_t *ny, const int32_t *nz, const double *rx,
                        const double *ry, const double *rz, const int64_t *cell_offsets,
                        const void *delta, const void *vx, const void *vy, const void *vz,
                        const void *temp, int gas_dtype, const uint8_t *cr0amr,
                        const uint8_t *solapst, int n_levels, const double *res,
                        const double *res3, double rho_B, double mass_scale,
                        double vel_scale, int64_t n_queries, const double *boxes,
                        const double *centres, const double *radii, int64_t max_found,
                        int64_t *offsets, double *out, int64_t *n_found, void *workspace,
                        size_t workspace_bytes, void *stream);

/* halo_shape (particle_subroutines.f90:160-214) of many groups: offsets (device, G+1)
 * over member-ordered FP32 x, y, z (relative to each group's centre, as
 * halo_properties.py:852-858 passes them) and masses; out (device, 3 FP32 per group): the
 * square roots of the eigenvalues of sum(m r r^T) / sum(m), largest first.  The FP32
 * terms are summed in double-double (the reference: FP32 OpenMP reduction) and the
 * eigenvalues come from the file's Jacobi variant in FP64. */
int halma_halo_shape(int64_t n_groups, const int64_t *offsets, const float *x, const float *y,
                     const float *z, const float *mass, float *out, void *stream);

/* sigma_projections (particle_subroutines.f90:217-461) of one halo: FP32 member arrays
 * (device, n; velocities relative to the halo's, as halo_properties.py:791-793 passes
 * them), grid (device, n_cell FP32 nodes), centre, R05 per axis and ll; out (device, 5
 * FP32): SIG_1D_x_05, SIG_1D_y_05, SIG_1D_z_05, V_sigma, lambda.  Per-particle sums are
 * FP64 rounded to FP32 (the reference: FP32 OpenMP reductions); the sums over cells are
 * sequential FP32 in the reference's loop order. */
size_t halma_sigma_projections_workspace_size(int64_t n, int n_cell);
int halma_sigma_projections(int64_t n, const float *x, const float *y, const float *z,
                            const float *vx, const float *vy, const float *vz,
                            const float *mass, const float *grid, int n_cell, float cx,
                            float cy, float cz, float R05x, float R05y, float R05z, float ll,
                            float *out, void *workspace, size_t workspace_bytes,
                            void *stream);

/* Progenitor matching (the merger section of pyHALMA.py:1211-1266, main_progenitors_finder
 * :1219-1257) for every new halo at once.  New haloes: CSR new_offsets (device, n_new+1,
 * new_offsets[0] = 0) over new_ids (device int64, the particle IDs of each halo's
 * members, st_oripa[halo]) and their masses new_mass (device FP64, n_new); old haloes
 * likewise (the previous iteration's oripas_before and omm).  An old halo intersects a
 * new one when they share an ID (np.in1d(oripas, ooripas) of :1230-1234); an ID may be
 * listed in several old haloes.  Outputs (device int32, n_new each): pro1 / pro2 (1-based
 * old index of the most and second most massive intersecting old halo, 0 for none; equal
 * masses rank the higher old index first), n_mergers (intersecting old haloes),
 * merger_type (-1 when 1.2 * new mass < pro1's mass; with two or more intersections 2
 * when 1/20 < second / first mass < 1/3, 3 otherwise -- the reference's type 1 is
 * overwritten by its following if / else; 0 without intersections or with one).  At most
 * 2^30-1 members per catalogue; IDs may not span the whole int64 range.  Synchronises
 * the stream. */
size_t halma_progenitors_workspace_size(int64_t n_new, int64_t n_new_members, int64_t n_old,
                                        int64_t n_old_members);
int halma_progenitors(int64_t n_new, int64_t n_new_members, const int64_t *new_offsets,
                      const int64_t *new_ids, const double *new_mass, int64_t n_old,
                      int64_t n_old_members, const int64_t *old_offsets,
                      const int64_t *old_ids, const double *old_mass, int32_t *pro1,
                      int32_t *pro2, int32_t *n_mergers, int32_t *merger_type,
                      void *workspace, size_t workspace_bytes, void *stream);

/* ------------------------------------------------------------------------------------
 * Batched iterative unbinding.  The energy test of halo_properties.py:342-359 (stars,
 * factor = FACTOR_V^2) and halo_gas.py:456-476 (gas, factor = 2):
 *   E_i = 1/2 |v_i - v_cm|^2 - factor * G * S_i,   G = 4.3e-9 (km/s)^2 Mpc / Msun,
 * S_i = sum over the halo's bound members j of m_j / r_ij (FP32-cast positions and
 * masses, particle_subroutines.f90:466-514), bound iff E_i <= 0; v_cm is the mass-weighted
 * velocity of the bound members.  Every halo is iterated until a pass removes nobody (at
 * most max_iter passes); max_iter = 1 is the reference's single test with v_cm over all
 * members.  Haloes are the CSR ranges offsets[h]..offsets[h+1] of `members` (particle
 * indices into the global arrays); offsets is a HOST array of n_halo+1 entries.
 * Decisions are exact with respect to the FP64 oracle: members whose FP32-compensated E
 * lies inside the rigorous error band are re-tested with an FP64 sum in the oracle's
 * fixed order.  Outputs (device): bound[k] per member position (1 bound, 0 not); iters
 * (n_halo int32, may be NULL) passes per halo; vcm (3 per halo, may be NULL); pot (per
 * member, the last S used, may be NULL).  stats (HOST, 13 int64, may be NULL): [0] passes,
 * [1] members re-tested in FP64, [2] symmetric-kernel queue items of the first pass,
 * [3] tests per lane, [4] sub-tile, [5] pair-kernel time in ns and [6] timed passes
 * (with HALMA_UNBIND_TIME_POT: hipEvents recorded on the stream around the pair
 * kernels), [7] pairs of the reference's loop (n_bound x n_sources summed over haloes
 * and passes), [8] symmetric-kernel time in ns (part of [5]), [9] its launches,
 * [10] pairs of the symmetric kernel's passes (part of [7]; it evaluates each unordered
 * pair once for two of them), [11] chunked-kernel launches, [12] members of [1] re-summed
 * because their S had an error bound above 1e-6 of S (a kept member whose incremental
 * updates took its rigorous bound past 1e-6 gets a fresh FP64 sum, so every S the call
 * keeps is within 1e-6 of the FP64 direct sum).  Full passes (first pass, or
 * more than a quarter of the live set removed) go to the symmetric kernel for haloes of
 * at most HALMA_TUNE_SYM_MAX members, to the chunked kernel otherwise; incremental passes
 * go to the chunked kernel.  Passes are queued one ahead of the host: the counters of
 * pass p are read back (pinned buffer, event) while pass p + 1 is already queued, so
 * the call ends with one extra pass that finds no active halo.  flags:
 * HALMA_UNBIND_GRAPHS captures each part of a pass (prepare, symmetric kernel, chunked
 * kernel, decisions) into a hipGraph replayed every pass; by default the parts are plain
 * launches (a few passes do not repay the capture).
 * ---------------------------------------------------------------------------------- */
#define HALMA_UNBIND_GRAPHS 1
#define HALMA_UNBIND_TIME_POT 2
size_t halma_unbind_workspace_size(int64_t n_halo, const int64_t *offsets);
int halma_unbind_batched(int64_t n_halo, const int64_t *offsets, const int32_t *members,
                         const void *x, const void *y, const void *z, int pos_dtype,
                         const void *vx, const void *vy, const void *vz, int vel_dtype,
                         const void *m, int mass_dtype, double factor, int max_iter, int flags,
                         uint8_t *bound, int32_t *iters, double *vcm, double *pot,
                         int64_t *stats, void *workspace, size_t workspace_bytes, void *stream);

/* The plain FP64 comparator of the same method: one halo at a time, one launch per step,
 * FP64 storage, sequential per-thread sums with FP64 sqrt and divide, no batching and no
 * graphs.  iters_host: HOST int32 per halo (may be NULL).  Synchronises often. */
size_t halma_unbind_plain_workspace_size(int64_t n_halo, const int64_t *offsets);
int halma_unbind_plain_f64(int64_t n_halo, const int64_t *offsets, const int32_t *members,
                           const void *x, const void *y, const void *z, int pos_dtype,
                           const void *vx, const void *vy, const void *vz, int vel_dtype,
                           const void *m, int mass_dtype, double factor, int max_iter,
                           uint8_t *bound, int32_t *iters_host, void *workspace,
                           size_t workspace_bytes, void *stream);

/* ------------------------------------------------------------------------------------
 * Batched single energy test with exact decisions.  Replaces the per-halo energy test
 * of halo_gas.RPS (halo_gas.py:456-476: sources = gas + dark matter + stars, tests =
 * gas cells, factor 2, v_h = the halo's velocity) and of
 * escape_velocity_unbinding_fortran (halo_properties.py:333-359, factor FACTOR_V^2),
 * for every halo in one call, without the reference's random subsampling of large
 * source groups (every source is summed).
 *   E_i = 1/2 |v_i - v_h|^2 - factor * G * S_i,  G = 4.3e-9,  bound[i] = (E_i <= 0)
 * with S_i = sum_j m_j / r_ij over halo h's sources [src_off[h], src_off[h+1]) whose
 * three coordinates all differ from test i's; tests [tst_off[h], tst_off[h+1]) (HOST
 * offsets).  halo_v: 3 doubles per halo (device).  bound and s_out (may be NULL) are
 * indexed like the tests.  Decisions equal those made with the FP64 sum in the oracle's
 * fixed order (band re-test, see halma_unbind_batched).  stats (HOST, 3 int64, may be
 * NULL; the call then synchronises the stream): [0] tests re-tested in FP64, [1] time of
 * the pair kernel in ns (hipEvents on the stream around its launch), [2] (test, source)
 * pairs summed.
 * ---------------------------------------------------------------------------------- */
size_t halma_energy_test_workspace_size(int64_t n_halo, const int64_t *src_off,
                                        const int64_t *tst_off);
int halma_energy_test_batched(int64_t n_halo, const int64_t *src_off, const float *sm,
                              const float *sx, const float *sy, const float *sz,
                              const int64_t *tst_off, const float *tx, const float *ty,
                              const float *tz, const void *vx, const void *vy, const void *vz,
                              int vel_dtype, const double *halo_v, double factor,
                              uint8_t *bound, double *s_out, int64_t *stats, void *workspace,
                              size_t workspace_bytes, void *stream);

/* The same test when each halo's tests are also its first sources (the gas cells of
 * halo_gas.py:285-492, which are their own sources, followed by dark matter and stars):
 * halo h's tests are [tst_off[h], tst_off[h+1]) of (tm, tx, ty, tz) and its sources are
 * those tests then its extra sources [ext_off[h], ext_off[h+1]) of (em, ex, ey, ez) (HOST
 * offsets; a halo may have no extra sources).  The tests' own pair sum runs on the
 * symmetric kernel (each pair distance once, halma_pot_sum_self_batched), the extra
 * sources on the chunked kernel; decisions, band re-test (FP64 in the oracle's order over
 * the tests then the extra sources) and outputs as halma_energy_test_batched with
 * src = tests ++ extras.  stats [2] counts (test, source) pairs of the reference's loop. */
size_t halma_energy_test_self_workspace_size(int64_t n_halo, const int64_t *tst_off,
                                             const int64_t *ext_off);
int halma_energy_test_self_batched(int64_t n_halo, const int64_t *tst_off, const float *tm,
                                   const float *tx, const float *ty, const float *tz,
                                   const int64_t *ext_off, const float *em, const float *ex,
                                   const float *ey, const float *ez, const void *vx,
                                   const void *vy, const void *vz, int vel_dtype,
                                   const double *halo_v, double factor, uint8_t *bound,
                                   double *s_out, int64_t *stats, void *workspace,
                                   size_t workspace_bytes, void *stream);

/* ------------------------------------------------------------------------------------
 * Softened potential and its gradient, for the potential-minimum centring of haloes
 * (the differentiable extension of the unbinding sum: no reference function computes
 * it; the reference's most-bound position, halo_gas.py:498-634, is its discrete
 * counterpart).  For each halo h (CSR offsets, HOST arrays of n_halo+1 entries) and each
 * test i of h (FP64 positions tx, ty, tz):
 *   pot[i]    = sum_j m_j (r_ij^2 + eps^2)^(-1/2)
 *   acc[3i+k] = sum_j m_j (x_j - t_i)_k (r_ij^2 + eps^2)^(-3/2)     (= d pot[i] / d t_i)
 *   wsum[i]   = sum_j m_j (r_ij^2 + eps^2)^(-3/2)
 * over the sources j of h (FP32 sm, sx, sy, sz) whose three coordinates all differ
 * from the test's (particle_subroutines.f90:499-501; drops the self pair).  Any of
 * pot, acc, wsum may be NULL; when pot and acc or wsum are asked for together, one pass
 * over the pairs computes them all.  FP32 compensated sums merged in FP64: pot within
 * halma_pot_error_bound() relative; acc and wsum within the same bound relative to
 * sum_j |term_j|.  A halo with at most 8 tests spreads each test's sources over the lanes
 * of a wavefront (chunks of 4096 sources); larger haloes hold 64 tests per wavefront.
 * Chunk sizes follow from each halo's own source and test counts and the chunk
 * partials are added in chunk order, so a test's bits depend on its halo's inputs only
 * (not on the other haloes of the call).  The pair terms are recomputed, never stored.
 * Stream-ordered.
 * ---------------------------------------------------------------------------------- */
size_t halma_soft_workspace_size(int64_t n_halo, const int64_t *src_off,
                                 const int64_t *tst_off);
int halma_pot_soft_batched(int64_t n_halo, const int64_t *src_off, const float *sm,
                           const float *sx, const float *sy, const float *sz,
                           const int64_t *tst_off, const double *tx, const double *ty,
                           const double *tz, double eps, double *pot, double *acc,
                           double *wsum, void *workspace, size_t workspace_bytes,
                           void *stream);
/* The same with the sources packed by the caller: src_xyzm holds (x, y, z, m) per source
 * as 4 consecutive floats (16-byte aligned), indexed like sm, sx, sy, sz above.  For
 * callers that evaluate the same sources many times (the centring iterations): the
 * sources are not repacked on every call.  Same workspace size, results bit-identical
 * to halma_pot_soft_batched on the same values. */
int halma_pot_soft_batched_packed(int64_t n_halo, const int64_t *src_off,
                                  const float *src_xyzm, const int64_t *tst_off,
                                  const double *tx, const double *ty, const double *tz,
                                  double eps, double *pot, double *acc, double *wsum,
                                  void *workspace, size_t workspace_bytes, void *stream);
/* The halo potential-energy form: the tests are the sources themselves (halo h's
 * members src_xyzm rows [offsets[h], offsets[h+1]), packed as above), so pot[k], acc and
 * wsum are the sums of member k over the other members of its halo (the coordinate rule
 * drops k itself).  Outputs indexed like the rows.  This is the forward (pot) and
 * backward (acc: dW/dx_k = -G m_k acc_k for W_h = -G/2 sum m_k pot_k) of the halo
 * potential energy. */
size_t halma_soft_self_workspace_size(int64_t n_halo, const int64_t *offsets);
int halma_pot_soft_self_batched(int64_t n_halo, const int64_t *offsets, const float *src_xyzm,
                                double eps, double *pot, double *acc, double *wsum,
                                void *workspace, size_t workspace_bytes, void *stream);

/* Tuning knobs of the FP32 pair kernels (set by the autotuner; read back with get).
 *   HALMA_TUNE_TESTS_PER_LANE : test particles per lane (1, 2, 4)
 *   HALMA_TUNE_SUBTILE        : sources per plain FP32 sub-tile before the Kahan add
 *                               (4 or 8)
 *   HALMA_TUNE_MIN_CHUNK      : smallest source chunk when a halo's source range is
 *                               split across waves (multiple of 64, >= 64)
 *   HALMA_TUNE_ITEMS_PER_WAVE : target queue items per resident wave
 *   HALMA_TUNE_WAVES_PER_CU   : persistent waves launched per compute unit (4..32)
 *   HALMA_TUNE_SYM_MAX        : largest halo whose full unbinding passes use the symmetric
 *                               kernel (0..65536; its partial slots take (n/512 + 1) * n * 8
 *                               bytes); larger haloes use the chunked kernel */
#define HALMA_TUNE_TESTS_PER_LANE 1
#define HALMA_TUNE_SUBTILE 2
#define HALMA_TUNE_MIN_CHUNK 3
#define HALMA_TUNE_ITEMS_PER_WAVE 4
#define HALMA_TUNE_WAVES_PER_CU 5
#define HALMA_TUNE_SYM_MAX 6
int halma_tune_set(int key, int value);
int halma_tune_get(int key);

#ifdef __cplusplus
}
#endif
#endif /* HALMA_H */
