"""halo_shape and sigma_projections oracle -- test infrastructure only.

Restates particle_subroutines.f90:160-214 (halo_shape: mass-weighted second-moment
tensor, Jacobi eigenvalues, square roots in decreasing order) and :217-461
(sigma_projections: projected 2D cell grids of mean velocity, mass, count and 1D
dispersion; mean dispersion inside R05 per projection; V/sigma and lambda over the
cells within R05 + 2 ll), in the reference's FP32 arithmetic for every single operation.

The reference accumulates its per-particle sums in FP32 through OpenMP reductions, in an
order that varies with the thread schedule; here those sums are FP64 (math.fsum) rounded
to FP32 at the end, so the reference's own results scatter around these by its FP32
summation error (parity to that tolerance).  The sums over cells at the end are
sequential FP32 in the reference's loop order, as there.  The Jacobi iteration follows
the file's variant of the cyclic method (Numerical Recipes' jacobi with a stopping rule
relative to the initial element sum, threshold 0.2 S^2 for the first three sweeps,
negligible elements zeroed after the fourth) in FP64, since its early stop decides the
last bits of the eigenvalues.
"""
import math

import numpy as np

f32 = np.float32


def jacobi_eigenvalues(mat):
    """Eigenvalues of a symmetric 3x3 FP32 matrix by the file's Jacobi variant
    (particle_subroutines.f90:11-122), returned as FP32 in diagonal order."""
    a = np.array(mat, np.float64).reshape(3, 3)
    n = 3
    b = np.diag(a).copy()
    d = b.copy()
    z = np.zeros(3)
    total = sum(abs(a[p, q]) for p in range(n) for q in range(p, n))
    for sweep in range(1, 101):
        sm = sum(abs(a[p, q]) for p in range(n - 1) for q in range(p + 1, n))
        if sm < float(f32(1e-4)) * total:
            break
        tresh = 0.2 * sm ** 2 if sweep < 4 else 0.0
        for p in range(n - 1):
            for q in range(p + 1, n):
                g = 100.0 * abs(a[p, q])
                if sweep > 4 and abs(d[p]) + g == abs(d[p]) and abs(d[q]) + g == abs(d[q]):
                    a[p, q] = 0.0
                elif abs(a[p, q]) > tresh:
                    h = d[q] - d[p]
                    if abs(h) + g == abs(h):
                        t = a[p, q] / h
                    else:
                        theta = 0.5 * h / a[p, q]
                        t = 1.0 / (abs(theta) + math.sqrt(1.0 + theta ** 2))
                        if theta < 0.0:
                            t = -t
                    c = 1.0 / math.sqrt(1.0 + t ** 2)
                    s = t * c
                    tau = s / (1.0 + c)
                    h = t * a[p, q]
                    z[p] -= h
                    z[q] += h
                    d[p] -= h
                    d[q] += h
                    a[p, q] = 0.0
                    for j in range(p):
                        g1, h1 = a[j, p], a[j, q]
                        a[j, p] = g1 - s * (h1 + g1 * tau)
                        a[j, q] = h1 + s * (g1 - h1 * tau)
                    for j in range(p + 1, q):
                        g1, h1 = a[p, j], a[j, q]
                        a[p, j] = g1 - s * (h1 + g1 * tau)
                        a[j, q] = h1 + s * (g1 - h1 * tau)
                    for j in range(q + 1, n):
                        g1, h1 = a[p, j], a[q, j]
                        a[p, j] = g1 - s * (h1 + g1 * tau)
                        a[q, j] = h1 + s * (g1 - h1 * tau)
        b += z
        d = b.copy()
        z[:] = 0.0
    return d.astype(f32)


def halo_shape(x, y, z, mass):
    """particle_subroutines.f90:160-214: sqrt of the eigenvalues of the mass-weighted
    second-moment tensor of the (already centred) points, largest first, FP32."""
    x, y, z, m = (np.asarray(a, f32) for a in (x, y, z, mass))
    r = (x, y, z)
    tensor = np.zeros((3, 3), f32)
    for i in range(3):
        for j in range(3):
            tensor[i, j] = f32(math.fsum(((m * r[i]) * r[j]).astype(np.float64).tolist()))
    mass_tot = f32(math.fsum(m.astype(np.float64).tolist()))
    tensor = tensor / mass_tot
    ev = sorted(jacobi_eigenvalues(tensor).tolist(), reverse=True)
    return np.sqrt(np.array(ev, f32))


def _nearest(grid, v):
    return int(np.argmin(np.abs(grid - v)))  # first of equal distances (Fortran minloc)


def sigma_projections(grid, part_list, st_x, st_y, st_z, st_vx, st_vy, st_vz, st_mass,
                      cx, cy, cz, R05x, R05y, R05z, ll):
    """particle_subroutines.f90:217-461 with 0-based part_list and FP32 inputs (the
    velocities already relative to the halo's, as halo_properties.py:791-793 passes
    them): (SIG_1D_x_05, SIG_1D_y_05, SIG_1D_z_05, V_sigma, lambda), FP32."""
    grid = np.asarray(grid, f32)
    n = grid.size
    cx, cy, cz, R05x, R05y, R05z, ll = (f32(v) for v in (cx, cy, cz, R05x, R05y, R05z, ll))
    p = np.asarray(part_list, np.int64)
    X, Y, Z, VX, VY, VZ, M = (np.asarray(a, f32)[p] for a in (st_x, st_y, st_z, st_vx, st_vy,
                                                               st_vz, st_mass))
    dx, dy, dz = X - cx, Y - cy, Z - cz
    ix = np.array([_nearest(grid, v) for v in dx], np.int64)
    iy = np.array([_nearest(grid, v) for v in dy], np.int64)
    iz = np.array([_nearest(grid, v) for v in dz], np.int64)
    proj = {"x": (iy, iz, VX), "y": (ix, iz, VY), "z": (ix, iy, VZ)}
    vcm, sd, q, sig = {}, {}, {}, {}
    for k, (a, b, v) in proj.items():
        acc_v = [[[] for _ in range(n)] for _ in range(n)]
        acc_m = [[[] for _ in range(n)] for _ in range(n)]
        for t in range(p.size):
            acc_v[a[t]][b[t]].append(float(v[t] * M[t]))
            acc_m[a[t]][b[t]].append(float(M[t]))
        sv = np.array([[f32(math.fsum(acc_v[i][j])) for j in range(n)] for i in range(n)], f32)
        sm = np.array([[f32(math.fsum(acc_m[i][j])) for j in range(n)] for i in range(n)], f32)
        qq = np.array([[len(acc_m[i][j]) for j in range(n)] for i in range(n)], np.int64)
        with np.errstate(divide="ignore", invalid="ignore"):
            vc = np.where(sm != 0, sv / sm, sv).astype(f32)
        acc_s = [[[] for _ in range(n)] for _ in range(n)]
        for t in range(p.size):
            dv = f32(v[t] - vc[a[t], b[t]])
            acc_s[a[t]][b[t]].append(float(f32(dv * dv)))
        ss = np.array([[f32(math.fsum(acc_s[i][j])) for j in range(n)] for i in range(n)], f32)
        with np.errstate(divide="ignore", invalid="ignore"):
            sg = np.where(qq != 0, np.sqrt((ss / qq.astype(f32)).astype(f32)), ss).astype(f32)
        vcm[k], sd[k], q[k], sig[k] = vc, sm, qq, sg
    dist = {"x": np.sqrt(dy * dy + dz * dz), "y": np.sqrt(dx * dx + dz * dz),
            "z": np.sqrt(dx * dx + dy * dy)}
    R05 = {"x": R05x, "y": R05y, "z": R05z}
    sig05 = {}
    for k, (a, b, _) in proj.items():
        sel = dist[k] < R05[k]
        cnt = int(sel.sum())
        tot = f32(math.fsum(sig[k][a[sel], b[sel]].astype(np.float64).tolist()))
        sig05[k] = f32(tot / f32(cnt)) if cnt > 0 else f32(0.0)
    vs, lam = {}, {}
    # planes: (projection, grid axis of the inner loop, of the outer loop)
    for k in ("z", "y", "x"):
        sumV = sumS = sumU = sumD = f32(0.0)
        for o in range(n):          # outer loop (second index)
            for i in range(n):      # inner loop (first index)
                rbin = f32(np.sqrt(f32(grid[i] * grid[i]) + f32(grid[o] * grid[o])))
                if rbin < f32(R05[k] + f32(2 * ll)):
                    v, s, m = vcm[k][i, o], sig[k][i, o], sd[k][i, o]
                    sumV = f32(sumV + f32(f32(v * v) * m))
                    sumS = f32(sumS + f32(f32(s * s) * m))
                    sumU = f32(sumU + f32(f32(m * rbin) * abs(v)))
                    sumD = f32(sumD + f32(f32(m * rbin) * f32(np.sqrt(f32(f32(v * v) + f32(s * s))))))
        vs[k] = f32(np.sqrt(f32(sumV / sumS))) if sumS > 0 else f32(0.0)
        lam[k] = f32(sumU / sumD) if sumD > 0 else f32(0.0)
    V_sigma = f32(f32(f32(vs["x"] + vs["y"]) + vs["z"]) / f32(3.0))
    lam_all = f32(f32(f32(lam["x"] + lam["y"]) + lam["z"]) / f32(3.0))
    return sig05["x"], sig05["y"], sig05["z"], V_sigma, lam_all
