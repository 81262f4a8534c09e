"""Source gathering oracle -- test infrastructure only.

Restates the source gathering of the gas path (halo_gas.py:9-52 patch_to_particles,
:56-141 AMRgrid_to_particles, :218-220 parallel_inside, :223-277
st_gas_dm_particles_inside) and the dark-matter mass inside a sphere of
pyHALMA.py:1172-1185:

  * a gas particle per AMR cell of a level >= 1 patch whose centre lies strictly inside
    the box [c - R, c + R] per axis and that is neither refined nor overlapped
    (cr0amr and solapst), cells in (ix, iy, iz) loop order, patches in ascending index;
    centre x0 + ix res with x0 = rx - res / 2, res = (L / ncoarse) / 2^level, mass
    (1 + delta) rho_B res^3, velocity x 3e5; then only the particles with
    sqrt(dx^2 + dy^2 + dz^2) < R, their mass x rete^3;
  * dark matter and stars within R of the centre by scipy's KD-tree
    (query_ball_point: distance <= R) -- scipy is the reference's own dependency and is
    used here as the oracle, with a brute-force check beside it; the indices are
    returned in ascending order (the KD-tree's own order for a single point is
    unspecified).

``res ** 3`` is Python's float power; the reference's compiled ``patch_res**3`` may round
its last bit differently (masses agree to 1 ulp there, parity unpinned below that).

The reference selects the patches with masclet_framework's which_patches_inside_box
and numbers their levels with create_vector_levels (masclet_framework, unpinned, not
vendored).  Here the caller passes each patch's level, and every level >= 1 patch is
visited: a patch that the reference's selection skips has no cell inside the box, so the
output is the same set in the same order as long as that selection keeps every patch
with a cell inside the box.  The unit constant masclet units.mass_to_sun (DM and star
masses) is a parameter.
"""
import numpy as np


def patch_to_particles(res, rx, ry, rz, nx, ny, nz, delta, cr0amr, solapst, vx, vy, vz, temp,
                       box, rho_B):
    """halo_gas.py:9-52: arrays (x, y, z, vx, vy, vz, mass, temp) of the patch's cells."""
    out = [[] for _ in range(8)]
    x0 = rx - res / 2
    y0 = ry - res / 2
    z0 = rz - res / 2
    for ix in range(nx):
        x = x0 + ix * res
        if not (x > box[0] and x < box[1]):
            continue
        for iy in range(ny):
            y = y0 + iy * res
            if not (y > box[2] and y < box[3]):
                continue
            for iz in range(nz):
                z = z0 + iz * res
                if not (z > box[4] and z < box[5]):
                    continue
                if cr0amr[ix, iy, iz] and solapst[ix, iy, iz]:
                    vals = (x, y, z, float(vx[ix, iy, iz]), float(vy[ix, iy, iz]),
                            float(vz[ix, iy, iz]),
                            (1 + float(delta[ix, iy, iz])) * rho_B * res ** 3,
                            float(temp[ix, iy, iz]))
                    for a, v in zip(out, vals):
                        a.append(v)
    return tuple(np.array(a, np.float64) for a in out)


def amr_to_particles(L, ncoarse, levels, rx, ry, rz, gas, R, cx, cy, cz, rho_B):
    """halo_gas.py:56-141 with the patches given explicitly: ``levels``, ``rx/ry/rz`` per
    patch, ``gas`` = (delta, vx, vy, vz, temp, cr0amr, solapst), each a list of per-patch
    (nx, ny, nz) arrays."""
    box = np.array([cx - R, cx + R, cy - R, cy + R, cz - R, cz + R])
    parts = [[] for _ in range(8)]
    for p, lev in enumerate(levels):
        if lev < 1:
            continue
        res = (L / ncoarse) / 2 ** int(lev)
        delta, vx, vy, vz, temp, cr, sol = (g[p] for g in gas)
        nx, ny, nz = delta.shape
        got = patch_to_particles(res, float(rx[p]), float(ry[p]), float(rz[p]), nx, ny, nz,
                                 delta, cr, sol, vx, vy, vz, temp, box, rho_B)
        for a, v in zip(parts, got):
            a.append(v)
    out = [np.concatenate(a) if a else np.zeros(0) for a in parts]
    for k in (3, 4, 5):
        out[k] = out[k] * 3e5
    return tuple(out)


def inside(x, y, z, R, cx, cy, cz):
    """halo_gas.py:218-220."""
    return np.sqrt((x - cx) ** 2 + (y - cy) ** 2 + (z - cz) ** 2) < R


def ball(x, y, z, cx, cy, cz, R):
    """Indices (ascending) of the points within R of the centre, by scipy's KD-tree
    (query_ball_point, as pyHALMA.py:545, :695 build them)."""
    from scipy.spatial import KDTree
    data = np.array((x, y, z)).T.astype(np.float64)
    return np.array(sorted(KDTree(data).query_ball_point([cx, cy, cz], R)), np.int64)


def ball_brute(x, y, z, cx, cy, cz, R):
    """The same set by brute force: (dx^2 + dy^2) + dz^2 <= R^2 in FP64."""
    dx = np.asarray(x, np.float64) - cx
    dy = np.asarray(y, np.float64) - cy
    dz = np.asarray(z, np.float64) - cz
    return np.nonzero((dx * dx + dy * dy) + dz * dz <= R * R)[0]


def gather_sources(rete, L, ncoarse, levels, rx, ry, rz, gas, dm, st, cx, cy, cz, R, rho_B,
                   mass_to_sun):
    """halo_gas.py:223-277 for one halo: (gas x, y, z, vx, vy, vz, mass, temp, dm x, y, z,
    mass, st x, y, z, mass, oripa); dm = (x, y, z, mass), st = (x, y, z, mass, oripa) in
    code units as masclet reads them."""
    g = amr_to_particles(L, ncoarse, levels, rx, ry, rz, gas, R, cx, cy, cz, rho_B)
    keep = inside(g[0], g[1], g[2], R, cx, cy, cz)
    g = [a[keep] for a in g]
    g[6] = g[6] * rete ** 3
    di = ball(dm[0], dm[1], dm[2], cx, cy, cz, R)
    si = ball(st[0], st[1], st[2], cx, cy, cz, R)
    dm_out = [np.asarray(a)[di] for a in dm[:3]] + [np.asarray(dm[3])[di] * mass_to_sun]
    st_out = ([np.asarray(a)[si] for a in st[:3]] + [np.asarray(st[3])[si] * mass_to_sun]
              + [np.asarray(st[4])[si]])
    return tuple(g) + tuple(dm_out) + tuple(st_out)
