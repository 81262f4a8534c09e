"""Per-group reduction oracles (test infrastructure only).

Restatements of halo_properties.py:16-24 (total_mass), :26-43 (center_of_mass),
:45-60 (CM_velocity), :98-111 (furthest_particle) and :518-556 (density_peak).
The reference runs the sums in numba ``prange`` with ``fastmath=True`` (unspecified
order), so the sums here are correctly rounded (``math.fsum``) and parity with them is
tolerance-based; ``density_peak`` follows the reference's serial loop exactly.
"""
import math

import numpy as np


def _idx(part_list):
    return np.asarray(part_list, dtype=np.int64)


def total_mass(part_list, st_mass):
    """halo_properties.py:16-24."""
    p = _idx(part_list)
    return math.fsum(np.asarray(st_mass, np.float64)[p].tolist())


def center_of_mass(part_list, st_x, st_y, st_z, st_mass):
    """halo_properties.py:26-43: (cx, cy, cz, M), or zeros when M <= 0."""
    p = _idx(part_list)
    m = np.asarray(st_mass, np.float64)[p]
    M = math.fsum(m.tolist())
    if not M > 0:
        return 0.0, 0.0, 0.0, 0.0
    c = [math.fsum((m * np.asarray(a, np.float64)[p]).tolist()) / M
         for a in (st_x, st_y, st_z)]
    return c[0], c[1], c[2], M


def cm_velocity(M, part_list, st_vx, st_vy, st_vz, st_mass):
    """halo_properties.py:45-60: mass-weighted mean velocity, zeros when M <= 0."""
    p = _idx(part_list)
    m = np.asarray(st_mass, np.float64)[p]
    if not M > 0:
        return 0.0, 0.0, 0.0
    return tuple(math.fsum((m * np.asarray(a, np.float64)[p]).tolist()) / M
                 for a in (st_vx, st_vy, st_vz))


def furthest_particle(cx, cy, cz, part_list, st_x, st_y, st_z):
    """halo_properties.py:98-111: max |x - c| (0 for an empty list)."""
    p = _idx(part_list)
    if p.size == 0:
        return 0.0
    dx = cx - np.asarray(st_x, np.float64)[p]
    dy = cy - np.asarray(st_y, np.float64)[p]
    dz = cz - np.asarray(st_z, np.float64)[p]
    d = np.sqrt((dx * dx + dy * dy) + dz * dz)
    return float(max(0.0, d.max()))


def linspace_ref(start, stop, num=50):
    """np.linspace(start, stop, num) as used at halo_properties.py:521-523."""
    return np.linspace(float(start), float(stop), num)


def density_peak(part_list, st_x, st_y, st_z, st_mass, ll, n_cells=50):
    """halo_properties.py:518-556 restated: nearest-node binning on a 50^3 grid spanning
    [min - ll, max + ll] per axis (np.linspace), mass summed per node in particle order,
    the first node in (ix, iy, iz) order holding the largest mass strictly above zero
    wins; the node coordinates are returned (not a particle position).  If no node
    holds positive mass the centre node (n_cells // 2 per axis) is returned.  An empty
    part_list raises ValueError (np.min of an empty array)."""
    p = _idx(part_list)
    if p.size == 0:
        raise ValueError("zero-size array to reduction operation minimum which has no identity")
    ax = []
    for a in (st_x, st_y, st_z):
        v = np.asarray(a)[p]
        ax.append((v.astype(np.float64), linspace_ref(float(v.min()) - ll,
                                                        float(v.max()) + ll, n_cells)))
    idx = []
    for v, g in ax:
        idx.append(np.argmin(np.abs(g[None, :] - v[:, None]), axis=1))
    m = np.asarray(st_mass, np.float64)[p]
    cell = np.zeros((n_cells, n_cells, n_cells))
    np.add.at(cell, (idx[0], idx[1], idx[2]), m)  # sequential, in particle order
    flat = cell.ravel()
    best = int(np.argmax(flat))  # first index of the maximum in C order
    if not flat[best] > 0.0:
        ix = iy = iz = n_cells // 2
    else:
        ix, rem = divmod(best, n_cells * n_cells)
        iy, iz = divmod(rem, n_cells)
    return ax[0][1][ix], ax[1][1][iy], ax[2][1][iz]
