"""Drop-in for ``python_scripts.fof`` (friends-of-friends), on the GPU.

Same names, arguments and return types as the reference:

  friends_of_friends(data, linking_length)                        fof.py:18-79
  friends_of_friends_serial(st_x, st_y, st_z, linking_length)     fof.py:87-89
  friends_of_friends_parallel(st_x, st_y, st_z, linking_length, L, minp, st_mass)
                                                                   fof.py:95-176
  pyfof_friends_of_friends_serial(st_x, st_y, st_z, linking_length)   fof.py:184-187
  pyfof_friends_of_friends_parallel(st_x, st_y, st_z, linking_length, L, minp, st_mass)
                                                                   fof.py:192-272

The groups are the connected components of the graph with an edge wherever
((dx*dx + dy*dy) + dz*dz) <= ll*ll in FP64 -- the test scipy's KDTree applies for the
reference (coordinates widened to FP64, inclusive, non-periodic) -- computed by
libhalma's cell-binned union-find.  ``friends_of_friends`` returns the groups of at
least two members as sorted lists, ordered by their smallest member, exactly as the
reference's breadth-first search emits them.

The parallel variants keep the reference's domain decomposition (cubes of side L/n with
a 0.1 Mpc buffer, a strict ``> minp`` size filter and the centre-of-mass ownership test),
so they return what the reference returns, including its duplicates and splits at
subdomain borders; the subdomain count follows fof.py:98-99 from ``ncores`` (default: the
host's CPU count, numba's default thread count).

The pyfof variants replace a third-party C++ library that is not part of the reference
tree; they return every component (singletons included) as sorted lists ordered by the
smallest member.  pyfof's own member order and tie rule could not be checked here.
"""
import os

import numpy as np

from . import engine


def _components(x, y, z, linking_length, min_size):
    x = np.asarray(x)
    if x.shape[0] == 0:
        return []
    _, groups = engine.fof_groups(x, y, z, float(linking_length), min_size=min_size)
    return groups


def friends_of_friends(data, linking_length):
    data = np.asarray(data)
    if data.ndim != 2 or data.shape[1] != 3:
        raise ValueError("data must have shape (N, 3)")
    if data.dtype not in (np.float32, np.float64):
        data = data.astype(np.float64)
    g = _components(data[:, 0], data[:, 1], data[:, 2], linking_length, 2)
    return g.to_lists() if g else []


def friends_of_friends_serial(st_x, st_y, st_z, linking_length):
    data = np.array((st_x, st_y, st_z)).T
    return friends_of_friends(data, linking_length)


def _ncores(ncores):
    if ncores is None:
        ncores = os.cpu_count() or 1
    mandatories = np.array([1, 8, 64, 512])
    return int(np.max(mandatories[mandatories <= ncores]))


def _parallel(st_x, st_y, st_z, linking_length, L, minp, st_mass, ncores, min_size):
    from .halo_properties import _reduce_one

    st_x = np.asarray(st_x)
    st_y = np.asarray(st_y)
    st_z = np.asarray(st_z)
    n = _ncores(ncores)
    nsides = int(np.round(n ** (1 / 3)))
    sub_L = L / nsides
    buffer = 0.1  # Mpc, fof.py:114
    boxes = []
    for ix in range(nsides):
        for iy in range(nsides):
            for iz in range(nsides):
                lo = (-L / 2 + ix * sub_L, -L / 2 + iy * sub_L, -L / 2 + iz * sub_L)
                hi = (-L / 2 + (ix + 1) * sub_L, -L / 2 + (iy + 1) * sub_L,
                      -L / 2 + (iz + 1) * sub_L)
                boxes.append((lo, hi))
    everyone = np.arange(st_x.shape[0])
    final = []
    for lo, hi in boxes:
        mask = ((st_x >= lo[0] - buffer) & (st_x < hi[0] + buffer)
                & (st_y >= lo[1] - buffer) & (st_y < hi[1] + buffer)
                & (st_z >= lo[2] - buffer) & (st_z < hi[2] + buffer))
        sub = everyone[mask]
        if sub.size == 0:
            continue
        g = _components(st_x[mask], st_y[mask], st_z[mask], linking_length, min_size)
        if not g:
            continue
        for grp in g.to_arrays():
            if len(grp) > minp:
                glob = sub[grp]
                cx, cy, cz, _ = _reduce_one(glob, st_x, st_y, st_z, st_mass)[:4]
                if (lo[0] <= cx <= hi[0] and lo[1] <= cy <= hi[1] and lo[2] <= cz <= hi[2]):
                    final.append(glob)
    return final


def friends_of_friends_parallel(st_x, st_y, st_z, linking_length, L, minp, st_mass,
                                ncores=None):
    return _parallel(st_x, st_y, st_z, linking_length, L, minp, st_mass, ncores, 2)


def pyfof_friends_of_friends_serial(st_x, st_y, st_z, linking_length):
    data = np.array((st_x, st_y, st_z)).T.astype(np.float64)
    g = _components(data[:, 0], data[:, 1], data[:, 2], linking_length, 1)
    return g.to_lists() if g else []


def pyfof_friends_of_friends_parallel(st_x, st_y, st_z, linking_length, L, minp, st_mass,
                                      ncores=None):
    return _parallel(np.asarray(st_x, np.float64), np.asarray(st_y, np.float64),
                     np.asarray(st_z, np.float64), linking_length, L, minp, st_mass, ncores, 1)


def good_groups(groups, minp):
    """pyHALMA.py:148-149: mask of the groups with at least minp members."""
    return np.array([len(groups[ig]) >= minp for ig in range(len(groups))], dtype=bool)
