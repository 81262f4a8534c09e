"""Substructure (bridge) splitting oracle -- test infrastructure only.

Restates the SUBSTRUCTURE IDENTIFICATION BLOCK of pyHALMA.py:703-766 and the
half-mass radius of halo_properties.py:366-392, on the oracle's own FoF and
centre-of-mass routines.

Points where the reference leaves the result open (parity unpinned there):
  * ``half_mass_radius`` sorts with ``np.argsort`` (an unstable quicksort): among equal
    distances the mass order, and so the radius at a crossing inside a tie, is not
    fixed.  Here ties keep member order (a stable sort).
  * a halo whose every member goes to new groups makes the reference call ``np.min`` on
    an empty array (an exception); here its loop stops and it ends empty.
"""
import math

import numpy as np

from .reductions import center_of_mass
from .fof import labels_to_groups
from .native import fof_labels_c


def half_mass_radius(cx, cy, cz, M, part_list, st_x, st_y, st_z, st_mass):
    """halo_properties.py:366-392: the distance to the centre of the last member
    (members by increasing distance) before the running mass first exceeds M / 2; the
    smallest distance if the first member already does; the largest if none does."""
    p = np.asarray(part_list, np.int64)
    if p.size == 0:
        return 0.0
    dx = cx - np.asarray(st_x, np.float64)[p]
    dy = cy - np.asarray(st_y, np.float64)[p]
    dz = cz - np.asarray(st_z, np.float64)[p]
    rad = np.sqrt((dx * dx + dy * dy) + dz * dz)
    order = np.argsort(rad, kind="stable")
    rad_sorted = rad[order]
    m_sorted = np.asarray(st_mass, np.float64)[p][order]
    rad05 = float(rad_sorted[0])
    mass_sum = 0.0
    for ip in range(p.size):
        mass_sum += float(m_sorted[ip])
        if mass_sum > 0.5 * M:
            break
        rad05 = float(rad_sorted[ip])
    return rad05


def sub_fof(part_list, st_x, st_y, st_z, ll):
    """fof.friends_of_friends_serial on the members (pyHALMA.py:723): groups of local
    indices into part_list, ordered by their smallest index, members ascending."""
    p = np.asarray(part_list, np.int64)
    if p.size == 0:
        return []
    lab = fof_labels_c(np.asarray(st_x)[p], np.asarray(st_y)[p], np.asarray(st_z)[p], ll)
    return labels_to_groups(lab, 2)


def split_substructure(groups, st_x, st_y, st_z, st_mass, ll, minp_sub, dec_factor, min_ll,
                       dist_factor, minp):
    """pyHALMA.py:703-766.  ``groups``: list of ascending particle-index arrays.  Returns
    (list of groups after splitting and the final ``good_groups(..., minp)`` filter,
    number of substructures found)."""
    groups_before = [np.asarray(g, np.int64) for g in groups]
    out = [g.tolist() for g in groups_before]
    counter = 0
    for ihal, g0 in enumerate(groups_before):
        part_list = g0.copy()
        starting = len(part_list)
        new_groups = []
        sub_ll = ll / dec_factor
        while sub_ll > min_ll:
            if part_list.size == 0:
                break
            cx, cy, cz, mass = center_of_mass(part_list, st_x, st_y, st_z, st_mass)
            r12 = half_mass_radius(cx, cy, cz, mass, part_list, st_x, st_y, st_z, st_mass)
            subs = [s for s in sub_fof(part_list, st_x, st_y, st_z, sub_ll) if len(s) >= minp_sub]
            for grp in subs:
                frac = len(grp) / len(part_list)
                if 0.6 > frac > 0.01:
                    members = part_list[np.asarray(grp, np.int64)]
                    cxs, cys, czs, _ = center_of_mass(members, st_x, st_y, st_z, st_mass)
                    dist = math.sqrt((cx - cxs) ** 2 + (cy - cys) ** 2 + (cz - czs) ** 2)
                    if dist > dist_factor * r12:
                        new_groups.append(members.tolist())
                        counter += 1
            for ng in new_groups:
                part_list = np.setdiff1d(part_list, ng)
            sub_ll = sub_ll / dec_factor
        if starting == 0 or len(part_list) / starting < 0.3:
            out[ihal] = []
        else:
            out[ihal] = part_list.tolist()
        if new_groups:
            out.extend(new_groups)
    return [g for g in out if len(g) >= minp], counter
