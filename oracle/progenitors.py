"""Progenitor matching oracle -- test infrastructure only.

Restates main_progenitors_finder of pyHALMA.py:1219-1257 for every new halo:

  * old halo oih intersects new halo ih when np.in1d(oripas, ooripas,
    assume_unique=True) finds a common particle ID (pyHALMA.py:1230-1234);
  * mass_intersections[oih] = omm[oih] for those, 0 otherwise; the candidates are
    ranked by np.flip(np.argsort(mass_intersections)) (pyHALMA.py:1236-1237);
  * pro1 = the first + 1; merger type -1 when 1.2 * mass < omm[pro1 - 1]; otherwise,
    with two or more intersections, pro2 = the second + 1 and, with f = second / first
    mass, type 2 when 1/20 < f < 1/3 and 3 in every other case -- the type 1 set for
    f > 1/3 is overwritten by the else branch of the following if, as the reference's
    code does (pyHALMA.py:1247-1255).

np.argsort's default sort leaves the order of equal masses to the implementation; here
equal masses rank as np.flip of a stable argsort does (the higher old index first).
Cases with two old haloes of exactly equal mass competing for pro1 or pro2 are
therefore parity unpinned.
"""
import numpy as np


def main_progenitors(oripas, mass_ih, oripas_before, omm):
    """pyHALMA.py:1219-1257 for one new halo: (pro1, pro2, nmergs, merger_type)."""
    omm = np.asarray(omm, np.float64)
    mass_intersections = np.zeros(len(omm))
    nmergs = 0
    for oih, ooripas in enumerate(oripas_before):
        if np.count_nonzero(np.isin(oripas, ooripas, assume_unique=True)) > 0:
            mass_intersections[oih] = omm[oih]
            nmergs += 1
    order = np.flip(np.argsort(mass_intersections, kind="stable"))
    ranked = mass_intersections[order]
    pro1 = pro2 = mtype = 0
    if nmergs > 0:
        pro1 = int(order[0]) + 1
        if 1.2 * mass_ih < omm[pro1 - 1]:
            mtype = -1
        elif nmergs > 1:
            pro2 = int(order[1]) + 1
            frac = ranked[1] / ranked[0]
            if frac > 1 / 3:
                mtype = 1
            if 1 / 20 < frac < 1 / 3:
                mtype = 2
            else:
                mtype = 3
    return pro1, pro2, nmergs, mtype


def progenitors(new_groups, st_oripa, masses, oripas_before, omm):
    """The merger section of pyHALMA.py:1211-1266 over all new haloes: int32 arrays
    (pro1, pro2, n_mergers, merger_type)."""
    st_oripa = np.asarray(st_oripa)
    out = np.zeros((4, len(new_groups)), np.int32)
    for ih, halo in enumerate(new_groups):
        out[:, ih] = main_progenitors(st_oripa[np.asarray(halo, np.int64)], masses[ih],
                                      oripas_before, omm)
    return out[0], out[1], out[2], out[3]
