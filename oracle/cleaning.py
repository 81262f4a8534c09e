"""Phase-space cleaning oracle -- test infrastructure only.

Restates calc_cell and clean_cell (halo_properties.py:113-278) and the per-halo driver
of pyHALMA.py:789-843, following their loops and quirks:
  * the grid is np.arange(-(R + ll), R + ll, 2 ll) about the centre of mass, each member
    in the nearest node per axis (np.argmin: the first of equal distances);
  * the first cleaning round scales the deviation only in cells of more than two members,
    later rounds in cells of two or more;
  * in rounds after the first, the decision about the ip-th remaining member clears entry
    ip of the control array (which indexes the original member list), and ``remaining``
    is decremented for every removal counted, as the reference does;
  * ``fac_sig``: 2 in the first round, then x 0.75 (no floor after the first round),
    then x 0.75 with a floor of 1.

Sums: the reference adds each cell's terms in member order in FP64.  Here every cell
sum is correctly rounded (``math.fsum``), which the GPU's double-double sums reproduce;
against the reference's own rounding the sums agree to a few ulp, so a member whose
test value lies that close to the threshold could be decided differently there
(parity within that tolerance).  v * m is formed in FP32 when both arrays are FP32, as
numba does.  Centres use the oracle's correctly rounded ``center_of_mass``.
"""
import math

import numpy as np

from .reductions import center_of_mass, furthest_particle


def grid_of(R, ll):
    return np.arange(-(R + ll), R + ll, 2 * ll)


def _vm(v, m, i):
    if v.dtype == np.float32 and m.dtype == np.float32:
        return float(np.float32(v[i]) * np.float32(m[i]))
    return float(np.float64(v[i]) * np.float64(m[i]))


def calc_cell(cx, cy, cz, grid, part_list, st_x, st_y, st_z, st_vx, st_vy, st_vz, st_mass):
    """halo_properties.py:113-159.  Returns (cells, which): cells maps (ix, iy, iz) to
    (vcm (3 floats), mass, count, sigma); which is an (n, 3) int array of node indices."""
    p = np.asarray(part_list, np.int64)
    grid = np.asarray(grid, np.float64)
    vx, vy, vz, m = (np.asarray(a) for a in (st_vx, st_vy, st_vz, st_mass))
    which = np.zeros((p.size, 3), np.int64)
    for a, (arr, c) in enumerate(((st_x, cx), (st_y, cy), (st_z, cz))):
        h = np.asarray(arr, np.float64)[p] - c
        which[:, a] = np.abs(grid[None, :] - h[:, None]).argmin(axis=1) if p.size else 0
    members = {}
    for ip in range(p.size):
        members.setdefault(tuple(which[ip]), []).append(int(p[ip]))
    cells = {}
    for key, idx in members.items():
        mass = math.fsum(float(m[i]) for i in idx)
        vc = [math.fsum(_vm(v, m, i) for i in idx) / mass for v in (vx, vy, vz)]
        terms = []
        for i in idx:
            dx = vc[0] - float(vx[i])
            dy = vc[1] - float(vy[i])
            dz = vc[2] - float(vz[i])
            terms.append((dx * dx + dy * dy) + dz * dz)
        q = float(len(idx))
        cells[key] = (vc, mass, q, math.sqrt(math.fsum(terms) / 3.0 / q))
    return cells, which


def _removed(cells, cell, vel, q_fil, fac, sig_fil, first):
    vc, _, q, sig = cells[cell]
    if not q > q_fil:
        return True
    dx = vc[0] - vel[0]
    dy = vc[1] - vel[1]
    dz = vc[2] - vel[2]
    bas = math.sqrt((dx * dx + dy * dy) + dz * dz)
    scaled = q > 2 if first else q >= 2
    if scaled:
        bas = bas / sig if sig != 0.0 else (math.inf if bas > 0 else math.nan)
    else:
        bas = 0.0
    return bas >= fac * sig_fil


def clean_cell(part_list, st_x, st_y, st_z, st_vx, st_vy, st_vz, st_mass, cells, which, ll,
               q_fil, sig_fil):
    """halo_properties.py:162-278 after the driver's first calc_cell: returns
    (cx, cy, cz, M, R, control, which of the last calc_cell, rounds run)."""
    part_list = np.asarray(part_list, np.int64)
    vx, vy, vz = (np.asarray(a) for a in (st_vx, st_vy, st_vz))
    npart = part_list.size
    control = np.ones(npart, np.int32)
    cleaned = 0
    remaining = npart
    fac_sig = 2.0
    for ip in range(npart):
        i = part_list[ip]
        if _removed(cells, tuple(which[ip]), (float(vx[i]), float(vy[i]), float(vz[i])), q_fil,
                    fac_sig, sig_fil, True):
            control[ip] = 0
            cleaned += 1
            remaining -= 1
    clean = part_list[np.nonzero(control)[0]]
    fac_sig *= 0.75
    rounds = 1
    if remaining > 0:
        frac = cleaned / remaining
        while frac > 0.1:
            cx, cy, cz, _ = center_of_mass(clean, st_x, st_y, st_z, st_mass)
            R = furthest_particle(cx, cy, cz, clean, st_x, st_y, st_z)
            cells, which = calc_cell(cx, cy, cz, grid_of(R, ll), clean, st_x, st_y, st_z, st_vx,
                                     st_vy, st_vz, st_mass)
            cleaned = 0
            for ip in range(clean.size):
                i = clean[ip]
                if _removed(cells, tuple(which[ip]), (float(vx[i]), float(vy[i]), float(vz[i])),
                            q_fil, fac_sig, sig_fil, False):
                    control[ip] = 0
                    cleaned += 1
                    remaining -= 1
            clean = part_list[np.nonzero(control)[0]]
            rounds += 1
            if remaining == 0:
                return 0.0, 0.0, 0.0, 0.0, 0.0, control, which, rounds
            frac = cleaned / remaining
            fac_sig *= 0.75
            fac_sig = max(1.0, fac_sig)
    cx, cy, cz, M = center_of_mass(clean, st_x, st_y, st_z, st_mass)
    R = furthest_particle(cx, cy, cz, clean, st_x, st_y, st_z)
    cells, which = calc_cell(cx, cy, cz, grid_of(R, ll), clean, st_x, st_y, st_z, st_vx, st_vy,
                             st_vz, st_mass)
    return cx, cy, cz, M, R, control, which, rounds


def phase_space_clean(groups, st_x, st_y, st_z, st_vx, st_vy, st_vz, st_mass, ll, q_fil,
                      sig_fil, psc=True, min_part=50, minp=50):
    """The cleaning block of pyHALMA.py:789-843 over a list of groups.  Returns, per input
    group, a dict with the control array, the clean_cell results, the initial grid's
    n_cell and the last calc_cell's node indices; and the kept groups (npart >= minp)
    with the driver's recomputed centre, mass and radius."""
    per_group, kept = [], []
    for part_list in groups:
        part_list = np.asarray(part_list, np.int64)
        cx, cy, cz, _ = center_of_mass(part_list, st_x, st_y, st_z, st_mass)
        R = furthest_particle(cx, cy, cz, part_list, st_x, st_y, st_z)
        grid = grid_of(R, ll)
        cells, which = calc_cell(cx, cy, cz, grid, part_list, st_x, st_y, st_z, st_vx, st_vy,
                                 st_vz, st_mass)
        ccx, ccy, ccz, cM, cR, control, which, rounds = clean_cell(
            part_list, st_x, st_y, st_z, st_vx, st_vy, st_vz, st_mass, cells, which, ll, q_fil,
            sig_fil)
        raw = control.astype(bool)
        if not psc or part_list.size < min_part:
            control = np.ones(part_list.size, np.int32)
        control = control.astype(bool)
        members = part_list[control]
        fcx, fcy, fcz, fM = center_of_mass(members, st_x, st_y, st_z, st_mass)
        fR = furthest_particle(fcx, fcy, fcz, members, st_x, st_y, st_z)
        per_group.append(dict(control=control, control_clean=raw, clean=(ccx, ccy, ccz, cM, cR),
                              n_cell=len(grid), which=which, rounds=rounds))
        if members.size >= minp:
            kept.append(dict(members=members, centre=(fcx, fcy, fcz), mass=fM, rmax=fR,
                             n_cell=len(grid)))
    return per_group, kept
