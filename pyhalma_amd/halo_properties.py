"""Drop-in for the reduction and energy-test part of ``python_scripts.halo_properties``.

Same names, arguments and return values as the reference (scalars are Python floats):

  total_mass(part_list, st_mass)                                halo_properties.py:16-24
  center_of_mass(part_list, st_x, st_y, st_z, st_mass)          :26-43
  CM_velocity(M, part_list, st_vx, st_vy, st_vz, st_mass)       :45-60
  furthest_particle(cx, cy, cz, part_list, st_x, st_y, st_z)    :98-111
  density_peak(part_list, st_x, st_y, st_z, st_mass, ll)        :518-556
  escape_velocity_unbinding_fortran(...)                        :282-361
  half_mass_radius(cx, cy, cz, M, part_list, st_x, ..., st_mass) :366-392
  calc_cell(cx, cy, cz, grid, part_list, ..., sig3D_cell)        :113-159
  clean_cell(cx, cy, cz, M, RRHH, grid, part_list, ...)          :162-278

The reference's numba ``prange`` sums have no fixed order; here they are FP64
double-double sums (within about an ulp of the correctly rounded value).  density_peak
is bit-exact to the reference's serial loop (see libhalma's halma_density_peak).
Each call here handles one group; the batched forms over all groups at once are
``pyhalma_amd.engine.group_reduce`` / ``density_peak`` / ``unbind``.

The remaining functions of the reference module (angular momentum, morphology, star
formation, dispersions, shapes, Sersic fits, SMBH) are halo-property calculations
outside the accelerated path.
"""
import numpy as np
import torch

from . import engine

_g = 4.3 * 1e-3  # (km/s)^2 pc / Msun, halo_properties.py:345
_g *= 1e-6  # (km/s)^2 Mpc / Msun, halo_properties.py:346
G_CONST = _g


def _one_group(part_list, arrays):
    """Gather one group's particles on the host and hand them to the GPU as a single
    group (members 0..n-1)."""
    p = np.asarray(part_list, dtype=np.int64).ravel()
    sub = [np.ascontiguousarray(np.asarray(a)[p]) for a in arrays]
    n = p.size
    dev = torch.device("cuda", torch.cuda.current_device())
    members = torch.arange(max(n, 1), dtype=torch.int32, device=dev)
    offsets = torch.tensor([0, n], dtype=torch.int64, device=dev)
    groups = engine.Groups(members, offsets, np.array([0, n], dtype=np.int64))
    return groups, sub


def _reduce_one(part_list, st_x, st_y, st_z, st_mass, vel=None, centre=None):
    arrays = [st_x, st_y, st_z, st_mass] + (list(vel) if vel is not None else [])
    groups, sub = _one_group(part_list, arrays)
    if groups.n_members == 0:
        return np.zeros(16)
    x, y, z, m = sub[:4]
    v = sub[4:] if vel is not None else (None, None, None)
    out = engine.group_reduce(groups, x, y, z, m, *v,
                              centres=None if centre is None else [centre])
    r = out[0].cpu().numpy()
    return (r[1], r[2], r[3], r[0]) + tuple(r[4:])


def total_mass(part_list, st_mass):
    p = np.asarray(part_list, dtype=np.int64).ravel()
    if p.size == 0:
        return 0.0
    z = np.zeros(p.size)
    groups, sub = _one_group(np.arange(p.size), [z, z, z, np.asarray(st_mass)[p]])
    out = engine.group_reduce(groups, *sub)
    return float(out[0, 0].item())


def center_of_mass(part_list, st_x, st_y, st_z, st_mass):
    if np.asarray(part_list).size == 0:
        return 0.0, 0.0, 0.0, 0.0
    r = _reduce_one(part_list, st_x, st_y, st_z, st_mass)
    return float(r[0]), float(r[1]), float(r[2]), float(r[3])


def CM_velocity(M, part_list, st_vx, st_vy, st_vz, st_mass):
    if not M > 0.0 or np.asarray(part_list).size == 0:
        return 0.0, 0.0, 0.0
    # the mass-weighted velocity sums, divided by the caller's M as the reference does
    zero = np.zeros(len(st_mass))
    r = _reduce_one(part_list, zero, zero, zero, st_mass, vel=(st_vx, st_vy, st_vz))
    Mk = r[3]
    return tuple(float(v) if Mk == M else float(v * Mk / M) for v in r[4:7])


def furthest_particle(cx, cy, cz, part_list, st_x, st_y, st_z):
    if np.asarray(part_list).size == 0:
        return 0.0
    ones = np.ones(len(st_x))
    r = _reduce_one(part_list, st_x, st_y, st_z, ones, centre=(cx, cy, cz))
    return float(r[7])


def half_mass_radius(cx, cy, cz, M, part_list, st_x, st_y, st_z, st_mass):
    """halo_properties.py:366-392 (the np.min of an empty list raises, as there)."""
    p = np.asarray(part_list)
    if p.size == 0:
        raise ValueError("zero-size array to reduction operation minimum which has no "
                         "identity")
    groups, sub = _one_group(part_list, [st_x, st_y, st_z, st_mass])
    red = torch.zeros((1, 16), dtype=torch.float64, device=groups.members.device)
    red[0, 0] = float(M)
    red[0, 1], red[0, 2], red[0, 3] = float(cx), float(cy), float(cz)
    return float(engine.half_mass_radius(groups, *sub, reduced=red)[0].item())


def density_peak(part_list, st_x, st_y, st_z, st_mass, ll):
    p = np.asarray(part_list)
    if p.size == 0:
        raise ValueError("zero-size array to reduction operation minimum which has no "
                         "identity")
    groups, sub = _one_group(part_list, [st_x, st_y, st_z, st_mass])
    out = engine.density_peak(groups, *sub, ll=float(ll))
    r = out[0].cpu().numpy()
    return float(r[0]), float(r[1]), float(r[2])


def escape_velocity_unbinding(part_list, st_x, st_y, st_z, st_vx, st_vy, st_vz, st_mass,
                              vx, vy, vz, factor_v, extra_sources=None):
    """The star energy test of halo_properties.py:333-359 with the sources already
    gathered: sources = (extra_sources gas + DM as (x, y, z, mass) in Msun, then the halo's
    stars), tests = the halo's stars; returns the bool mask bound (E <= 0)."""
    from . import halo_gas

    p = np.asarray(part_list, dtype=np.int64)
    sx, sy, sz = (np.asarray(a)[p] for a in (st_x, st_y, st_z))
    svx, svy, svz = (np.asarray(a)[p] for a in (st_vx, st_vy, st_vz))
    sm = np.asarray(st_mass)[p]
    if extra_sources is not None:
        ex, ey, ez, em = (np.asarray(a) for a in extra_sources)
        tx = np.concatenate((ex, sx))
        ty = np.concatenate((ey, sy))
        tz = np.concatenate((ez, sz))
        tm = np.concatenate((em, sm))
    else:
        tx, ty, tz, tm = sx, sy, sz, sm
    be = halo_gas.brute_force_binding_energy_fortran(tm, tx, ty, tz, sx, sy, sz)
    be = -be
    be *= G_CONST
    be *= factor_v ** 2
    kin = 0.5 * ((svx - vx) ** 2 + (svy - vy) ** 2 + (svz - vz) ** 2)
    return kin + be <= 0.0


def escape_velocity_unbinding_fortran(rete, L, ncoarse, grid_data, gas_data, masclet_dm_data,
                                      cx, cy, cz, vx, vy, vz, Rmax, part_list, st_x, st_y,
                                      st_z, st_vx, st_vy, st_vz, st_mass, factor_v, rho_B,
                                      gas_particles=None, dm_mass_to_sun=None):
    """Reference signature and steps (halo_properties.py:282-361): the dark matter of
    masclet_dm_data within Rmax (masses x masclet's mass unit, ``gathering.MASS_TO_SUN``
    unless ``dm_mass_to_sun`` is given), the gas particles of the AMR cells within Rmax
    (``halo_gas.AMRgrid_to_particles`` on the GPU, masses x rete**3) and the halo's stars
    are the sources, in the reference's order gas, stars, dark matter; the stars are the
    tests.  Returns the bool mask bound (E <= 0) over part_list.  A caller without the
    MASCLET grid arrays may pass ``gas_particles`` = (x, y, z, mass in Msun, comoving)
    instead; the Rmax cut and the rete**3 scaling are applied to them as to the AMR
    particles."""
    from . import halo_gas
    from .gathering import MASS_TO_SUN

    unit = MASS_TO_SUN if dm_mass_to_sun is None else dm_mass_to_sun
    # dark matter (halo_properties.py:288-300)
    dm_x = masclet_dm_data[0]
    dm_y = masclet_dm_data[1]
    dm_z = masclet_dm_data[2]
    dm_mass = masclet_dm_data[3] * unit
    inside = np.sqrt((dm_x - cx) ** 2 + (dm_y - cy) ** 2 + (dm_z - cz) ** 2) < Rmax
    dm_x, dm_y, dm_z, dm_mass = dm_x[inside], dm_y[inside], dm_z[inside], dm_mass[inside]
    # the halo's stars (:303-309)
    p = np.asarray(part_list, dtype=np.int64)
    st_x_halo, st_y_halo, st_z_halo = (np.asarray(a)[p] for a in (st_x, st_y, st_z))
    st_vx_halo, st_vy_halo, st_vz_halo = (np.asarray(a)[p] for a in (st_vx, st_vy, st_vz))
    st_mass_halo = np.asarray(st_mass)[p]
    # gas particles of the AMR cells (:313-328)
    if gas_particles is None:
        g = halo_gas.AMRgrid_to_particles(L, ncoarse, grid_data, gas_data, Rmax, cx, cy, cz,
                                          rho_B)
        gas_x, gas_y, gas_z, gas_mass = g[0], g[1], g[2], g[6]
    else:
        gas_x, gas_y, gas_z, gas_mass = (np.asarray(a) for a in gas_particles)
    gin = np.sqrt((gas_x - cx) ** 2 + (gas_y - cy) ** 2 + (gas_z - cz) ** 2) < Rmax
    gas_x, gas_y, gas_z = gas_x[gin], gas_y[gin], gas_z[gin]
    gas_mass = gas_mass[gin]
    gas_mass *= rete ** 3
    # sources gas, stars, dark matter; tests the stars (:335-340)
    total_x = np.concatenate((gas_x, st_x_halo, dm_x))
    total_y = np.concatenate((gas_y, st_y_halo, dm_y))
    total_z = np.concatenate((gas_z, st_z_halo, dm_z))
    total_mass = np.concatenate((gas_mass, st_mass_halo, dm_mass))
    be = halo_gas.brute_force_binding_energy_fortran(total_mass, total_x, total_y, total_z,
                                                     st_x_halo, st_y_halo, st_z_halo)
    be = -be
    be *= G_CONST
    be *= factor_v ** 2
    kin = 0.5 * ((st_vx_halo - vx) ** 2 + (st_vy_halo - vy) ** 2 + (st_vz_halo - vz) ** 2)
    return kin + be <= 0.0


def _arange_grid(grid):
    """(start, second node, delta, length) of a grid made by np.arange (node k = start,
    the second node, start + k * delta); ValueError for any other grid."""
    grid = np.asarray(grid, dtype=np.float64).ravel()
    n = grid.size
    if n == 0:
        raise ValueError("attempt to get argmin of an empty sequence")
    start = float(grid[0])
    g1 = float(grid[1]) if n > 1 else start + 1.0
    d = g1 - start
    nodes = start + np.arange(n, dtype=np.float64) * d
    nodes[0] = start
    if n > 1:
        nodes[1] = g1
    if not (d > 0 and np.array_equal(nodes, grid)):
        raise ValueError("grid must be an increasing np.arange grid")
    return start, g1, d, n


def calc_cell(cx, cy, cz, grid, part_list, st_x, st_y, st_z, st_vx, st_vy, st_vz, st_mass,
              vcm_cell, mass_cell, quantas_cell, sig3D_cell):
    """halo_properties.py:113-159.  The cell arrays are filled for the occupied cells (the
    reference adds into them; every caller passes them zero-filled) and returned with the
    per-member node indices (int32).  The grid must be an np.arange grid, as the
    reference's callers build it."""
    start, g1, d, n = _arange_grid(grid)
    groups, sub = _one_group(part_list, [st_x, st_y, st_z, st_vx, st_vy, st_vz, st_mass])
    npart = groups.n_members
    ijk, st = engine.cell_statistics(*sub, groups.members[:npart], (cx, cy, cz), start, g1, d,
                                     n)
    ijk = ijk.cpu().numpy()
    st = st.cpu().numpy()
    if npart:
        keys = np.unique((ijk[:, 0].astype(np.int64) * n + ijk[:, 1]) * n + ijk[:, 2])
        ix, rem = np.divmod(keys, n * n)
        iy, iz = np.divmod(rem, n)
        vcm_cell[ix, iy, iz, :] = st[:, 0:3]
        mass_cell[ix, iy, iz] = st[:, 3]
        quantas_cell[ix, iy, iz] = st[:, 4]
        sig3D_cell[ix, iy, iz] = st[:, 5]
    wx, wy, wz = (np.ascontiguousarray(ijk[:, a]) for a in range(3))
    return vcm_cell, mass_cell, quantas_cell, sig3D_cell, wx, wy, wz


def clean_cell(cx, cy, cz, M, RRHH, grid, part_list, st_x, st_y, st_z, st_vx, st_vy, st_vz,
               st_mass, vcm_cell, mass_cell, quantas_cell, sig3D_cell, ll, q_fil, sig_fil,
               which_cell_x, which_cell_y, which_cell_z):
    """halo_properties.py:162-278: (cx, cy, cz, M, R, control (int32), which_cell_x, _y,
    _z).  The first round's cells are recomputed on the GPU from (cx, cy, cz) and the
    grid, which must be np.arange(-(RRHH + ll), RRHH + ll, 2 ll) as the reference's
    driver builds it; the cell arrays and node indices passed in (calc_cell's output for
    that grid) are not read."""
    want = engine._grid(np.array([float(RRHH)]), float(ll))
    start, g1, d, n = _arange_grid(grid)
    if n != int(want[3][0]) or start != want[0][0] or (n > 1 and g1 != want[1][0]):
        raise ValueError("grid must be np.arange(-(RRHH + ll), RRHH + ll, 2 ll)")
    groups, sub = _one_group(part_list, [st_x, st_y, st_z, st_vx, st_vy, st_vz, st_mass])
    res = engine.phase_space_clean(groups, *sub, float(ll), float(q_fil), float(sig_fil),
                                   centres=[(float(cx), float(cy), float(cz))],
                                   radii=[float(RRHH)])
    control = res.control_clean.cpu().numpy().astype(np.int32)
    w = res.which_of(0)
    wx, wy, wz = (np.ascontiguousarray(w[:, a]) for a in range(3))
    ccx, ccy, ccz, cM, cR = (float(v) for v in res.clean[0])
    return ccx, ccy, ccz, cM, cR, control, wx, wy, wz
