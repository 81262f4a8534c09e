"""Drop-in for the reduction and energy-test part of ``python_scripts.halo_properties``.

Same names, arguments and return values as the reference (scalars are Python floats):

  total_mass(part_list, st_mass)                                halo_properties.py:16-24
  center_of_mass(part_list, st_x, st_y, st_z, st_mass)          :26-43
  CM_velocity(M, part_list, st_vx, st_vy, st_vz, st_mass)       :45-60
  furthest_particle(cx, cy, cz, part_list, st_x, st_y, st_z)    :98-111
  density_peak(part_list, st_x, st_y, st_z, st_mass, ll)        :518-556
  escape_velocity_unbinding_fortran(...)                        :282-361

The reference's numba ``prange`` sums have no fixed order; here they are FP64
double-double sums (within about an ulp of the correctly rounded value).  density_peak
is bit-exact to the reference's serial loop (see libhalma's halma_density_peak).
Each call here handles one group; the batched forms over all groups at once are
``pyhalma_amd.engine.group_reduce`` / ``density_peak`` / ``unbind``.

The remaining functions of the reference module (phase-space cleaning, half-mass radii,
angular momentum, morphology, star formation, dispersions, shapes, Sersic fits, SMBH)
are halo-property calculations outside the accelerated path.
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
    """Reference signature (halo_properties.py:282-361).  The DM selection inside Rmax is
    done here; the AMR gas cells must be passed already converted to particles
    (``gas_particles`` = (x, y, z, mass in Msun, comoving) and the MASCLET mass unit as
    ``dm_mass_to_sun``), because the AMR reader (halo_gas.AMRgrid_to_particles) needs
    MASCLET grid data."""
    if gas_particles is None or dm_mass_to_sun is None:
        raise NotImplementedError(
            "pass gas_particles=(x, y, z, mass) and dm_mass_to_sun: converting MASCLET AMR "
            "cells to particles (halo_gas.py:56-141) is outside this package")
    dm_x, dm_y, dm_z = (np.asarray(a) for a in masclet_dm_data[:3])
    dm_m = np.asarray(masclet_dm_data[3]) * dm_mass_to_sun
    inside = np.sqrt((dm_x - cx) ** 2 + (dm_y - cy) ** 2 + (dm_z - cz) ** 2) < Rmax
    gx, gy, gz, gm = (np.asarray(a) for a in gas_particles)
    gin = np.sqrt((gx - cx) ** 2 + (gy - cy) ** 2 + (gz - cz) ** 2) < Rmax
    gm = gm[gin] * rete ** 3
    sources = (np.concatenate((gx[gin], dm_x[inside])), np.concatenate((gy[gin], dm_y[inside])),
               np.concatenate((gz[gin], dm_z[inside])), np.concatenate((gm, dm_m[inside])))
    # the reference orders the sources gas, stars, DM; the sum is order-insensitive up to
    # rounding, and the GPU sum is compensated
    return escape_velocity_unbinding(part_list, st_x, st_y, st_z, st_vx, st_vy, st_vz,
                                     st_mass, vx, vy, vz, factor_v, extra_sources=sources)
