"""Drop-in for the pair-potential part of ``python_scripts.halo_gas``.

Same names, arguments, return values and error behaviour as the reference functions;
the O(N^2) sums run on the GPU through ``pyhalma_amd.particle`` (libhalma).

  brute_force_binding_energy_fortran          halo_gas.py:164-187
  serial_brute_force_binding_energy_fortran   halo_gas.py:191-215
  parallel_inside                             halo_gas.py:218-220
  RPS                                         halo_gas.py:285-492
  most_bound_particle                         halo_gas.py:498-634

RPS and most_bound_particle keep the reference's sequence of source groups and its
subsampling above BRUTE_FORCE_LIM (np.random.choice with replacement from the global
numpy RNG, drawn in the same order), so with a seeded ``np.random`` they draw the same
samples the reference would.  ``RPS_batched`` is the whole-catalogue form of RPS (every
halo in one call, every source summed, exact decisions).  The AMR source-gathering functions of the same module
(patch_to_particles, AMRgrid_to_particles, st_gas_dm_particles_inside) need MASCLET
grid data and are outside this package.
"""
import numpy as np

from .particle import particle

_g = 4.3 * 1e-3  # (km/s)^2 pc / Msun, halo_gas.py:459
_g *= 1e-6  # (km/s)^2 Mpc / Msun, halo_gas.py:460
G_CONST = _g


def _ncores():
    import os
    return np.int32(os.cpu_count() or 1)


def brute_force_binding_energy_fortran(total_mass, total_x, total_y, total_z, test_x,
                                       test_y, test_z):
    """Positive sum_j m_j/r_ij per test particle (float32); empty input -> np.array([])."""
    ntotal = np.int32(len(total_mass))
    ntest = np.int32(len(test_x))
    if ntest == 0:
        return np.array([])
    f32 = np.float32
    return particle.brute_force_binding_energy(
        _ncores(), ntotal, f32(total_mass), f32(total_x), f32(total_y), f32(total_z),
        ntest, f32(test_x), f32(test_y), f32(test_z))


def serial_brute_force_binding_energy_fortran(total_mass, total_x, total_y, total_z,
                                              test_x, test_y, test_z):
    ntotal = np.int32(len(total_mass))
    ntest = np.int32(len(test_x))
    if ntest == 0:
        return np.array([])
    f32 = np.float32
    return particle.serial_brute_force_binding_energy(
        ntotal, f32(total_mass), f32(total_x), f32(total_y), f32(total_z), ntest,
        f32(test_x), f32(test_y), f32(test_z))


def parallel_inside(array_x, array_y, array_z, R, cx, cy, cz):
    return np.sqrt((array_x - cx) ** 2 + (array_y - cy) ** 2 + (array_z - cz) ** 2) < R


def _source_sum(mass, x, y, z, tx, ty, tz, limit):
    """One source group's contribution: exact above nothing, subsampled (with
    replacement, rescaled by n/nsample) above ``limit``; None when the group is empty."""
    n = len(x)
    if n > limit:
        nsample = np.max([limit, int(0.01 * n)])
        pick = np.random.choice(np.arange(n), nsample, replace=True)
        s = brute_force_binding_energy_fortran(mass[pick], x[pick], y[pick], z[pick],
                                               tx, ty, tz)
        return s * n / nsample
    if n > 0:
        return brute_force_binding_energy_fortran(mass, x, y, z, tx, ty, tz)
    return None


def _dm_split(dm_x, dm_y, dm_z, dm_mass, mass_dm_part):
    """l <= 1 particles (mass >= 0.9 m_dm/8) are always summed exactly."""
    mand = dm_mass >= 0.9 * (mass_dm_part / 8)
    other = np.logical_not(mand)
    return ((dm_x[mand], dm_y[mand], dm_z[mand], dm_mass[mand]),
            (dm_x[other], dm_y[other], dm_z[other], dm_mass[other]))


def RPS(gas_x, gas_y, gas_z, gas_vx, gas_vy, gas_vz, gas_mass, gas_temp,
        dm_x, dm_y, dm_z, dm_mass, st_x, st_y, st_z, st_mass, vx, vy, vz, BRUTE_FORCE_LIM,
        mass_dm_part, num_dm_species):
    """Bound/unbound gas masses: (total gas mass, cold bound fraction, cold unbound mass,
    hot unbound mass); bound iff 1/2|v - v_h|^2 - 2 G S <= 0, cold iff T < 5e4 K."""
    ngas = len(gas_x)
    be = np.zeros((ngas,), dtype=np.float32)

    def add(part):
        nonlocal be
        if part is not None:
            be += part

    add(_source_sum(gas_mass, gas_x, gas_y, gas_z, gas_x, gas_y, gas_z, BRUTE_FORCE_LIM))
    if num_dm_species > 1:
        (mx, my, mz, mm), (ox, oy, oz, om) = _dm_split(dm_x, dm_y, dm_z, dm_mass,
                                                       mass_dm_part)
        if len(mx) > 0:
            add(brute_force_binding_energy_fortran(mm, mx, my, mz, gas_x, gas_y, gas_z))
        add(_source_sum(om, ox, oy, oz, gas_x, gas_y, gas_z, BRUTE_FORCE_LIM))
    else:
        add(_source_sum(dm_mass, dm_x, dm_y, dm_z, gas_x, gas_y, gas_z, BRUTE_FORCE_LIM))
    add(_source_sum(st_mass, st_x, st_y, st_z, gas_x, gas_y, gas_z, BRUTE_FORCE_LIM))

    be = -be
    be *= G_CONST
    be *= 2.0
    gas_v2 = 0.5 * ((gas_vx - vx) ** 2 + (gas_vy - vy) ** 2 + (gas_vz - vz) ** 2)
    total_energy = gas_v2 + be
    unbound = total_energy > 0.0
    bound = total_energy <= 0.0
    cold = gas_temp < 5 * 1e4
    hot = gas_temp >= 5 * 1e4
    total_gas_mass = np.sum(gas_mass)
    cold_bound = np.sum(gas_mass[cold * bound])
    frac = cold_bound / total_gas_mass if total_gas_mass != 0.0 else 0.0
    return (total_gas_mass, frac, np.sum(gas_mass[unbound * cold]),
            np.sum(gas_mass[unbound * hot]))


def most_bound_particle(gas_x, gas_y, gas_z, gas_mass, dm_x, dm_y, dm_z, dm_mass,
                        st_x, st_y, st_z, st_mass, st_oripa, BRUTE_FORCE_LIM, mass_dm_part):
    """Position and id of the star with the deepest potential (argmin of -S)."""
    nst = len(st_x)
    be = np.zeros((nst,), dtype=np.float32)

    def add(part):
        nonlocal be
        if part is not None:
            be += part

    add(_source_sum(gas_mass, gas_x, gas_y, gas_z, st_x, st_y, st_z, BRUTE_FORCE_LIM))
    (mx, my, mz, mm), (ox, oy, oz, om) = _dm_split(dm_x, dm_y, dm_z, dm_mass, mass_dm_part)
    if len(mx) > 0:
        add(brute_force_binding_energy_fortran(mm, mx, my, mz, st_x, st_y, st_z))
    add(_source_sum(om, ox, oy, oz, st_x, st_y, st_z, BRUTE_FORCE_LIM))
    add(_source_sum(st_mass, st_x, st_y, st_z, st_x, st_y, st_z, BRUTE_FORCE_LIM))
    be = -be
    k = np.argmin(be)
    return st_x[k], st_y[k], st_z[k], st_oripa[k]


def patch_to_particles(*args, **kwargs):
    raise NotImplementedError("AMR source gathering (halo_gas.py:9-52) needs MASCLET data")


def AMRgrid_to_particles(*args, **kwargs):
    raise NotImplementedError("AMR source gathering (halo_gas.py:56-141) needs MASCLET data")


def st_gas_dm_particles_inside(*args, **kwargs):
    raise NotImplementedError("source gathering (halo_gas.py:223-277) needs MASCLET data")


def _source_csr(groups, H):
    """Per-halo concatenation [group 0 | group 1 | ...] of CSR source groups (x, y, z,
    mass, offsets); returns the source offsets (host) and the packed device arrays."""
    import torch

    from .device import to_dev
    counts = np.zeros(H, np.int64)
    for g in groups:
        counts += np.diff(np.asarray(g[4], np.int64))
    off = np.zeros(H + 1, np.int64)
    off[1:] = np.cumsum(counts)
    n = int(off[-1])
    dev = to_dev(np.zeros(1, np.float32)).device
    out = [torch.empty(max(n, 1), dtype=torch.float32, device=dev) for _ in range(4)]
    base = torch.from_numpy(off[:-1].copy()).to(dev)
    filled = torch.zeros(H, dtype=torch.int64, device=dev)
    for g in groups:
        goff = np.asarray(g[4], np.int64)
        gcnt = torch.from_numpy(np.diff(goff)).to(dev)
        ng = int(goff[-1] - goff[0])
        if ng == 0:
            continue
        halo = torch.repeat_interleave(torch.arange(H, device=dev), gcnt)
        local = torch.arange(ng, device=dev) - torch.from_numpy(goff[:-1] - goff[0]).to(dev)[halo]
        dest = base[halo] + filled[halo] + local
        for k in range(4):
            col = g[k]
            if not isinstance(col, torch.Tensor):
                col = np.asarray(col)
            out[k][dest] = to_dev(col[goff[0]:goff[-1]], np.float32)
        filled += gcnt
    return off, out


def RPS_batched(gas, gas_offsets, halo_v, dm=None, stars=None):
    """RPS (halo_gas.py:285-492) for every halo at once.  ``gas`` = (x, y, z, vx, vy, vz,
    mass, temp) in halo order with CSR host ``gas_offsets``; ``dm`` / ``stars`` =
    (x, y, z, mass, offsets) or None; ``halo_v`` (H, 3) the haloes' velocities.  Sources
    of halo h = its gas, then its DM, then its stars, all summed (no subsampling).
    Returns (H, 4) float64: total gas mass, cold bound fraction, cold unbound mass, hot
    unbound mass (cold iff T < 5e4 K), and the per-cell bound mask (device uint8)."""
    import torch

    from . import engine
    from .device import to_dev
    gas_offsets = np.asarray(gas_offsets, np.int64)
    H = len(gas_offsets) - 1
    gx, gy, gz, gvx, gvy, gvz, gm, gt = gas
    groups = [(gx, gy, gz, gm, gas_offsets)] + [g for g in (dm, stars) if g is not None]
    src_off, (sx, sy, sz, sm) = _source_csr(groups, H)
    bound, _, _ = engine.energy_test_batched(src_off, sm, sx, sy, sz, gas_offsets, gx, gy, gz,
                                             gvx, gvy, gvz, halo_v, 2.0)
    mass = to_dev(gm, np.float64)
    temp = to_dev(gt, np.float64)
    dev = mass.device
    halo = torch.repeat_interleave(torch.arange(H, device=dev),
                                   torch.from_numpy(np.diff(gas_offsets)).to(dev))
    b = bound.to(torch.bool)
    cold = temp < 5 * 1e4

    def seg(v):
        return torch.zeros(H, dtype=torch.float64, device=dev).index_add(0, halo, v)

    total = seg(mass)
    cold_bound = seg(torch.where(cold & b, mass, torch.zeros_like(mass)))
    unb_cold = seg(torch.where(cold & ~b, mass, torch.zeros_like(mass)))
    unb_hot = seg(torch.where(~cold & ~b, mass, torch.zeros_like(mass)))
    frac = torch.where(total != 0, cold_bound / torch.where(total != 0, total, 1.0),
                       torch.zeros_like(total))
    return torch.stack([total, frac, unb_cold, unb_hot], 1).cpu().numpy(), bound
