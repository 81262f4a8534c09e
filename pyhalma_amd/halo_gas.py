"""Drop-in for the pair-potential part of ``python_scripts.halo_gas``.

Same names, arguments, return values and error behaviour as the reference functions;
the O(N^2) sums run on the GPU through ``pyhalma_amd.particle`` (libhalma).

  brute_force_binding_energy_fortran          halo_gas.py:164-187
  serial_brute_force_binding_energy_fortran   halo_gas.py:191-215
  parallel_inside                             halo_gas.py:218-220
  patch_to_particles, AMRgrid_to_particles    halo_gas.py:9-141
  st_gas_dm_particles_inside                  halo_gas.py:223-277
  RPS                                         halo_gas.py:285-492
  most_bound_particle                         halo_gas.py:498-634

RPS and most_bound_particle keep the reference's sequence of source groups and its
subsampling above BRUTE_FORCE_LIM (np.random.choice with replacement from the global
numpy RNG, drawn in the same order), so with a seeded ``np.random`` they draw the same
samples the reference would.  ``RPS_batched`` is the whole-catalogue form of RPS (every
halo in one call, every source summed, exact decisions).  The source gathering of the
same module (patch_to_particles, AMRgrid_to_particles, st_gas_dm_particles_inside,
halo_gas.py:9-141, :223-277) runs on the GPU through ``pyhalma_amd.gathering``, whose
``gather_sources`` does it for every halo at once.
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


_GAS_FIELDS = ("x", "y", "z", "vx", "vy", "vz", "mass", "temp")


def patch_to_particles(patch_res, patch_rx, patch_ry, patch_rz, patch_nx, patch_ny, patch_nz,
                       patch_delta, patch_cr0amr, patch_solapst, patch_vx, patch_vy, patch_vz,
                       patch_temp, box, rho_B):
    """halo_gas.py:9-52 on the GPU: the (x, y, z, vx, vy, vz, mass, temp) arrays of the
    patch's unrefined, non-overlapped cells strictly inside the box."""
    from .gathering import AmrGrid, amr_particles
    nx, ny, nz = int(patch_nx), int(patch_ny), int(patch_nz)
    cells = [[np.asarray(a)[:nx, :ny, :nz]] for a in (patch_delta, patch_vx, patch_vy,
                                                       patch_vz, patch_temp, patch_cr0amr,
                                                       patch_solapst)]
    # one level-1 patch whose cell side is patch_res: L / ncoarse = 2 patch_res
    grid = AmrGrid(2.0 * float(patch_res), 1, [1], [nx], [ny], [nz], [patch_rx], [patch_ry],
                   [patch_rz], cells)
    _, g = amr_particles(grid, rho_B, [np.asarray(box, np.float64)], vel_scale=1.0)
    return tuple(g[k].cpu().numpy() for k in _GAS_FIELDS)


def AMRgrid_to_particles(L, ncoarse, grid_data, gas_data, Rrps, cx, cy, cz, rho_B):
    """halo_gas.py:56-141 on the GPU (the snapshot is uploaded once per grid_data /
    gas_data pair, see ``pyhalma_amd.gathering``)."""
    from .gathering import amr_particles, cached_grid
    grid = cached_grid(L, ncoarse, grid_data, gas_data)
    box = np.array([cx - Rrps, cx + Rrps, cy - Rrps, cy + Rrps, cz - Rrps, cz + Rrps])
    _, g = amr_particles(grid, rho_B, [box])
    return tuple(g[k].cpu().numpy() for k in _GAS_FIELDS)


def st_gas_dm_particles_inside(rete, L, ncoarse, grid_data, gas_data, masclet_dm_data,
                               masclet_st_data, st_kdtree, dm_kdtree, cx, cy, cz, R, rho_B):
    """halo_gas.py:223-277 on the GPU.  The KD-trees are not used: the dark matter and
    stars within R come from a sphere query of masclet_dm_data / masclet_st_data (same
    set; indices in ascending order, where the KD-tree leaves the order open).  Masses
    in Msun with masclet's mass unit ``gathering.MASS_TO_SUN``."""
    from .gathering import cached_grid, gather_sources
    grid = cached_grid(L, ncoarse, grid_data, gas_data)
    dm = tuple(masclet_dm_data[k] for k in (0, 1, 2, 3))
    st = tuple(masclet_st_data[k] for k in (0, 1, 2, 6, 9))
    out = gather_sources(grid, rho_B, rete, dm, st, [[cx, cy, cz]], [R])
    gas = tuple(out["gas"][k].cpu().numpy() for k in _GAS_FIELDS)
    di = out["dm_index"].cpu().numpy()
    si = out["st_index"].cpu().numpy()
    dm_xyz = tuple(np.asarray(masclet_dm_data[k])[di] for k in (0, 1, 2))
    st_xyz = tuple(np.asarray(masclet_st_data[k])[si] for k in (0, 1, 2))
    return (gas + dm_xyz + (out["dm_mass"].cpu().numpy(),) + st_xyz
            + (out["st_mass"].cpu().numpy(), out["st_oripa"].cpu().numpy()))


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


def RPS_batched(gas, gas_offsets, halo_v, dm=None, stars=None, stats=None):
    """RPS (halo_gas.py:285-492) for every halo at once.  ``gas`` = (x, y, z, vx, vy, vz,
    mass, temp) in halo order with CSR host ``gas_offsets``; ``dm`` / ``stars`` =
    (x, y, z, mass, offsets) or None; ``halo_v`` (H, 3) the haloes' velocities.  Sources
    of halo h = its gas, then its DM, then its stars, all summed (no subsampling).
    Returns (H, 4) float64: total gas mass, cold bound fraction, cold unbound mass, hot
    unbound mass (cold iff T < 5e4 K), and the per-cell bound mask (device uint8).
    The gas self-term runs on the symmetric kernel (each cell pair once), the dark matter
    and stars against the cells on the chunked kernel (halma_energy_test_self_batched).
    ``stats`` (a dict) receives the energy test's counters (engine.energy_test_batched)."""
    import torch

    from . import engine
    from .device import to_dev
    gas_offsets = np.asarray(gas_offsets, np.int64)
    H = len(gas_offsets) - 1
    gx, gy, gz, gvx, gvy, gvz, gm, gt = gas
    # the gas cells are their own first sources (symmetric kernel); dark matter and stars
    # follow as extra sources
    extra = [g for g in (dm, stars) if g is not None]
    ext = dict(ext_off=None)
    if extra:
        ext_off, (ex, ey, ez, em) = _source_csr(extra, H)
        ext = dict(ext_off=ext_off, em=em, ex=ex, ey=ey, ez=ez)
    bound, _, _ = engine.energy_test_self_batched(gas_offsets, gm, gx, gy, gz, gvx, gvy, gvz,
                                                  halo_v, 2.0, stats=stats, **ext)
    mass = to_dev(gm, np.float64)
    temp = to_dev(gt, np.float64)
    dev = mass.device
    b = bound.to(torch.bool)
    cold = temp < 5 * 1e4
    zero = torch.zeros_like(mass)
    # the four per-halo sums with the library's group reduction over the CSR ranges
    # (double-double, fixed order, no atomics on the few hundred per-halo totals)
    n = int(mass.numel())
    grp = engine.Groups(torch.arange(max(n, 1), dtype=torch.int32, device=dev),
                        to_dev(gas_offsets, np.int64), gas_offsets)
    cx, cy, cz = engine._coords(gx, gy, gz)

    def seg(v):
        return engine.group_reduce(grp, cx, cy, cz, v)[:, 0]

    total = seg(mass)
    cold_bound = seg(torch.where(cold & b, mass, zero))
    unb_cold = seg(torch.where(cold & ~b, mass, zero))
    unb_hot = seg(torch.where(~cold & ~b, mass, zero))
    frac = torch.where(total != 0, cold_bound / torch.where(total != 0, total, 1.0),
                       torch.zeros_like(total))
    return torch.stack([total, frac, unb_cold, unb_hot], 1).cpu().numpy(), bound
