"""Source gathering for the gas path on the GPU (halo_gas.py:9-277, pyHALMA.py:1172-1185).

``AmrGrid`` holds a MASCLET AMR snapshot on the device (the level >= 1 patches, their
cells flattened in (ix, iy, iz) order); ``amr_particles`` turns the cells inside many
boxes (and spheres) into gas particles at once; ``gather_sources`` is
st_gas_dm_particles_inside (halo_gas.py:223-277) for every halo in one call: gas from the
AMR cells, dark matter and stars from batched sphere queries.

Two values come from masclet_framework, which is not part of this package: the level of
each patch (``levels_of``, the layout of masclet's tools.create_vector_levels: patch 0
is the level-0 base grid, then npatch[1] patches of level 1, npatch[2] of level 2, ...)
and the mass unit ``MASS_TO_SUN`` (masclet's units.mass_to_sun, 9.1717e18 Msun).  Both
can be given explicitly.
"""
import numpy as np
import torch

from . import _lib, engine
from .device import ptr, stream_ptr, to_dev, workspace

MASS_TO_SUN = 9.1717e18


def _native(a):
    """A device tensor of ``a`` in its own dtype (tensors stay where they are on the GPU)."""
    return to_dev(a if isinstance(a, torch.Tensor) else np.asarray(a), None)


def levels_of(npatch):
    """Level of every patch: [0] + [1] * npatch[1] + [2] * npatch[2] + ..."""
    npatch = np.asarray(npatch, np.int64)
    return np.concatenate([[0]] + [np.full(int(k), lev + 1, np.int32)
                                   for lev, k in enumerate(npatch[1:])]).astype(np.int32)


class AmrGrid:
    """The level >= 1 patches of an AMR snapshot, resident on the GPU.

    ``levels``, ``nx``, ``ny``, ``nz``, ``rx``, ``ry``, ``rz`` are per patch (all patches,
    level 0 included); ``cells`` = (delta, vx, vy, vz, temp, cr0amr, solapst), each a
    list of per-patch (nx, ny, nz) arrays indexed [ix, iy, iz].  Cell sides are
    (L / ncoarse) / 2**level and their cubes ``res**3``, computed here as the reference
    computes them."""

    def __init__(self, L, ncoarse, levels, nx, ny, nz, rx, ry, rz, cells):
        levels = np.asarray(levels, np.int32)
        keep = np.nonzero(levels >= 1)[0]
        self.patch_index = keep
        self.L, self.ncoarse = float(L), int(ncoarse)
        top = int(levels.max()) + 1 if levels.size else 1
        base = self.L / self.ncoarse
        self.res = np.array([base / 2 ** lev for lev in range(top)], np.float64)
        self.res3 = np.array([r ** 3 for r in self.res], np.float64)
        shape = np.stack([np.asarray(a, np.int64)[keep] for a in (nx, ny, nz)], 1) \
            if keep.size else np.zeros((0, 3), np.int64)
        sizes = shape.prod(1) if keep.size else np.zeros(0, np.int64)
        off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
        self.n_patches = int(keep.size)
        self.n_cells = int(off[-1])
        self.level = to_dev(levels[keep], np.int32)
        self.nx, self.ny, self.nz = (to_dev(shape[:, d], np.int32) for d in range(3))
        self.rx, self.ry, self.rz = (to_dev(np.asarray(a, np.float64)[keep], np.float64)
                                     for a in (rx, ry, rz))
        self.cell_offsets = to_dev(off, np.int64)
        delta, vx, vy, vz, temp, cr0amr, solapst = cells
        floats = [delta, vx, vy, vz, temp]
        dt = np.float32 if all(np.asarray(f[p]).dtype == np.float32 for f in floats
                               for p in keep) else np.float64
        self.dtype = dt

        def flat(lst, dtype):
            if not keep.size:
                return to_dev(np.zeros(1, dtype), dtype)
            return to_dev(np.concatenate([np.ascontiguousarray(np.asarray(lst[p]), dtype).ravel()
                                          for p in keep]), dtype)

        self.delta, self.vx, self.vy, self.vz, self.temp = (flat(f, dt) for f in floats)
        self.cr0amr = flat([np.asarray(a) != 0 for a in cr0amr], np.uint8)
        self.solapst = flat([np.asarray(a) != 0 for a in solapst], np.uint8)

    @classmethod
    def from_masclet(cls, L, ncoarse, grid_data, gas_data, levels=None):
        """From the reader outputs AMRgrid_to_particles takes (halo_gas.py:69-86):
        grid_data[5] npatch, [6:9] patchnx/ny/nz, [12:15] patchrx/ry/rz; gas_data[0:7]
        delta, vx, vy, vz, temp, cr0amr, solapst per patch."""
        if levels is None:
            levels = levels_of(grid_data[5])
        return cls(L, ncoarse, levels, grid_data[6], grid_data[7], grid_data[8],
                   grid_data[12], grid_data[13], grid_data[14], gas_data[:7])


def amr_particles(grid, rho_B, boxes, centres=None, radii=None, mass_scale=1.0,
                  vel_scale=3e5):
    """Gas particles of many queries (halma_amr_particles): cells strictly inside each
    box (6 per query), and within the sphere when ``centres`` is given.  Returns
    (offsets (numpy int64, Q+1), dict x, y, z, vx, vy, vz, mass, temp of FP64 device
    tensors)."""
    L = _lib.load()
    boxes_d = to_dev(np.asarray(boxes, np.float64).reshape(-1, 6), np.float64)
    Q = int(boxes_d.shape[0])
    cen = rad = None
    if centres is not None:
        cen = to_dev(np.asarray(centres, np.float64).reshape(Q, 3), np.float64)
        rad = to_dev(np.asarray(radii, np.float64).reshape(Q), np.float64)
    dev = boxes_d.device
    off = torch.zeros(Q + 1, dtype=torch.int64, device=dev)
    found = engine.ctypes_i64()
    gflag = _lib.F32 if grid.dtype == np.float32 else _lib.F64
    res = np.ascontiguousarray(grid.res)
    res3 = np.ascontiguousarray(grid.res3)
    max_found = 1 << 16
    ws = workspace(L.halma_amr_particles_workspace_size(Q))
    for _ in range(2):
        out = torch.empty(8 * max_found, dtype=torch.float64, device=dev)
        rc = L.halma_amr_particles(grid.n_patches, ptr(grid.level), ptr(grid.nx), ptr(grid.ny),
                                   ptr(grid.nz), ptr(grid.rx), ptr(grid.ry), ptr(grid.rz),
                                   ptr(grid.cell_offsets), ptr(grid.delta), ptr(grid.vx),
                                   ptr(grid.vy), ptr(grid.vz), ptr(grid.temp), gflag,
                                   ptr(grid.cr0amr), ptr(grid.solapst), len(res),
                                   res.ctypes.data, res3.ctypes.data, float(rho_B),
                                   float(mass_scale), float(vel_scale), Q, ptr(boxes_d),
                                   ptr(cen), ptr(rad), max_found, ptr(off), ptr(out), found,
                                   ptr(ws), ws.numel(), stream_ptr())
        if rc == _lib.E_LIMIT and found.value > max_found:
            max_found = int(found.value)
            continue
        _lib.check(rc, "halma_amr_particles")
        break
    F = int(found.value)
    names = ("x", "y", "z", "vx", "vy", "vz", "mass", "temp")
    return off.cpu().numpy(), {k: out[i * max_found: i * max_found + F]
                               for i, k in enumerate(names)}


def gather_sources(grid, rho_B, rete, dm, st, centres, radii, mass_to_sun=MASS_TO_SUN,
                   cell=0.0):
    """st_gas_dm_particles_inside (halo_gas.py:223-277) for every halo at once.

    ``dm`` = (x, y, z, mass) and ``st`` = (x, y, z, mass, oripa) as masclet reads them
    (masses in code units); ``centres`` (H, 3), ``radii`` (H,).  Returns a dict:
    ``gas_offsets`` and ``gas`` (the AMRgrid_to_particles output inside each sphere, mass
    x rete^3), ``dm_offsets`` / ``dm_index`` and ``st_offsets`` / ``st_index``
    (ascending indices of the particles within R), ``dm_mass`` / ``st_mass`` (their masses
    in Msun, in index order) and ``st_oripa``."""
    c = np.asarray(centres, np.float64).reshape(-1, 3)
    r = np.asarray(radii, np.float64).reshape(-1)
    boxes = np.stack([c[:, 0] - r, c[:, 0] + r, c[:, 1] - r, c[:, 1] + r, c[:, 2] - r,
                      c[:, 2] + r], 1)
    goff, gas = amr_particles(grid, rho_B, boxes, c, r, mass_scale=float(rete) ** 3)
    out = {"gas_offsets": goff, "gas": gas}
    for name, src in (("dm", dm), ("st", st)):
        off, idx, _ = engine.ball_query(src[0], src[1], src[2], c, r, cell=cell)
        out[f"{name}_offsets"] = off
        out[f"{name}_index"] = idx
        mass = _native(src[3])  # scaled in its own precision, as numpy scales it
        out[f"{name}_mass"] = mass[idx.long()] * mass_to_sun
    out["st_oripa"] = _native(st[4])[out["st_index"].long()]
    return out


def dark_matter_mass(dm_x, dm_y, dm_z, dm_mass, centres, radii, mass_to_sun=MASS_TO_SUN,
                     cell=0.0):
    """The dark-matter mass within radii[h] of centres[h] for every halo
    (pyHALMA.py:1172-1185): the masses are scaled to Msun in their own precision, as
    numpy scales them, and summed in double-double over the particles within R
    (distance <= R, the KD-tree's query_ball_point).  numpy's np.sum of the float32
    values rounds each pairwise partial sum to float32 instead: the two agree to the
    float32 summation error (parity to that tolerance)."""
    m = _native(dm_mass) * mass_to_sun
    _, _, sums = engine.ball_query(dm_x, dm_y, dm_z, centres, radii, m=m, cell=cell,
                                   count_only=True)
    return sums


_cache = {}


def cached_grid(L, ncoarse, grid_data, gas_data):
    """The AmrGrid of (grid_data, gas_data), uploaded once and kept for later calls with
    the same objects (call ``clear_cache`` after changing them in place)."""
    key = (id(grid_data), id(gas_data), float(L), int(ncoarse))
    hit = _cache.get("grid")
    if hit is None or hit[0] != key:
        grid = AmrGrid.from_masclet(L, ncoarse, grid_data, gas_data)
        _cache["grid"] = (key, grid, grid_data, gas_data)  # holds the objects: ids stay theirs
        return grid
    return hit[1]


def clear_cache():
    _cache.clear()
