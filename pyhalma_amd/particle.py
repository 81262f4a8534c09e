"""Drop-in for the f2py extension ``fortran_modules.particle``.

The reference reaches its Fortran through ``particle.particle.<routine>``
(halo_gas.py:6,182,208; halo_properties.py:13); this module exposes the same attribute
path, argument order and return type for the pair-potential routines, computed by the
libhalma GPU kernels:

  particle.particle.brute_force_binding_energy(ncores, ntotal, total_mass, total_x,
      total_y, total_z, ntest, test_x, test_y, test_z) -> float32[ntest]
      (particle_subroutines.f90:466-514; ``ncores`` is accepted and ignored: the GPU
      decides its own parallelism)
  particle.particle.serial_brute_force_binding_energy(ntotal, total_mass, total_x,
      total_y, total_z, ntest, test_x, test_y, test_z) -> float32[ntest]
      (particle_subroutines.f90:517-556)

Like f2py, array arguments are converted to float32 and a length that does not match
the declared dimension raises ValueError.  The sum is FP32 with compensated
accumulation; its relative error is below ``pyhalma_amd.engine.pot_error_bound()``
(the reference's plain FP32 loop has no such bound).

The same Fortran module also holds ``halo_shape`` and ``sigma_projections``
(particle_subroutines.f90:160-461), which halo_properties.py:805,861 reach through the
same attribute path; both run on the GPU here (libhalma's halma_halo_shape and
halma_sigma_projections), with the reference's FP32 per-operation arithmetic and FP64
sums where the reference sums in FP32 through OpenMP reductions.
"""
import numpy as np

from . import engine


def _as_f32(a, n, name, routine):
    arr = np.asarray(a)
    if arr.ndim != 1:
        raise ValueError(f"{routine}: {name} must be rank-1, got rank {arr.ndim}")
    if arr.shape[0] != n:
        raise ValueError(f"{routine}: 0-th dimension of {name} must be fixed to {n} "
                         f"but got {arr.shape[0]}")
    return np.ascontiguousarray(arr, dtype=np.float32)


def _pot(routine, ntotal, total_mass, total_x, total_y, total_z, ntest, test_x, test_y,
         test_z):
    ntotal = int(ntotal)
    ntest = int(ntest)
    if ntotal < 0 or ntest < 0:
        raise ValueError(f"{routine}: negative dimension")
    m = _as_f32(total_mass, ntotal, "total_mass", routine)
    x = _as_f32(total_x, ntotal, "total_x", routine)
    y = _as_f32(total_y, ntotal, "total_y", routine)
    z = _as_f32(total_z, ntotal, "total_z", routine)
    tx = _as_f32(test_x, ntest, "test_x", routine)
    ty = _as_f32(test_y, ntest, "test_y", routine)
    tz = _as_f32(test_z, ntest, "test_z", routine)
    if ntest == 0:
        return np.zeros(0, np.float32)
    out = engine.pot_sum(m, x, y, z, tx, ty, tz)
    return out.cpu().numpy()


class _FortranModuleParticle:
    """The Fortran ``MODULE particle`` as seen through f2py."""

    @staticmethod
    def brute_force_binding_energy(ncores, ntotal, total_mass, total_x, total_y, total_z,
                                   ntest, test_x, test_y, test_z):
        del ncores  # OpenMP team size in the reference; the GPU path ignores it
        return _pot("brute_force_binding_energy", ntotal, total_mass, total_x, total_y,
                    total_z, ntest, test_x, test_y, test_z)

    @staticmethod
    def serial_brute_force_binding_energy(ntotal, total_mass, total_x, total_y, total_z,
                                          ntest, test_x, test_y, test_z):
        return _pot("serial_brute_force_binding_energy", ntotal, total_mass, total_x,
                    total_y, total_z, ntest, test_x, test_y, test_z)

    @staticmethod
    def halo_shape(ncore, npart, x, y, z, mass):
        """particle_subroutines.f90:160-214 -> float32[3]: square roots of the eigenvalues of
        the mass-weighted second-moment tensor of the (centred) points, largest first."""
        del ncore
        n = int(npart)
        arrs = [_as_f32(a, n, name, "halo_shape")
                for a, name in ((x, "x"), (y, "y"), (z, "z"), (mass, "mass"))]
        return engine.halo_shapes(np.array([0, n], np.int64), *arrs).cpu().numpy()[0]

    @staticmethod
    def sigma_projections(ncore, npart, grid, n_cell, part_list, st_x, st_y, st_z, st_vx,
                          st_vy, st_vz, st_mass, cx, cy, cz, R05x, R05y, R05z, ll):
        """particle_subroutines.f90:217-461 -> (SIG_1D_x_05, SIG_1D_y_05, SIG_1D_z_05,
        V_sigma, lambda); part_list is 1-based, as f2py passes it."""
        del ncore
        n = int(npart)
        nc = int(n_cell)
        g = _as_f32(grid, nc, "grid", "sigma_projections")
        pl = np.asarray(part_list).ravel()
        if pl.shape[0] != n:
            raise ValueError(f"sigma_projections: 0-th dimension of part_list must be fixed "
                             f"to {n} but got {pl.shape[0]}")
        idx = pl.astype(np.int64) - 1
        arrs = [np.ascontiguousarray(np.asarray(a, np.float32)[idx])
                for a in (st_x, st_y, st_z, st_vx, st_vy, st_vz, st_mass)]
        out = engine.sigma_projections(g, *arrs, cx, cy, cz, R05x, R05y, R05z, ll)
        return tuple(float(v) for v in out)


particle = _FortranModuleParticle()
