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
(particle_subroutines.f90:160-461), halo-property routines outside the accelerated
path; calling them here raises NotImplementedError.
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
    def halo_shape(*args, **kwargs):
        raise NotImplementedError(
            "particle.halo_shape (particle_subroutines.f90:160-214) is a halo-property "
            "routine outside the accelerated FoF/reduction/unbinding path")

    @staticmethod
    def sigma_projections(*args, **kwargs):
        raise NotImplementedError(
            "particle.sigma_projections (particle_subroutines.f90:217-461) is a "
            "halo-property routine outside the accelerated FoF/reduction/unbinding path")


particle = _FortranModuleParticle()
