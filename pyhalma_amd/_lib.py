"""ctypes binding of libhalma.so (the C ABI declared in include/halma.h).

torch is imported first so that the HIP runtime already loaded by torch
(libamdhip64.so.7) is the one libhalma binds to: device pointers and streams then
come straight from torch tensors and ``torch.cuda.current_stream()``.

There is no fallback: if the library is missing or a GPU call fails, an exception is
raised.
"""
import ctypes
import os

import numpy as np
import torch  # noqa: F401  (loads the HIP runtime libhalma binds to)

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libhalma.so")

_c = ctypes
_p = _c.c_void_p
_i64 = _c.c_int64
_i32 = _c.c_int32
_dbl = _c.c_double
_flt = _c.c_float
_size = _c.c_size_t
_i64p = _c.POINTER(_c.c_int64)

# name: (restype, [argtypes])
SIGNATURES = {
    "halma_last_error": (_c.c_char_p, []),
    "halma_version": (_c.c_int, []),
    "halma_device_cu_count": (_c.c_int, []),
    "halma_pot_sum_workspace_size": (_size, [_i64, _i64]),
    "halma_soft_workspace_size": (_size, [_i64, _i64p, _i64p]),
    "halma_pot_sum_self_workspace_size": (_size, [_i64, _i64p]),
    "halma_pot_sum_self_batched": (_c.c_int, [_i64, _i64p, _p, _p, _p, _p, _p, _p, _size, _p]),
    "halma_pot_sym_error_bound": (_dbl, []),
    "halma_energy_test_workspace_size": (_size, [_i64, _i64p, _i64p]),
    "halma_energy_test_batched": (_c.c_int, [_i64, _i64p, _p, _p, _p, _p, _i64p, _p, _p, _p,
                                             _p, _p, _p, _c.c_int, _p, _dbl, _p, _p, _i64p,
                                             _p, _size, _p]),
    "halma_energy_test_self_workspace_size": (_size, [_i64, _i64p, _i64p]),
    "halma_energy_test_self_batched": (_c.c_int, [_i64, _i64p, _p, _p, _p, _p, _i64p, _p, _p,
                                                  _p, _p, _p, _p, _p, _c.c_int, _p, _dbl, _p,
                                                  _p, _i64p, _p, _size, _p]),
    "halma_pot_soft_batched": (_c.c_int, [_i64, _i64p, _p, _p, _p, _p, _i64p, _p, _p, _p,
                                          _dbl, _p, _p, _p, _p, _size, _p]),
    "halma_pot_soft_batched_packed": (_c.c_int, [_i64, _i64p, _p, _i64p, _p, _p, _p, _dbl, _p,
                                                 _p, _p, _p, _size, _p]),
    "halma_soft_self_workspace_size": (_size, [_i64, _i64p]),
    "halma_pot_soft_self_batched": (_c.c_int, [_i64, _i64p, _p, _dbl, _p, _p, _p, _p, _size,
                                               _p]),
    "halma_pot_sum_f32": (_c.c_int, [_p, _p, _p, _p, _i64, _p, _p, _p, _i64, _p, _p, _p,
                                     _size, _p]),
    "halma_pot_error_bound": (_dbl, []),
    "halma_pot_sum_noalloc_max_bytes": (_size, []),
    "halma_pot_sum_f32_noalloc": (_c.c_int, [_p, _p, _p, _p, _i64, _p, _p, _p, _i64, _p, _p,
                                             _p]),
    "halma_pot_sum_batched_workspace_size": (_size, [_i64, _i64p, _i64p]),
    "halma_pot_sum_batched_f32": (_c.c_int, [_i64, _i64p, _p, _p, _p, _p, _i64p, _p, _p, _p,
                                             _p, _p, _size, _p]),
    "halma_pot_sum_f64_ordered": (_c.c_int, [_p, _p, _p, _p, _i64, _p, _p, _p, _i64, _p, _p]),
    "halma_pot_sum_f64_plain": (_c.c_int, [_p, _p, _p, _p, _i64, _p, _p, _p, _i64, _p, _p]),
    "halma_fof_workspace_size": (_size, [_i64, _c.c_int]),
    "halma_fof_labels": (_c.c_int, [_p, _p, _p, _c.c_int, _i64, _dbl, _p, _p, _size, _p]),
    "halma_groups_workspace_size": (_size, [_i64]),
    "halma_groups_from_labels": (_c.c_int, [_p, _i64, _i32, _p, _p, _i64p, _i64p, _p, _size,
                                            _p]),
    "halma_group_reduce": (_c.c_int, [_i64, _p, _p, _p, _p, _p, _c.c_int, _p, _p, _p, _c.c_int,
                                      _p, _c.c_int, _p, _p, _p]),
    "halma_gather_workspace_size": (_size, [_i64, _c.c_int]),
    "halma_gather_members": (_c.c_int, [_p, _i64, _i64, _p, _p, _p, _c.c_int, _p, _p, _p,
                                        _c.c_int, _p, _c.c_int, _c.c_int, _p, _p, _p, _p, _p,
                                        _p, _p, _p, _size, _p]),
    "halma_density_peak_workspace_size": (_size, [_i64]),
    "halma_density_peak": (_c.c_int, [_i64, _p, _p, _p, _p, _p, _c.c_int, _p, _c.c_int, _dbl,
                                      _p, _p, _p, _size, _p]),
    "halma_half_mass_radius_workspace_size": (_size, [_i64]),
    "halma_half_mass_radius": (_c.c_int, [_i64, _i64, _p, _p, _p, _p, _p, _c.c_int, _p,
                                          _c.c_int, _p, _p, _p, _size, _p]),
    "halma_halo_shape": (_c.c_int, [_i64, _p, _p, _p, _p, _p, _p, _p]),
    "halma_sigma_projections_workspace_size": (_size, [_i64, _c.c_int]),
    "halma_sigma_projections": (_c.c_int, [_i64, _p, _p, _p, _p, _p, _p, _p, _p, _c.c_int,
                                           _flt, _flt, _flt, _flt, _flt, _flt, _flt, _p, _p,
                                           _size, _p]),
    "halma_ball_query_workspace_size": (_size, [_i64, _i64, _i64]),
    "halma_ball_query": (_c.c_int, [_i64, _p, _p, _p, _c.c_int, _p, _c.c_int, _i64, _p, _p,
                                    _c.c_int, _dbl, _c.c_int, _i64, _p, _p, _p, _i64p, _p,
                                    _size, _p]),
    "halma_amr_particles_workspace_size": (_size, [_i64]),
    "halma_amr_particles": (_c.c_int, [_i64, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p,
                                       _c.c_int, _p, _p, _c.c_int, _p, _p, _dbl, _dbl, _dbl,
                                       _i64, _p, _p, _p, _i64, _p, _p, _i64p, _p, _size, _p]),
    "halma_progenitors_workspace_size": (_size, [_i64, _i64, _i64, _i64]),
    "halma_progenitors": (_c.c_int, [_i64, _i64, _p, _p, _p, _i64, _i64, _p, _p, _p, _p, _p,
                                     _p, _p, _p, _size, _p]),
    "halma_clean_select_workspace_size": (_size, [_i64]),
    "halma_clean_select": (_c.c_int, [_i64, _i64, _p, _p, _p, _p, _p, _p, _p, _p, _p, _i64p,
                                      _p, _size, _p]),
    "halma_clean_workspace_size": (_size, [_i64]),
    "halma_clean_cells": (_c.c_int, [_i64, _i64, _p, _p, _p, _p, _p, _c.c_int, _p, _p, _p,
                                     _c.c_int, _p, _c.c_int, _p, _p, _p, _p, _p, _c.c_int,
                                     _p, _p, _i64p, _p, _size, _p]),
    "halma_clean_decide": (_c.c_int, [_i64, _p, _p, _p, _p, _p, _p, _c.c_int, _p, _p, _p,
                                      _dbl, _dbl, _p, _p, _p, _size, _p]),
    "halma_unbind_workspace_size": (_size, [_i64, _i64p]),
    "halma_unbind_batched": (_c.c_int, [_i64, _i64p, _p, _p, _p, _p, _c.c_int, _p, _p, _p,
                                        _c.c_int, _p, _c.c_int, _dbl, _c.c_int, _c.c_int, _p,
                                        _p, _p, _p, _i64p, _p, _size, _p]),
    "halma_unbind_plain_workspace_size": (_size, [_i64, _i64p]),
    "halma_unbind_plain_f64": (_c.c_int, [_i64, _i64p, _p, _p, _p, _p, _c.c_int, _p, _p, _p,
                                          _c.c_int, _p, _c.c_int, _dbl, _c.c_int, _p,
                                          _c.POINTER(_c.c_int32), _p, _size, _p]),
    "halma_tune_set": (_c.c_int, [_c.c_int, _c.c_int]),
    "halma_tune_get": (_c.c_int, [_c.c_int]),
}

TUNE_TESTS_PER_LANE = 1
TUNE_SUBTILE = 2
TUNE_MIN_CHUNK = 3
TUNE_ITEMS_PER_WAVE = 4
TUNE_WAVES_PER_CU = 5
TUNE_SYM_MAX = 6

F32 = 0
F64 = 1
E_LIMIT = -3
GROUP_FIELDS = 16
UNBIND_GRAPHS = 1
UNBIND_TIME_POT = 2

_LIB = None


class HalmaError(RuntimeError):
    """A libhalma entry point returned a non-zero code."""


def load():
    """Load libhalma.so (raises if it has not been built)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"libhalma.so not found at {LIB_PATH}; build it with "
                "`python -m pyhalma_amd.build` (hipcc, gfx950)")
        lib = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = lib
    return _LIB


def check(rc, what=""):
    if rc != 0:
        msg = load().halma_last_error().decode(errors="replace")
        raise HalmaError(f"{what}: code {rc}: {msg}")


def i64_array(values):
    """A host int64 array for an ``int64_t*`` argument (no copy when ``values`` is
    already a contiguous int64 numpy array; the pointer keeps the array alive)."""
    arr = np.ascontiguousarray(values, dtype=np.int64)
    return arr.ctypes.data_as(_i64p)
