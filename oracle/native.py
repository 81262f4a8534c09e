"""ctypes bindings of the oracle's C restatement (``oracle_c.c``) and of the
reference's own FP32 pair-potential routine built from its Fortran source into
``oracle/_ref`` (``make -C oracle ref``).  Test infrastructure only."""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_REF = None

_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
_i64 = ctypes.c_int64


def lib_path():
    return os.path.join(_HERE, "liboracle.so")


def _lib():
    global _LIB
    if _LIB is None:
        path = lib_path()
        if not os.path.exists(path):
            raise RuntimeError(
                f"oracle library missing: {path} (build it with `make -C oracle`)")
        lib = ctypes.CDLL(path)
        lib.orc_pot_sum_f32.argtypes = [ctypes.c_int, _i64, _f32p, _f32p, _f32p, _f32p,
                                        _i64, _f32p, _f32p, _f32p, _f32p]
        lib.orc_pot_sum_f64.argtypes = [ctypes.c_int, _i64, _f32p, _f32p, _f32p, _f32p,
                                        _i64, _f32p, _f32p, _f32p, _f64p]
        lib.orc_pot_soft_f64.argtypes = [ctypes.c_int, _i64, _f32p, _f32p, _f32p, _f32p,
                                         _i64, _f64p, _f64p, _f64p, ctypes.c_double, _f64p]
        lib.orc_acc_soft_f64.argtypes = [ctypes.c_int, _i64, _f32p, _f32p, _f32p, _f32p,
                                         _i64, _f64p, _f64p, _f64p, ctypes.c_double,
                                         _f64p, _f64p, _f64p]
        lib.orc_fof_labels.argtypes = [_i64, _f64p, _f64p, _f64p, ctypes.c_double, _i64p]
        lib.orc_fof_labels.restype = ctypes.c_int
        lib.orc_fof_labels_cells.argtypes = [_i64, _f64p, _f64p, _f64p, ctypes.c_double,
                                             _i64p]
        lib.orc_fof_labels_cells.restype = ctypes.c_int
        _LIB = lib
    return _LIB


def _f32(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float32))


def _f64(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64))


def pot_sum_f32(m, x, y, z, tx, ty, tz, nthreads=0):
    """FP32 restatement of particle_subroutines.f90:466-514 (returns float32)."""
    m, x, y, z, tx, ty, tz = map(_f32, (m, x, y, z, tx, ty, tz))
    out = np.zeros(len(tx), np.float32)
    _lib().orc_pot_sum_f32(int(nthreads), len(m), m, x, y, z, len(tx), tx, ty, tz, out)
    return out


def pot_sum_f64(m, x, y, z, tx, ty, tz, nthreads=0):
    """FP64 direct sum of particle_subroutines.f90:466-514 on FP32-cast inputs
    (halo_gas.py:172-178), sequential in j, no contraction (returns float64)."""
    m, x, y, z, tx, ty, tz = map(_f32, (m, x, y, z, tx, ty, tz))
    out = np.zeros(len(tx), np.float64)
    _lib().orc_pot_sum_f64(int(nthreads), len(m), m, x, y, z, len(tx), tx, ty, tz, out)
    return out


def pot_soft_f64(m, x, y, z, tx, ty, tz, eps, nthreads=0):
    m, x, y, z = map(_f32, (m, x, y, z))
    tx, ty, tz = map(_f64, (tx, ty, tz))
    out = np.zeros(len(tx), np.float64)
    _lib().orc_pot_soft_f64(int(nthreads), len(m), m, x, y, z, len(tx), tx, ty, tz,
                            float(eps), out)
    return out


def acc_soft_f64(m, x, y, z, tx, ty, tz, eps, nthreads=0):
    """d/dx_test of sum_j m_j (r^2+eps^2)^-1/2, FP64; returns (n,3)."""
    m, x, y, z = map(_f32, (m, x, y, z))
    tx, ty, tz = map(_f64, (tx, ty, tz))
    gx = np.zeros(len(tx))
    gy = np.zeros(len(tx))
    gz = np.zeros(len(tx))
    _lib().orc_acc_soft_f64(int(nthreads), len(m), m, x, y, z, len(tx), tx, ty, tz,
                            float(eps), gx, gy, gz)
    return np.stack([gx, gy, gz], axis=1)


def fof_labels_c(x, y, z, ll):
    """C cell-list FoF; labels = min index of each component (singletons included)."""
    x, y, z = map(_f64, (x, y, z))
    lab = np.empty(len(x), np.int64)
    rc = _lib().orc_fof_labels(len(x), x, y, z, float(ll), lab)
    if rc != 0:
        raise RuntimeError("orc_fof_labels failed")
    return lab


def fof_labels_cells(x, y, z, ll):
    """The same labels through sub-cells of side ll/sqrt(3) (each one connected) joined by
    one passing pair per neighbouring cell pair: fast on dense clumps."""
    x, y, z = map(_f64, (x, y, z))
    lab = np.empty(len(x), np.int64)
    rc = _lib().orc_fof_labels_cells(len(x), x, y, z, float(ll), lab)
    if rc != 0:
        raise RuntimeError("orc_fof_labels_cells failed")
    return lab


# ----------------------------------------------------------------- reference ----
def _ref_path():
    return os.path.join(_HERE, "_ref", "libparticle_ref.so")


def have_ref():
    return os.path.exists(_ref_path())


def _ref():
    global _REF
    if _REF is None:
        path = _ref_path()
        if not os.path.exists(path):
            raise RuntimeError(f"reference routine not built: {path}")
        lib = ctypes.CDLL(path)
        fn = lib._QMparticlePbrute_force_binding_energy
        i32p = ctypes.POINTER(ctypes.c_int32)
        fn.argtypes = [i32p, i32p, _f32p, _f32p, _f32p, _f32p, i32p, _f32p, _f32p, _f32p,
                       _f32p]
        fn.restype = None
        _REF = fn
    return _REF


def ref_pot_sum_f32(m, x, y, z, tx, ty, tz, nthreads=1):
    """The reference's own brute_force_binding_energy (particle_subroutines.f90:466-514),
    compiled from its Fortran source with amdflang; Fortran passes every argument by
    reference, in the order of the SUBROUTINE statement."""
    m, x, y, z, tx, ty, tz = map(_f32, (m, x, y, z, tx, ty, tz))
    out = np.zeros(len(tx), np.float32)
    nc = ctypes.c_int32(int(nthreads))
    nt = ctypes.c_int32(len(m))
    ns = ctypes.c_int32(len(tx))
    _ref()(ctypes.byref(nc), ctypes.byref(nt), m, x, y, z, ctypes.byref(ns), tx, ty, tz, out)
    return out
