"""Device-resident batched API over libhalma (torch tensors in, torch tensors out).

These functions are what the drop-in mirrors (fof, halo_properties, halo_gas, particle)
and bench.py call.  Inputs are torch device tensors (or numpy arrays, copied to the
device); every call runs on torch's current stream.
"""
import numpy as np
import torch

from . import _lib
from .device import ptr, stream_ptr, to_dev, workspace, empty


def _lib_():
    return _lib.load()


def pot_error_bound():
    """Rigorous relative error bound of the FP32 compensated pair sums."""
    return float(_lib_().halma_pot_error_bound())


def pot_sum(m, x, y, z, tx, ty, tz, out_dtype=np.float32, with_f64=False):
    """S_i = sum_j m_j / r_ij over sources with all three coordinates different from
    test i's (particle_subroutines.f90:466-514), FP32 compensated on the GPU.
    Returns a device tensor (float32 by default; float64 if out_dtype is float64).
    With with_f64=True returns (out, out_f64)."""
    L = _lib_()
    m, x, y, z = (to_dev(a, np.float32) for a in (m, x, y, z))
    tx, ty, tz = (to_dev(a, np.float32) for a in (tx, ty, tz))
    ns, nt = int(m.numel()), int(tx.numel())
    if not (x.numel() == y.numel() == z.numel() == ns):
        raise ValueError("source arrays must have equal lengths")
    if not (ty.numel() == tz.numel() == nt):
        raise ValueError("test arrays must have equal lengths")
    out32 = empty(nt, np.float32)
    out64 = empty(nt, np.float64) if (with_f64 or out_dtype == np.float64) else None
    ws_bytes = L.halma_pot_sum_workspace_size(ns, nt)
    ws = workspace(ws_bytes)
    rc = L.halma_pot_sum_f32(ptr(m), ptr(x), ptr(y), ptr(z), ns, ptr(tx), ptr(ty), ptr(tz),
                             nt, ptr(out32), ptr(out64), ptr(ws), ws.numel(), stream_ptr())
    _lib.check(rc, "halma_pot_sum_f32")
    if with_f64:
        return out32, out64
    return out64 if out_dtype == np.float64 else out32


def pot_sum_batched(src_off, m, x, y, z, tst_off, tx, ty, tz):
    """Batched pair sums over H haloes in CSR layout (host offsets).  FP64 output."""
    L = _lib_()
    src_off = np.asarray(src_off, np.int64)
    tst_off = np.asarray(tst_off, np.int64)
    H = len(src_off) - 1
    if len(tst_off) != H + 1:
        raise ValueError("offset arrays must have the same length")
    m, x, y, z = (to_dev(a, np.float32) for a in (m, x, y, z))
    tx, ty, tz = (to_dev(a, np.float32) for a in (tx, ty, tz))
    so = _lib.i64_array(src_off)
    to = _lib.i64_array(tst_off)
    out = torch.zeros(int(tx.numel()), dtype=torch.float64, device=tx.device)
    ws = workspace(L.halma_pot_sum_batched_workspace_size(H, so, to))
    rc = L.halma_pot_sum_batched_f32(H, so, ptr(m), ptr(x), ptr(y), ptr(z), to, ptr(tx),
                                     ptr(ty), ptr(tz), ptr(out), ptr(ws), ws.numel(),
                                     stream_ptr())
    _lib.check(rc, "halma_pot_sum_batched_f32")
    return out


def pot_sum_f64_ordered(m, x, y, z, tx, ty, tz):
    """FP64 pair sums in the oracle's fixed summation order (bitwise reproducible)."""
    L = _lib_()
    m, x, y, z = (to_dev(a, np.float32) for a in (m, x, y, z))
    tx, ty, tz = (to_dev(a, np.float32) for a in (tx, ty, tz))
    out = empty(int(tx.numel()), np.float64)
    rc = L.halma_pot_sum_f64_ordered(ptr(m), ptr(x), ptr(y), ptr(z), int(m.numel()), ptr(tx),
                                     ptr(ty), ptr(tz), int(tx.numel()), ptr(out), stream_ptr())
    _lib.check(rc, "halma_pot_sum_f64_ordered")
    return out


def pot_sum_f64_plain(m, x, y, z, tx, ty, tz):
    """Plain FP64 comparator kernel (one thread per test, sequential sum)."""
    L = _lib_()
    m, x, y, z = (to_dev(a, np.float64) for a in (m, x, y, z))
    tx, ty, tz = (to_dev(a, np.float64) for a in (tx, ty, tz))
    out = empty(int(tx.numel()), np.float64)
    rc = L.halma_pot_sum_f64_plain(ptr(m), ptr(x), ptr(y), ptr(z), int(m.numel()), ptr(tx),
                                   ptr(ty), ptr(tz), int(tx.numel()), ptr(out), stream_ptr())
    _lib.check(rc, "halma_pot_sum_f64_plain")
    return out


def _coords(x, y, z):
    """Device coordinate tensors sharing one float dtype (float32 kept, else float64)."""
    from .device import float_tensor
    x, y, z = float_tensor(x), float_tensor(y), float_tensor(z)
    if not (x.dtype == y.dtype == z.dtype):
        x, y, z = x.double(), y.double(), z.double()
    if not (x.numel() == y.numel() == z.numel()):
        raise ValueError("coordinate arrays must have equal lengths")
    return x, y, z


def fof_labels(x, y, z, linking_length):
    """Friends-of-friends labels (int32 device tensor): the minimum particle index of
    each particle's component (fof.py:18-79 semantics; singletons label themselves)."""
    from .device import dtype_flag
    L = _lib_()
    x, y, z = _coords(x, y, z)
    n = int(x.numel())
    labels = torch.empty(n, dtype=torch.int32, device=x.device)
    if n == 0:
        return labels
    flag = dtype_flag(x)
    ws = workspace(L.halma_fof_workspace_size(n, flag))
    rc = L.halma_fof_labels(ptr(x), ptr(y), ptr(z), flag, n, float(linking_length),
                            ptr(labels), ptr(ws), ws.numel(), stream_ptr())
    _lib.check(rc, "halma_fof_labels")
    return labels


def tune_set(key, value):
    _lib.check(_lib_().halma_tune_set(int(key), int(value)), "halma_tune_set")


def tune_get(key):
    return int(_lib_().halma_tune_get(int(key)))
