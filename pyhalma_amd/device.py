"""Device-memory plumbing: numpy <-> torch device tensors, workspaces, streams.

torch only provides memory and streams here; every computation is a libhalma call.
"""
import numpy as np
import torch

from . import _lib


def device():
    if not torch.cuda.is_available():
        raise RuntimeError("pyhalma_amd needs a ROCm GPU (torch.cuda.is_available() is False)")
    return torch.device("cuda", torch.cuda.current_device())


def stream_ptr(stream=None):
    s = stream if stream is not None else torch.cuda.current_stream()
    return int(s.cuda_stream)


def ptr(t):
    """Device pointer of a tensor (None -> NULL)."""
    if t is None:
        return None
    return int(t.data_ptr())


def to_dev(a, dtype=None):
    """numpy array (or tensor) -> contiguous device tensor of the given numpy dtype."""
    if isinstance(a, torch.Tensor):
        t = a
        if dtype is not None:
            t = t.to(dtype=_torch_dtype(dtype))
        return t.to(device()).contiguous()
    arr = np.asarray(a)
    if dtype is not None:
        arr = arr.astype(dtype, copy=False)
    arr = np.ascontiguousarray(arr)
    return torch.from_numpy(arr).to(device(), non_blocking=False)


def _torch_dtype(npdtype):
    return {np.dtype(np.float32): torch.float32, np.dtype(np.float64): torch.float64,
            np.dtype(np.int32): torch.int32, np.dtype(np.int64): torch.int64,
            np.dtype(np.uint8): torch.uint8}[np.dtype(npdtype)]


def empty(n, npdtype):
    return torch.empty(int(n), dtype=_torch_dtype(npdtype), device=device())


def workspace(nbytes):
    """Caller-owned workspace for a libhalma call (at least 256 bytes)."""
    return torch.empty(max(int(nbytes), 256), dtype=torch.uint8, device=device())


def dtype_flag(t):
    if t.dtype == torch.float32:
        return _lib.F32
    if t.dtype == torch.float64:
        return _lib.F64
    raise TypeError(f"unsupported dtype {t.dtype} (float32 or float64)")


def float_tensor(a):
    """Device tensor keeping float32/float64 as given; other dtypes become float64
    (numpy's promotion of integer coordinates)."""
    if isinstance(a, torch.Tensor):
        t = a.to(device())
    else:
        arr = np.asarray(a)
        if arr.dtype not in (np.float32, np.float64):
            arr = arr.astype(np.float64)
        t = torch.from_numpy(np.ascontiguousarray(arr)).to(device())
    if t.dtype not in (torch.float32, torch.float64):
        t = t.to(torch.float64)
    return t.contiguous()
