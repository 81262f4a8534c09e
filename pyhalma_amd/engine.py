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


def pot_sym_error_bound():
    """Rigorous relative error bound of the symmetric self-potential kernel."""
    return float(_lib_().halma_pot_sym_error_bound())


def pot_sum_self(offsets, m, x, y, z):
    """Self-potential of every halo (CSR host offsets; tests = sources = the members),
    each pair distance computed once for both terms.  FP64 output per member."""
    L = _lib_()
    offsets = np.asarray(offsets, np.int64)
    H = len(offsets) - 1
    m, x, y, z = (to_dev(a, np.float32) for a in (m, x, y, z))
    out = torch.zeros(int(x.numel()), dtype=torch.float64, device=x.device)
    if H <= 0:
        return out
    off = _lib.i64_array(offsets)
    ws = workspace(L.halma_pot_sum_self_workspace_size(H, off))
    rc = L.halma_pot_sum_self_batched(H, off, ptr(m), ptr(x), ptr(y), ptr(z), ptr(out), ptr(ws),
                                      ws.numel(), stream_ptr())
    _lib.check(rc, "halma_pot_sum_self_batched")
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


def pot_soft_batched(src_off, m, x, y, z, tst_off, tx, ty, tz, eps, pot=True, acc=False,
                     wsum=False, packed=None):
    """Softened sums per halo (CSR host offsets; see halma_pot_soft_batched):
    pot_i = sum_j m_j (r^2+eps^2)^-1/2, acc_i = d pot_i / d t_i (n_test x 3),
    wsum_i = sum_j m_j (r^2+eps^2)^-3/2.  Sources FP32, tests FP64.  ``packed``: the
    sources already as an (N, 4) contiguous FP32 device tensor of (x, y, z, m) rows
    (halma_pot_soft_batched_packed; m, x, y, z are then not read).  Returns a dict of
    float64 device tensors for the requested outputs."""
    L = _lib_()
    src_off = np.asarray(src_off, np.int64)
    tst_off = np.asarray(tst_off, np.int64)
    H = len(src_off) - 1
    if len(tst_off) != H + 1:
        raise ValueError("offset arrays must have the same length")
    if packed is not None:
        if not isinstance(packed, torch.Tensor) or packed.dim() != 2 or packed.shape[1] != 4:
            raise ValueError("packed sources must be an (N, 4) tensor of (x, y, z, m) rows")
        if int(packed.shape[0]) < int(src_off[-1]):  # rows src_off[0] .. src_off[-1] - 1
            raise ValueError("packed sources shorter than the source offsets")
        packed = to_dev(packed, np.float32)  # contiguous FP32 on the device
    else:
        m, x, y, z = (to_dev(a, np.float32) for a in (m, x, y, z))
    tx, ty, tz = (to_dev(a, np.float64) for a in (tx, ty, tz))
    nt = int(tx.numel())
    dev = tx.device
    out = {}
    if pot:
        out["pot"] = torch.zeros(nt, dtype=torch.float64, device=dev)
    if acc:
        out["acc"] = torch.zeros((nt, 3), dtype=torch.float64, device=dev)
    if wsum:
        out["wsum"] = torch.zeros(nt, dtype=torch.float64, device=dev)
    if H == 0 or nt == 0:
        return out
    so = _lib.i64_array(src_off)
    to = _lib.i64_array(tst_off)
    ws = workspace(L.halma_soft_workspace_size(H, so, to))
    if packed is not None:
        rc = L.halma_pot_soft_batched_packed(H, so, ptr(packed), to, ptr(tx), ptr(ty), ptr(tz),
                                             float(eps), ptr(out.get("pot")),
                                             ptr(out.get("acc")), ptr(out.get("wsum")),
                                             ptr(ws), ws.numel(), stream_ptr())
        _lib.check(rc, "halma_pot_soft_batched_packed")
        return out
    rc = L.halma_pot_soft_batched(H, so, ptr(m), ptr(x), ptr(y), ptr(z), to, ptr(tx), ptr(ty),
                                  ptr(tz), float(eps), ptr(out.get("pot")),
                                  ptr(out.get("acc")), ptr(out.get("wsum")), ptr(ws),
                                  ws.numel(), stream_ptr())
    _lib.check(rc, "halma_pot_soft_batched")
    return out


def pot_soft_self(offsets, packed, eps, pot=True, acc=False, wsum=False):
    """Softened sums of every member of every halo against the other members of its halo
    (halma_pot_soft_self_batched): the sources are the (N, 4) contiguous FP32 device tensor
    ``packed`` of (x, y, z, m) rows, haloes the CSR host ``offsets``; the tests are the
    same rows.  Returns a dict of float64 device tensors (acc: N x 3) for the requested
    outputs."""
    L = _lib_()
    offsets = np.asarray(offsets, np.int64)
    H = len(offsets) - 1
    if not isinstance(packed, torch.Tensor) or packed.dim() != 2 or packed.shape[1] != 4:
        raise ValueError("packed sources must be an (N, 4) tensor of (x, y, z, m) rows")
    if H < 0 or int(packed.shape[0]) < int(offsets[-1]):
        raise ValueError("packed sources shorter than the offsets")
    packed = to_dev(packed, np.float32)
    n = int(offsets[-1]) if H >= 0 else 0
    dev = packed.device
    out = {}
    if pot:
        out["pot"] = torch.zeros(n, dtype=torch.float64, device=dev)
    if acc:
        out["acc"] = torch.zeros((n, 3), dtype=torch.float64, device=dev)
    if wsum:
        out["wsum"] = torch.zeros(n, dtype=torch.float64, device=dev)
    if H <= 0 or n == int(offsets[0]):
        return out
    oo = _lib.i64_array(offsets)
    ws = workspace(L.halma_soft_self_workspace_size(H, oo))
    rc = L.halma_pot_soft_self_batched(H, oo, ptr(packed), float(eps), ptr(out.get("pot")),
                                       ptr(out.get("acc")), ptr(out.get("wsum")), ptr(ws),
                                       ws.numel(), stream_ptr())
    _lib.check(rc, "halma_pot_soft_self_batched")
    return out


def energy_test_batched(src_off, m, x, y, z, tst_off, tx, ty, tz, vx, vy, vz, halo_v,
                        factor, with_pot=False, stats=None):
    """One energy test per test particle, every halo at once (halma_energy_test_batched):
    bound iff 1/2|v - v_h|^2 - factor*G*S <= 0, S over the halo's sources (CSR host
    offsets).  Returns (bound uint8 tensor, S float64 tensor or None, tests re-tested in
    FP64); ``stats`` (a dict) receives retested, pair_ns (the pair kernel's time) and
    pairs."""
    import ctypes
    from .device import dtype_flag
    L = _lib_()
    src_off = np.asarray(src_off, np.int64)
    tst_off = np.asarray(tst_off, np.int64)
    H = len(src_off) - 1
    if len(tst_off) != H + 1:
        raise ValueError("offset arrays must have the same length")
    m, x, y, z = (to_dev(a, np.float32) for a in (m, x, y, z))
    tx, ty, tz = (to_dev(a, np.float32) for a in (tx, ty, tz))
    vx, vy, vz = _coords(vx, vy, vz)
    hv = to_dev(np.asarray(halo_v, np.float64).reshape(-1, 3) if not isinstance(halo_v, torch.Tensor)
                else halo_v.double().reshape(-1, 3), np.float64).contiguous()
    if hv.shape[0] != H:
        raise ValueError("halo_v must hold 3 values per halo")
    nt = int(tx.numel())
    bound = torch.zeros(max(nt, 1), dtype=torch.uint8, device=tx.device)
    S = torch.zeros(max(nt, 1), dtype=torch.float64, device=tx.device) if with_pot else None
    st = (ctypes.c_int64 * 3)()
    if H > 0 and nt > 0:
        so = _lib.i64_array(src_off)
        to = _lib.i64_array(tst_off)
        ws = workspace(L.halma_energy_test_workspace_size(H, so, to))
        rc = L.halma_energy_test_batched(H, so, ptr(m), ptr(x), ptr(y), ptr(z), to, ptr(tx),
                                         ptr(ty), ptr(tz), ptr(vx), ptr(vy), ptr(vz),
                                         dtype_flag(vx), ptr(hv), float(factor), ptr(bound),
                                         ptr(S), st, ptr(ws), ws.numel(), stream_ptr())
        _lib.check(rc, "halma_energy_test_batched")
    if stats is not None:
        stats.update(retested=int(st[0]), pair_ns=int(st[1]), pairs=int(st[2]))
    return bound[:nt], (S[:nt] if S is not None else None), int(st[0])


def energy_test_self_batched(tst_off, m, x, y, z, vx, vy, vz, halo_v, factor, ext_off=None,
                             em=None, ex=None, ey=None, ez=None, with_pot=False, stats=None):
    """The energy test with each halo's tests as its first sources
    (halma_energy_test_self_batched): halo h's tests (m, x, y, z, v over CSR host tst_off)
    are summed against each other on the symmetric kernel and against its extra sources
    (em, ex, ey, ez over CSR host ext_off; none when ext_off is None) on the chunked
    kernel.  Returns as energy_test_batched."""
    import ctypes
    from .device import dtype_flag
    L = _lib_()
    tst_off = np.asarray(tst_off, np.int64)
    H = len(tst_off) - 1
    if ext_off is None:
        ext_off = np.zeros(H + 1, np.int64)
        em = ex = ey = ez = np.zeros(1, np.float32)
    ext_off = np.asarray(ext_off, np.int64)
    if len(ext_off) != H + 1:
        raise ValueError("offset arrays must have the same length")
    m, x, y, z = (to_dev(a, np.float32) for a in (m, x, y, z))
    em, ex, ey, ez = (to_dev(a, np.float32) for a in (em, ex, ey, ez))
    vx, vy, vz = _coords(vx, vy, vz)
    hv = to_dev(np.asarray(halo_v, np.float64).reshape(-1, 3) if not isinstance(halo_v, torch.Tensor)
                else halo_v.double().reshape(-1, 3), np.float64).contiguous()
    if hv.shape[0] != H:
        raise ValueError("halo_v must hold 3 values per halo")
    nt = int(x.numel())
    bound = torch.zeros(max(nt, 1), dtype=torch.uint8, device=x.device)
    S = torch.zeros(max(nt, 1), dtype=torch.float64, device=x.device) if with_pot else None
    st = (ctypes.c_int64 * 3)()
    if H > 0 and nt > 0:
        to = _lib.i64_array(tst_off)
        eo = _lib.i64_array(ext_off)
        ws = workspace(L.halma_energy_test_self_workspace_size(H, to, eo))
        rc = L.halma_energy_test_self_batched(H, to, ptr(m), ptr(x), ptr(y), ptr(z), eo,
                                              ptr(em), ptr(ex), ptr(ey), ptr(ez), ptr(vx),
                                              ptr(vy), ptr(vz), dtype_flag(vx), ptr(hv),
                                              float(factor), ptr(bound), ptr(S), st, ptr(ws),
                                              ws.numel(), stream_ptr())
        _lib.check(rc, "halma_energy_test_self_batched")
    if stats is not None:
        stats.update(retested=int(st[0]), pair_ns=int(st[1]), pairs=int(st[2]))
    return bound[:nt], (S[:nt] if S is not None else None), int(st[0])


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


class Groups:
    """Groups in CSR form on the device: members[offsets[g]:offsets[g+1]] are the particle
    indices of group g (ascending), groups ordered by minimum member index."""

    def __init__(self, members, offsets, offsets_host):
        self.members = members            # int32 device tensor
        self.offsets = offsets            # int64 device tensor (G+1)
        self.offsets_host = offsets_host  # numpy int64 (G+1)

    @property
    def n_groups(self):
        return len(self.offsets_host) - 1

    @property
    def n_members(self):
        return int(self.offsets_host[-1]) if len(self.offsets_host) else 0

    def sizes(self):
        return np.diff(self.offsets_host)

    def to_lists(self):
        mem = self.members[: self.n_members].cpu().numpy()
        return [mem[a:b].tolist() for a, b in zip(self.offsets_host[:-1], self.offsets_host[1:])]

    def to_arrays(self):
        mem = self.members[: self.n_members].cpu().numpy().astype(np.int64)
        return [mem[a:b] for a, b in zip(self.offsets_host[:-1], self.offsets_host[1:])]


def groups_from_labels(labels, min_size=2):
    """Groups with at least min_size members from FoF labels (good_groups filter,
    pyHALMA.py:148-149)."""
    L = _lib_()
    labels = to_dev(labels, np.int32)
    n = int(labels.numel())
    members = torch.empty(max(n, 1), dtype=torch.int32, device=labels.device)
    # the call writes offsets[0..G] (G the group count it returns)
    offsets = torch.empty(n + 1, dtype=torch.int64, device=labels.device)
    ng = ctypes_i64()
    nm = ctypes_i64()
    if n > 0:
        ws = workspace(L.halma_groups_workspace_size(n))
        rc = L.halma_groups_from_labels(ptr(labels), n, int(min_size), ptr(members),
                                        ptr(offsets), ng, nm, ptr(ws), ws.numel(), stream_ptr())
        _lib.check(rc, "halma_groups_from_labels")
    G = int(ng.value)
    offsets = offsets[: G + 1]
    if G == 0:
        offsets = torch.zeros(1, dtype=torch.int64, device=labels.device)
    return Groups(members, offsets, offsets.cpu().numpy())


def fof_groups(x, y, z, linking_length, min_size=2):
    """friends-of-friends straight to groups (labels, Groups)."""
    lab = fof_labels(x, y, z, linking_length)
    return lab, groups_from_labels(lab, min_size)


def ctypes_i64(v=0):
    import ctypes
    return ctypes.c_int64(v)


def _vec(a):
    from .device import float_tensor
    return float_tensor(a)


class MemberData:
    """Member-ordered copies of the particle arrays of a Groups object, with `groups`
    rewritten to index them (members = 0..n_members-1, same offsets)."""

    def __init__(self, groups, x, y, z, m, vx, vy, vz):
        self.groups = groups
        self.x, self.y, self.z, self.m = x, y, z, m
        self.vx, self.vy, self.vz = vx, vy, vz


def gather_members(groups, x, y, z, m, vx=None, vy=None, vz=None):
    """One packed-record gather of the members' data (halma_gather_members)."""
    from .device import dtype_flag
    L = _lib_()
    x, y, z = _coords(x, y, z)
    m = _vec(m)
    if vx is not None:
        vx, vy, vz = _coords(vx, vy, vz)
    flags = [x.dtype, m.dtype] + ([vx.dtype] if vx is not None else [])
    odt = torch.float32 if all(f == torch.float32 for f in flags) else torch.float64
    nm = groups.n_members
    n = int(x.numel())
    dev = x.device
    outs = [torch.empty(max(nm, 1), dtype=odt, device=dev) for _ in range(7 if vx is not None else 4)]
    oflag = _lib.F32 if odt == torch.float32 else _lib.F64
    if nm > 0:
        ws = workspace(L.halma_gather_workspace_size(n, oflag))
        v = (ptr(vx), ptr(vy), ptr(vz)) if vx is not None else (None, None, None)
        ov = (ptr(outs[4]), ptr(outs[5]), ptr(outs[6])) if vx is not None else (None, None, None)
        rc = L.halma_gather_members(ptr(groups.members), nm, n, ptr(x), ptr(y), ptr(z),
                                    dtype_flag(x), *v,
                                    dtype_flag(vx) if vx is not None else _lib.F64,
                                    ptr(m), dtype_flag(m), oflag, ptr(outs[0]), ptr(outs[1]),
                                    ptr(outs[2]), ptr(outs[3]), *ov, ptr(ws), ws.numel(),
                                    stream_ptr())
        _lib.check(rc, "halma_gather_members")
    local = Groups(torch.arange(max(nm, 1), dtype=torch.int32, device=dev), groups.offsets,
                   groups.offsets_host)
    outs = [o[:nm] for o in outs]
    if vx is None:
        outs += [None, None, None]
    return MemberData(local, outs[0], outs[1], outs[2], outs[3], outs[4], outs[5], outs[6])


def group_reduce(groups, x, y, z, m, vx=None, vy=None, vz=None, centres=None):
    """Per-group reductions, float64 (G, 16): [M, cx, cy, cz, vx, vy, vz, rmax,
    xmin, ymin, zmin, xmax, ymax, zmax, sum|m|, exact-sum flag]."""
    from .device import dtype_flag
    L = _lib_()
    x, y, z = _coords(x, y, z)
    m = _vec(m)
    G = groups.n_groups
    out = torch.zeros((max(G, 1), _lib.GROUP_FIELDS), dtype=torch.float64, device=x.device)
    if G == 0:
        return out[:0]
    if vx is not None:
        vx, vy, vz = _coords(vx, vy, vz)
        vflag = dtype_flag(vx)
    else:
        vflag = _lib.F64
    cen = None
    if centres is not None:
        cen = to_dev(np.asarray(centres, np.float64).reshape(G, 3), np.float64)
    rc = L.halma_group_reduce(G, ptr(groups.offsets), ptr(groups.members), ptr(x), ptr(y),
                              ptr(z), dtype_flag(x), ptr(vx), ptr(vy), ptr(vz), vflag, ptr(m),
                              dtype_flag(m), ptr(cen), ptr(out), stream_ptr())
    _lib.check(rc, "halma_group_reduce")
    return out


def density_peak(groups, x, y, z, m, ll, reduced=None):
    """density_peak (halo_properties.py:518-556) of every group: (G, 3) float64."""
    from .device import dtype_flag
    L = _lib_()
    x, y, z = _coords(x, y, z)
    m = _vec(m)
    G = groups.n_groups
    if G == 0:
        return torch.zeros((0, 3), dtype=torch.float64, device=x.device)
    if reduced is None:
        reduced = group_reduce(groups, x, y, z, m)
    out = torch.zeros((G, 3), dtype=torch.float64, device=x.device)
    ws = workspace(L.halma_density_peak_workspace_size(G))
    rc = L.halma_density_peak(G, ptr(groups.offsets), ptr(groups.members), ptr(x), ptr(y),
                              ptr(z), dtype_flag(x), ptr(m), dtype_flag(m), float(ll),
                              ptr(reduced), ptr(out), ptr(ws), ws.numel(), stream_ptr())
    _lib.check(rc, "halma_density_peak")
    return out


def half_mass_radius(groups, x, y, z, m, reduced=None):
    """half_mass_radius (halo_properties.py:366-392) of every group: float64 (G,)."""
    from .device import dtype_flag
    L = _lib_()
    x, y, z = _coords(x, y, z)
    m = _vec(m)
    G = groups.n_groups
    out = torch.empty(max(G, 1), dtype=torch.float64, device=x.device)
    if G == 0:
        return out[:0]
    if reduced is None:
        reduced = group_reduce(groups, x, y, z, m)
    nm = groups.n_members
    ws = workspace(L.halma_half_mass_radius_workspace_size(nm))
    rc = L.halma_half_mass_radius(G, nm, ptr(groups.offsets), ptr(groups.members), ptr(x),
                                  ptr(y), ptr(z), dtype_flag(x), ptr(m), dtype_flag(m),
                                  ptr(reduced), ptr(out), ptr(ws), ws.numel(), stream_ptr())
    _lib.check(rc, "halma_half_mass_radius")
    return out[:G]


class Substructure:
    """Result of split_substructure: the groups after splitting (CSR ``groups``: the
    original groups in order, emptied ones dropped, then the new groups, all filtered by
    ``minp``), the number of substructures found, and for each output group the index of
    the original group it came from and whether it is a new (split-off) group."""

    def __init__(self, groups, n_found, parent, is_new):
        self.groups = groups
        self.n_found = n_found
        self.parent = parent    # numpy int64 per output group
        self.is_new = is_new    # numpy bool per output group

    def to_lists(self):
        return self.groups.to_lists()


def split_substructure(groups, x, y, z, m, ll, minp_sub=25, dec_factor=2.0, min_ll=0.5e-3,
                       dist_factor=0.75, minp=50):
    """The substructure (bridge) splitting of pyHALMA.py:703-766 for every group at once.

    For sub-linking lengths ll / dec_factor^k above ``min_ll``: the FoF groups (at least
    ``minp_sub`` members) of each group's remaining members that hold between 1 % and
    60 % of them and whose centre of mass lies further than ``dist_factor`` times the
    group's half-mass radius from the group's centre become new groups and leave the
    group.  A group left with under 30 % of its members is emptied.  Output order: the
    original groups, then the new ones by original group and discovery order; groups
    under ``minp`` members are dropped (pyHALMA.py:758-764).  One batched FoF per level
    runs over the members of all groups together: members of different groups are more
    than ``ll`` apart, so no sub-group crosses two groups.
    """
    x, y, z = _coords(x, y, z)
    m = _vec(m)
    dev = x.device
    H = groups.n_groups
    cur = groups.members[: groups.n_members].to(torch.int64)   # particle indices
    cur_off = np.asarray(groups.offsets_host, np.int64).copy()
    start = np.diff(cur_off)
    new_mem = []   # per level: particle indices of the new groups, in sub-group order
    new_tab = []   # rows (halo, level, order, level start, length, tensor id)
    sub_ll = ll / dec_factor
    level = 0
    while sub_ll > min_ll and cur.numel() > 0:
        sizes = np.diff(cur_off)
        sizes_d = torch.from_numpy(sizes).to(dev)
        halo_of_pos = torch.repeat_interleave(torch.arange(H, device=dev), sizes_d)
        cg = Groups(cur.to(torch.int32), torch.from_numpy(cur_off).to(dev), cur_off)
        red = group_reduce(cg, x, y, z, m)
        r12 = half_mass_radius(cg, x, y, z, m, red)
        md = gather_members(cg, x, y, z, m)
        lab = fof_labels(md.x, md.y, md.z, sub_ll)
        sub = groups_from_labels(lab, minp_sub)
        if sub.n_groups > 0:
            soh = np.asarray(sub.offsets_host, np.int64)
            soff = torch.from_numpy(soh).to(dev)
            smem = sub.members[: sub.n_members].to(torch.int64)     # concat positions
            h = halo_of_pos[smem[soff[:-1]]]
            ssize = torch.from_numpy(np.diff(soh)).to(dev).to(torch.float64)
            frac = ssize / sizes_d.to(torch.float64)[h]
            rs = group_reduce(Groups(sub.members, soff, soh), md.x, md.y, md.z, md.m)
            dist = torch.sqrt((red[h, 1] - rs[:, 1]) ** 2 + (red[h, 2] - rs[:, 2]) ** 2
                              + (red[h, 3] - rs[:, 3]) ** 2)
            sel = (frac < 0.6) & (frac > 0.01) & (dist > dist_factor * r12[h])
            sel_h = sel.cpu().numpy()
            if sel_h.any():
                sub_of_pos = torch.repeat_interleave(
                    torch.arange(sub.n_groups, device=dev), torch.from_numpy(np.diff(soh)).to(dev))
                taken = smem[sel[sub_of_pos]]                       # concat positions, in order
                keep = torch.ones(cur.numel(), dtype=torch.bool, device=dev)
                keep[taken] = False
                idx = np.nonzero(sel_h)[0]
                lens = np.diff(soh)[idx]
                starts = np.concatenate([[0], np.cumsum(lens)[:-1]])
                hh = h.cpu().numpy()[idx]
                for q in range(idx.size):
                    new_tab.append((int(hh[q]), level, q, int(starts[q]), int(lens[q]),
                                    len(new_mem)))
                new_mem.append(cur[taken])
                # kept members per group: a prefix sum read at the group offsets (the
                # groups are contiguous, so per-group atomics would queue on one address)
                cs = torch.cat([torch.zeros(1, dtype=torch.int64, device=dev),
                                torch.cumsum(keep.to(torch.int64), 0)])
                cur_off = cs[torch.from_numpy(cur_off).to(dev)].cpu().numpy().astype(np.int64)
                cur = cur[keep]
        sub_ll = sub_ll / dec_factor
        level += 1
    # assemble: original groups (emptied below 30 %), then the new ones
    sizes = np.diff(cur_off)
    # pyHALMA.py:752: len(part_list) / starting_npart < 0.3 empties the group
    ok = (start > 0) & (sizes >= minp)
    ok &= np.where(start > 0, sizes / np.maximum(start, 1) >= 0.3, False)
    sizes_d = torch.from_numpy(sizes).to(dev)
    halo_of_pos = torch.repeat_interleave(torch.arange(H, device=dev), sizes_d)
    parts = [cur[torch.from_numpy(ok).to(dev)[halo_of_pos]]] if cur.numel() else []
    lens = list(sizes[ok])
    parent = list(np.nonzero(ok)[0])
    is_new = [False] * len(lens)
    new_tab.sort(key=lambda t: (t[0], t[1], t[2]))
    rows = [t for t in new_tab if t[4] >= minp]
    if rows:
        base = np.concatenate([[0], np.cumsum([t.numel() for t in new_mem])]).astype(np.int64)
        allnew = torch.cat(new_mem)
        st = np.array([base[t[5]] + t[3] for t in rows], np.int64)
        ln = np.array([t[4] for t in rows], np.int64)
        gather = np.repeat(st - np.concatenate([[0], np.cumsum(ln)[:-1]]), ln) + np.arange(ln.sum())
        parts.append(allnew[torch.from_numpy(gather).to(dev)])
        lens += list(ln)
        parent += [t[0] for t in rows]
        is_new += [True] * len(rows)
    off = np.concatenate([[0], np.cumsum(np.asarray(lens, np.int64))]).astype(np.int64)
    members = torch.cat(parts) if parts else torch.zeros(0, dtype=torch.int64, device=dev)
    out = Groups(members.to(torch.int32) if members.numel() else
                 torch.zeros(1, dtype=torch.int32, device=dev),
                 torch.from_numpy(off).to(dev), off)
    return Substructure(out, len(new_tab), np.asarray(parent, np.int64),
                        np.asarray(is_new, bool))


class CleanResult:
    """Result of phase_space_clean for G input groups (member order of the input Groups).

    control_clean: bool per member, clean_cell's control array; control: the driver's
    (all members kept where ``psc`` is off or the group has fewer than ``min_part``
    members).  clean: numpy (G, 5) float64, clean_cell's (cx, cy, cz, M, R).  n_cell: the
    first grid's node count per axis.  rounds: test rounds per group.  which: device
    int32 (n_members, 3), the node indices of the last calc_cell, group g's rows at
    offsets[g] .. offsets[g] + which_len[g].  kept: the groups with at least ``minp``
    members left (CSR over the kept members), kept_index their input indices, and their
    recomputed centre (K, 3), mass, rmax, npart and n_cell (pyHALMA.py:823-843)."""

    def __init__(self, **kw):
        self.__dict__.update(kw)

    def which_of(self, g):
        a = int(self.offsets_host[g])
        return self.which[a: a + int(self.which_len[g])].cpu().numpy()


def _grid(R, ll):
    """np.arange(-(R + ll), R + ll, 2 ll) per group as (start, second node, delta, length):
    node k is start (k = 0), the second node (k = 1), start + k * delta (k >= 2)."""
    stop = R + ll
    start = -stop
    step = 2.0 * ll
    n = np.ceil((stop - start) / step).astype(np.int64)
    g1 = start + step
    return start, g1, g1 - start, n


def _centres(red, counts):
    """(centre (G, 3), M, R) of group_reduce rows, zeros for groups without mass (the
    reference's center_of_mass returns zeros when M <= 0, furthest_particle 0 when empty)."""
    M = red[:, 0].copy()
    c = red[:, 1:4].copy()
    R = red[:, 7].copy()
    nomass = ~(M > 0)
    c[nomass] = 0.0
    M[nomass] = 0.0
    R[counts == 0] = 0.0
    return c, M, R


def _clean_cells(L, ns, H, s_mem, s_grp, x, y, z, vx, vy, vz, m, c, grid, used, ijk, stats,
                 ws):
    """halma_clean_cells over the set (s_mem, s_grp) with per-group centres c (H, 3) and
    grids (start, second node, delta, length); ``used`` marks the groups with members in
    the set.  Returns the number of cells and the device copies of the per-group
    arguments (to be held until the stream has used them)."""
    from .device import dtype_flag
    dev = x.device
    start, g1, d, n = grid
    max_n = int(n[used].max())
    dv = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)  # noqa: E731
    held = (dv(c), dv(start), dv(g1), dv(d), dv(np.clip(n, 1, 2 ** 30).astype(np.int32)))
    ncells = ctypes_i64()
    rc = L.halma_clean_cells(ns, H, ptr(s_mem), ptr(s_grp), ptr(x), ptr(y), ptr(z),
                             dtype_flag(x), ptr(vx), ptr(vy), ptr(vz), dtype_flag(vx), ptr(m),
                             dtype_flag(m), *(ptr(t) for t in held), max_n, ptr(ijk), ptr(stats),
                             ncells, ptr(ws), ws.numel(), stream_ptr())
    _lib.check(rc, "halma_clean_cells")
    return int(ncells.value), held


def cell_statistics(x, y, z, vx, vy, vz, m, members, centre, grid_start, grid_g1, grid_d,
                    grid_n):
    """calc_cell (halo_properties.py:113-159) of one group: node indices (n, 3) int32 and
    the occupied cells' statistics (cells, 6): mass-weighted velocity (3), mass, member
    count, dispersion; cells in increasing (ix n + iy) n + iz order."""
    L = _lib_()
    x, y, z = _coords(x, y, z)
    vx, vy, vz = _coords(vx, vy, vz)
    m = _vec(m)
    dev = x.device
    members = to_dev(members, np.int32)
    ns = int(members.numel())
    ijk = torch.empty((max(ns, 1), 3), dtype=torch.int32, device=dev)
    stats = torch.empty(6 * max(ns, 1), dtype=torch.float64, device=dev)
    if ns == 0:
        return ijk[:0], stats[:0].reshape(0, 6)
    ws = workspace(L.halma_clean_workspace_size(ns))
    grp = torch.zeros(ns, dtype=torch.int32, device=dev)
    grid = tuple(np.array([v]) for v in (grid_start, grid_g1, grid_d)) + (
        np.array([int(grid_n)], np.int64),)
    nc, _ = _clean_cells(L, ns, 1, members, grp, x, y, z, vx, vy, vz, m,
                         np.asarray(centre, np.float64).reshape(1, 3), grid, np.ones(1, bool),
                         ijk, stats, ws)
    return ijk[:ns], stats[: 6 * nc].reshape(nc, 6)


def _select(L, H, goff, nm, members, control, live, force, out, ws):
    """halma_clean_select into out = (set_members, set_groups, set_target, set_offsets);
    returns the set's size."""
    n = ctypes_i64()
    rc = L.halma_clean_select(H, nm, ptr(goff), ptr(members), ptr(control), ptr(live),
                              ptr(force), *(ptr(t) for t in out), n, ptr(ws), ws.numel(),
                              stream_ptr())
    _lib.check(rc, "halma_clean_select")
    return int(n.value)


def phase_space_clean(groups, x, y, z, vx, vy, vz, m, ll, q_fil, sig_fil, psc=True,
                      min_part=50, minp=50, centres=None, radii=None):
    """The phase-space cleaning of pyHALMA.py:789-843 (calc_cell / clean_cell,
    halo_properties.py:113-278) for every group at once.

    Each round works on the set S of members still flagged of every group not finished
    (halma_clean_select): one group_reduce gives the centres and radii of the grids
    (np.arange(-(R + ll), R + ll, 2 ll) about the centre of mass), halma_clean_cells the
    cell statistics and halma_clean_decide the removals.  The members' data are gathered
    once into member order when the seven arrays share one dtype, so the rounds read
    them contiguously.  The per-group bookkeeping (remaining, the fraction removed,
    fac_sig: 2, 1.5, then x 0.75 with a floor of 1) is the reference's, quirks included:
    in rounds after the first the decision about the ip-th member of S clears entry ip
    of the group's control array.  A group that removes all of its members in the first
    round ends with zeros, as the reference's final step does with no members.  A cell
    with more than two members whose members share one velocity has sigma 0: its
    members then score inf (removed) or nan (kept) where the reference's compiled
    division raises.  ``centres`` (G, 3) and ``radii`` (G,), when given, replace the
    first round's centre of mass and furthest-member distance (the driver's own values,
    which clean_cell receives from its caller).
    """
    from .device import dtype_flag
    L = _lib_()
    if not ll > 0:
        raise ValueError("ll must be positive")
    x, y, z = _coords(x, y, z)
    vx, vy, vz = _coords(vx, vy, vz)
    m = _vec(m)
    dev = x.device
    H = groups.n_groups
    off = np.asarray(groups.offsets_host, np.int64)
    sizes = np.diff(off)
    nm = groups.n_members
    mem = groups.members[:nm]
    if H == 0:
        none = torch.zeros(0, dtype=torch.bool, device=dev)
        return CleanResult(control=none, control_clean=none, clean=np.zeros((0, 5)),
                           n_cell=np.zeros(0, np.int64), rounds=np.zeros(0, np.int64),
                           which=torch.zeros((0, 3), dtype=torch.int32, device=dev),
                           which_len=np.zeros(0, np.int64), offsets_host=off, kept=groups,
                           kept_index=np.zeros(0, np.int64), centre=np.zeros((0, 3)),
                           mass=np.zeros(0), rmax=np.zeros(0), npart=np.zeros(0, np.int64),
                           kept_n_cell=np.zeros(0, np.int64))
    goff = torch.from_numpy(off).to(dev)
    if x.dtype == vx.dtype == m.dtype and nm > 0:
        md = gather_members(groups, x, y, z, m, vx, vy, vz)
        A = (md.x, md.y, md.z, md.vx, md.vy, md.vz, md.m)
        by = None          # the set holds member positions into the gathered arrays
    else:
        A = (x, y, z, vx, vy, vz, m)
        by = mem           # the set holds particle indices
    control = torch.ones(max(nm, 1), dtype=torch.uint8, device=dev)
    sel = (torch.empty(max(nm, 1), dtype=torch.int32, device=dev),
           torch.empty(max(nm, 1), dtype=torch.int32, device=dev),
           torch.empty(max(nm, 1), dtype=torch.int64, device=dev),
           torch.empty(H + 1, dtype=torch.int64, device=dev))
    s_mem, s_grp, target, s_off = sel
    sel_ws = workspace(L.halma_clean_select_workspace_size(nm))
    phase = np.zeros(H, np.int32)  # 0 first round, 1 later round, 2 final grid, 3 done
    fac = np.full(H, 2.0)
    remaining = sizes.copy()
    rounds = np.zeros(H, np.int64)
    n_cell = np.zeros(H, np.int64)
    clean = np.zeros((H, 5))
    which = torch.zeros((max(nm, 1), 3), dtype=torch.int32, device=dev)
    which_len = np.zeros(H, np.int64)
    zero_ended = np.zeros(H, bool)  # emptied by a later round: clean_cell returned zeros
    ws = workspace(L.halma_clean_workspace_size(nm))
    ijk = torch.empty((max(nm, 1), 3), dtype=torch.int32, device=dev)
    stats = torch.empty(6 * max(nm, 1), dtype=torch.float64, device=dev)
    cleaned = torch.zeros(max(H, 1), dtype=torch.int32, device=dev)
    vf = dtype_flag(A[3])
    given = None
    if centres is not None:
        given = (np.asarray(centres, np.float64).reshape(H, 3).copy(),
                 np.asarray(radii, np.float64).reshape(H).copy())

    def dv(a):
        return torch.from_numpy(np.ascontiguousarray(a)).to(dev)

    while (phase < 3).any():
        live = dv((phase < 3).astype(np.uint8))
        ns = _select(L, H, goff, nm, by, control, live, None, sel, sel_ws)
        s_off_h = s_off.cpu().numpy()
        counts_h = np.diff(s_off_h)
        red = group_reduce(Groups(s_mem, s_off, s_off_h), A[0], A[1], A[2], A[6]).cpu().numpy()
        c, M, R = _centres(red, counts_h)
        if given is not None:
            c, R = given  # the caller's first grid (every group is in its first round)
            given = None
        grid = _grid(R, ll)
        first = phase == 0
        n_cell[first] = grid[3][first]
        if ns:
            _, held = _clean_cells(L, ns, H, s_mem, s_grp, *A, c, grid, counts_h > 0, ijk,
                                   stats, ws)
            which.index_copy_(0, target[:ns], ijk[:ns])
            cleaned.zero_()
            phase_d, fac_d = dv(phase), dv(fac)  # held until the kernel has read them
            rc = L.halma_clean_decide(ns, ptr(s_mem), ptr(s_grp), ptr(target), ptr(A[3]),
                                      ptr(A[4]), ptr(A[5]), vf, ptr(stats), ptr(phase_d),
                                      ptr(fac_d), float(q_fil), float(sig_fil), ptr(control),
                                      ptr(cleaned), ptr(ws), ws.numel(), stream_ptr())
            _lib.check(rc, "halma_clean_decide")
            cl = cleaned[:H].cpu().numpy().astype(np.int64)
        else:
            cl = np.zeros(H, np.int64)
        tested = (phase == 0) | (phase == 1)
        rounds[tested] += 1
        # groups ending this round keep its grid's node indices
        ends = phase == 2
        p0 = phase == 0
        p1 = phase == 1
        remaining[p0] = sizes[p0] - cl[p0]
        remaining[p1] -= cl[p1]
        zero = p1 & (remaining == 0)
        zero_ended |= zero
        ends |= zero
        which_len[ends] = counts_h[ends]
        fin = phase == 2
        clean[fin] = np.column_stack([c, M, R])[fin]
        phase[ends] = 3
        with np.errstate(divide="ignore", invalid="ignore"):
            frac = np.where(remaining != 0, cl / np.where(remaining != 0, remaining, 1), 0.0)
        go0 = p0 & (remaining > 0) & (frac > 0.1)
        phase[p0] = np.where(go0[p0], 1, 2)
        fac[p0] = 2.0 * 0.75
        live1 = p1 & ~zero
        phase[live1] = np.where(frac[live1] > 0.1, 1, 2)
        fac[live1] = np.maximum(1.0, fac[live1] * 0.75)
    # the driver (pyHALMA.py:819-843): every member kept where psc is off or the group is
    # small; centre, mass and radius again over the kept members; groups of minp or more.
    # A group that ended on its final grid keeps that grid's set, whose centre, mass and
    # radius are already known (same set, same reduction); the others are reduced again.
    keep_all = np.ones(H, bool) if not psc else sizes < min_part
    force = dv(keep_all.astype(np.uint8))
    control_clean = control[:nm].bool()
    kc, kM, kR = clean[:, 0:3].copy(), clean[:, 3].copy(), clean[:, 4].copy()
    kcount = which_len.copy()
    redo = keep_all | zero_ended
    if redo.any():
        ns = _select(L, H, goff, nm, by, control, dv(redo.astype(np.uint8)), force, sel,
                     sel_ws)
        cnt = np.diff(s_off.cpu().numpy())
        kred = group_reduce(Groups(s_mem, s_off, np.concatenate([[0], np.cumsum(cnt)])),
                            A[0], A[1], A[2], A[6]).cpu().numpy()
        rcen, rM, rR = _centres(kred, cnt)
        kc[redo], kM[redo], kR[redo], kcount[redo] = rcen[redo], rM[redo], rR[redo], cnt[redo]
    ok = kcount >= minp
    grp_of = torch.repeat_interleave(torch.arange(H, device=dev),
                                     torch.from_numpy(sizes).to(dev))
    final = control_clean | force.bool()[grp_of]
    ns = _select(L, H, goff, nm, mem, control, dv(ok.astype(np.uint8)), force, sel, sel_ws)
    out_off = np.concatenate([[0], np.cumsum(kcount[ok])]).astype(np.int64)
    kept = Groups(s_mem[:max(ns, 1)].clone(), torch.from_numpy(out_off).to(dev), out_off)
    return CleanResult(control=final, control_clean=control_clean, clean=clean, n_cell=n_cell,
                       rounds=rounds, which=which, which_len=which_len, offsets_host=off,
                       kept=kept, kept_index=np.nonzero(ok)[0], centre=kc[ok], mass=kM[ok],
                       rmax=kR[ok], npart=kcount[ok], kept_n_cell=n_cell[ok])


def progenitors(new_groups, oripa, masses, old_npart, old_oripas, omm):
    """The merger section of pyHALMA.py:1211-1266 for every new halo at once
    (halma_progenitors): new haloes ``new_groups`` (particle indices) with per-particle
    IDs ``oripa`` (st_oripa) and masses; old haloes as the checkpoint holds them
    (pyHALMA.py:419-428): members per halo ``old_npart``, their IDs concatenated
    ``old_oripas``, masses ``omm``.  Returns int32 numpy arrays (pro1, pro2, n_mergers,
    merger_type)."""
    L = _lib_()
    H = new_groups.n_groups
    out = [np.zeros(H, np.int32) for _ in range(4)]
    if H == 0:
        return tuple(out)
    nm = new_groups.n_members
    ids = to_dev(oripa, np.int64)
    new_ids = ids[new_groups.members[:nm].long()] if nm else ids[:0]
    new_mass = to_dev(np.asarray(masses, np.float64).reshape(H), np.float64)
    old_npart = np.asarray(old_npart, np.int64).ravel()
    G = old_npart.size
    old_off = to_dev(np.concatenate([[0], np.cumsum(old_npart)]).astype(np.int64), np.int64)
    no = int(old_npart.sum())
    if no == 0:
        old_ids = new_ids[:0]
    elif isinstance(old_oripas, torch.Tensor):
        old_ids = to_dev(old_oripas.reshape(-1)[:no], np.int64)
    else:
        old_ids = to_dev(np.asarray(old_oripas).ravel()[:no], np.int64)
    old_mass = to_dev(np.asarray(omm, np.float64).ravel()[:G], np.float64) if G else new_mass
    dev = new_mass.device
    res = [torch.empty(H, dtype=torch.int32, device=dev) for _ in range(4)]
    ws = workspace(L.halma_progenitors_workspace_size(H, nm, G, no))
    rc = L.halma_progenitors(H, nm, ptr(new_groups.offsets), ptr(new_ids), ptr(new_mass), G, no,
                             ptr(old_off), ptr(old_ids), ptr(old_mass), *(ptr(r) for r in res),
                             ptr(ws), ws.numel(), stream_ptr())
    _lib.check(rc, "halma_progenitors")
    return tuple(r.cpu().numpy() for r in res)


def ball_query(x, y, z, centres, radii, strict=False, m=None, cell=0.0, sorted=True,
               count_only=False):
    """The points within radii[h] of centres[h] for every sphere h (halma_ball_query): the
    KD-tree query_ball_point of halo_gas.py:255,269 and pyHALMA.py:1179 (distance <= R),
    or parallel_inside of halo_gas.py:218-220 (``strict``: distance < R).  Returns
    (offsets (numpy int64, H+1), indices (int32 device tensor, ascending per sphere when
    ``sorted``; None when ``count_only``), mass sums (numpy FP64 per sphere, or None
    without ``m``))."""
    from .device import dtype_flag
    L = _lib_()
    x, y, z = _coords(x, y, z)
    dev = x.device
    c = to_dev(np.asarray(centres, np.float64).reshape(-1, 3), np.float64)
    H = int(c.shape[0])
    r = to_dev(np.asarray(radii, np.float64).reshape(H), np.float64)
    n = int(x.numel())
    mm = _vec(m) if m is not None else None
    sums = torch.zeros(max(H, 1), dtype=torch.float64, device=dev) if m is not None else None
    off = torch.zeros(H + 1, dtype=torch.int64, device=dev)
    found = ctypes_i64()
    max_found = 0 if count_only else max(n, 1)
    for _ in range(2):
        idx = None if count_only else torch.empty(max(max_found, 1), dtype=torch.int32,
                                                  device=dev)
        ws = workspace(L.halma_ball_query_workspace_size(n, H, max_found))
        rc = L.halma_ball_query(n, ptr(x), ptr(y), ptr(z), dtype_flag(x), ptr(mm),
                                dtype_flag(mm) if mm is not None else _lib.F64, H, ptr(c), ptr(r),
                                int(bool(strict)), float(cell), int(bool(sorted)), max_found,
                                ptr(off), ptr(idx), ptr(sums), found, ptr(ws), ws.numel(),
                                stream_ptr())
        if rc == _lib.E_LIMIT and not count_only and found.value > max_found:
            max_found = int(found.value)  # more hits than points: size the output again
            continue
        _lib.check(rc, "halma_ball_query")
        break
    F = int(found.value)
    return (off.cpu().numpy(), None if count_only else idx[:F],
            sums[:H].cpu().numpy() if sums is not None else None)


def halo_shapes(offsets, x, y, z, mass):
    """halo_shape (particle_subroutines.f90:160-214) of many groups at once
    (halma_halo_shape): offsets (G+1) over member-ordered FP32 coordinates relative to
    each group's centre and masses.  Returns (G, 3) float32 device tensor."""
    L = _lib_()
    off = to_dev(np.asarray(offsets, np.int64), np.int64)
    G = int(off.numel()) - 1
    arrs = [to_dev(np.asarray(a, np.float32) if not isinstance(a, torch.Tensor) else a,
                   np.float32) for a in (x, y, z, mass)]
    out = torch.zeros((max(G, 1), 3), dtype=torch.float32, device=off.device)
    rc = L.halma_halo_shape(G, ptr(off), *(ptr(a) for a in arrs), ptr(out), stream_ptr())
    _lib.check(rc, "halma_halo_shape")
    return out[:G]


def sigma_projections(grid, x, y, z, vx, vy, vz, mass, cx, cy, cz, R05x, R05y, R05z, ll):
    """sigma_projections (particle_subroutines.f90:217-461) of one halo
    (halma_sigma_projections): FP32 member arrays (velocities relative to the halo's),
    the grid of cell nodes; returns numpy float32 (SIG_1D_x_05, SIG_1D_y_05,
    SIG_1D_z_05, V_sigma, lambda)."""
    L = _lib_()
    g = to_dev(np.asarray(grid, np.float32), np.float32)
    arrs = [to_dev(np.asarray(a, np.float32), np.float32) for a in (x, y, z, vx, vy, vz, mass)]
    n = int(arrs[0].numel())
    nc = int(g.numel())
    out = torch.zeros(5, dtype=torch.float32, device=g.device)
    ws = workspace(L.halma_sigma_projections_workspace_size(n, nc))
    f = [float(np.float32(v)) for v in (cx, cy, cz, R05x, R05y, R05z, ll)]
    rc = L.halma_sigma_projections(n, *(ptr(a) for a in arrs), ptr(g), nc, *f, ptr(out),
                                   ptr(ws), ws.numel(), stream_ptr())
    _lib.check(rc, "halma_sigma_projections")
    return out.cpu().numpy()


class UnbindResult:
    def __init__(self, bound, iters, vcm, pot, stats):
        self.bound = bound      # uint8 device tensor, member order
        self.iters = iters      # int32 device tensor per halo
        self.vcm = vcm          # float64 device tensor (H, 3)
        self.pot = pot          # float64 device tensor per member (last S) or None
        self.stats = stats      # dict


def unbind(groups, x, y, z, vx, vy, vz, m, factor=1.0, max_iter=100, use_graph=False,
           with_pot=False, time_pot=False):
    """Batched iterative unbinding of every group (see halma_unbind_batched)."""
    from .device import dtype_flag
    L = _lib_()
    x, y, z = _coords(x, y, z)
    vx, vy, vz = _coords(vx, vy, vz)
    m = _vec(m)
    H = groups.n_groups
    nm = groups.n_members
    dev = x.device
    # every entry of these is written by the call (no clearing launches)
    bound = torch.empty(max(nm, 1), dtype=torch.uint8, device=dev)
    iters = torch.empty(max(H, 1), dtype=torch.int32, device=dev)
    vcm = torch.empty((max(H, 1), 3), dtype=torch.float64, device=dev)
    pot = (torch.full((max(nm, 1),), float("nan"), dtype=torch.float64, device=dev)
           if with_pot else None)
    import ctypes
    st = (ctypes.c_int64 * 13)()
    if H > 0:
        off = _lib.i64_array(groups.offsets_host)
        ws = workspace(L.halma_unbind_workspace_size(H, off))
        rc = L.halma_unbind_batched(H, off, ptr(groups.members), ptr(x), ptr(y), ptr(z),
                                    dtype_flag(x), ptr(vx), ptr(vy), ptr(vz), dtype_flag(vx),
                                    ptr(m), dtype_flag(m), float(factor), int(max_iter),
                                    (_lib.UNBIND_GRAPHS if use_graph else 0)
                                    | (_lib.UNBIND_TIME_POT if time_pot else 0), ptr(bound),
                                    ptr(iters), ptr(vcm), ptr(pot), st, ptr(ws), ws.numel(),
                                    stream_ptr())
        _lib.check(rc, "halma_unbind_batched")
    s = dict(passes=int(st[0]), retested=int(st[1]), items=int(st[2]),
             tests_per_lane=int(st[3]), subtile=int(st[4]), pot_ns=int(st[5]),
             pot_launches=int(st[6]), pairs=int(st[7]), sym_ns=int(st[8]),
             sym_launches=int(st[9]), full_pairs=int(st[10]), f32_launches=int(st[11]),
             resummed=int(st[12]))
    return UnbindResult(bound[:nm], iters[:H], vcm[:H], None if pot is None else pot[:nm], s)


def unbind_plain_f64(groups, x, y, z, vx, vy, vz, m, factor=1.0, max_iter=100):
    """The plain FP64 comparator of the same method (per-halo launches)."""
    import ctypes
    from .device import dtype_flag
    L = _lib_()
    x, y, z = _coords(x, y, z)
    vx, vy, vz = _coords(vx, vy, vz)
    m = _vec(m)
    H = groups.n_groups
    nm = groups.n_members
    bound = torch.zeros(max(nm, 1), dtype=torch.uint8, device=x.device)
    iters = (ctypes.c_int32 * max(H, 1))()
    if H > 0:
        off = _lib.i64_array(groups.offsets_host)
        ws = workspace(L.halma_unbind_plain_workspace_size(H, off))
        rc = L.halma_unbind_plain_f64(H, off, ptr(groups.members), ptr(x), ptr(y), ptr(z),
                                      dtype_flag(x), ptr(vx), ptr(vy), ptr(vz), dtype_flag(vx),
                                      ptr(m), dtype_flag(m), float(factor), int(max_iter),
                                      ptr(bound), iters, ptr(ws), ws.numel(), stream_ptr())
        _lib.check(rc, "halma_unbind_plain_f64")
    return bound[:nm], np.array(iters[:H], dtype=np.int32)


def tune_set(key, value):
    _lib.check(_lib_().halma_tune_set(int(key), int(value)), "halma_tune_set")


def tune_get(key):
    return int(_lib_().halma_tune_get(int(key)))
