"""Differentiable halo potential and the potential-minimum centring.

The softened pair sum of a halo,

    S_i = sum_j m_j (|x_j - t_i|^2 + eps^2)^(-1/2)

(the sum of particle_subroutines.f90:466-514 with a softening length and the same
three-coordinate skip rule), is exposed as a torch.autograd.Function whose backward is
the library's hand-written gradient kernel (the per-test acceleration
sum_j m_j (x_j - t_i) (r^2 + eps^2)^(-3/2)); the pair terms are recomputed, never
stored.  Gradients with respect to the sources (positions and masses) reuse the same
kernels with tests and sources exchanged.

``potential_energy`` is the halo's softened potential energy
W_h = -G/2 sum_{i in h} m_i S_i as a differentiable function of the particle positions
and masses (dW/dx_k = -G m_k a_k, dW/dm_k = -G S_k, both by the symmetry of the pair
terms), through ``HaloPotentialEnergy``: the forward pass is the library's softened sum
of every member against its halo (tests = the members' own FP32 positions, so the rule
drops each self pair), the backward pass its hand-written acceleration kernel over the
same pairs, recomputed.  ``refine_centres`` moves a seed centre per halo downhill on
the potential Phi(c) = -G S(c): every step takes the autograd gradient of -S at the
current centres and scales it per halo by 1 / w(c), w = sum_j m_j (r^2+eps^2)^(-3/2),
which makes the step the fixed-point update c <- sum_j m_j w_j x_j / sum_j m_j w_j and
never increases Phi (the kernel (r^2 + eps^2)^(-1/2) is convex and decreasing in r^2).
The reference's discrete counterpart is the most-bound particle of
halo_gas.py:498-634.
"""
import numpy as np
import torch

from . import engine

G_CONST = 4.3 * 1e-3 * 1e-6  # (km/s)^2 Mpc / Msun, halo_gas.py:459-460


def _cols(p):
    return p[:, 0].contiguous(), p[:, 1].contiguous(), p[:, 2].contiguous()


_src_cache = {}


def _src_cols(src_pos):
    """FP32 columns of the source positions, kept for the next call on the same unchanged
    tensor (the centring iterations pass one source array to every forward and backward).
    The key holds the storage address, shape, strides, dtype and the version counter
    (bumped by every in-place change); the entry holds the tensor itself, so its storage
    cannot be freed and reused under the same address while it is cached."""
    p = src_pos.detach()
    key = (p.device, p.data_ptr(), tuple(p.shape), p.stride(), p.dtype, p._version)
    hit = _src_cache.get("src")
    if hit is not None and hit[0] == key:
        return hit[2]
    cols = _cols(p.float())
    _src_cache["src"] = (key, p, cols)
    return cols


def _key(t):
    t = t.detach()
    return (t.device, t.data_ptr(), tuple(t.shape), t.stride(), t.dtype, t._version)


def _src_packed(src_pos, src_m):
    """The sources as (x, y, z, m) FP32 rows for halma_pot_soft_batched_packed, kept like
    the columns above (keyed on both arrays, holding both)."""
    key = (_key(src_pos), _key(src_m))
    hit = _src_cache.get("packed")
    if hit is not None and hit[0] == key:
        return hit[2]
    packed = torch.cat([src_pos.detach().float(), src_m.detach().float()[:, None]],
                       1).contiguous()
    _src_cache["packed"] = (key, (src_pos.detach(), src_m.detach()), packed)
    return packed


def clear_cache():
    """Drop the cached source columns and packed sources (frees their device memory)."""
    _src_cache.clear()


class SoftPotential(torch.autograd.Function):
    """S (float64, one per test) for tests (T, 3) float64 against the sources of their
    halo: src_pos (N, 3), src_m (N,), CSR host offsets src_off / tst_off (H+1 each).
    ``aux`` (a dict or None) receives ``wsum`` (one per test) when the test gradient is
    computed."""

    @staticmethod
    def forward(ctx, tests, src_pos, src_m, src_off, tst_off, eps, aux=None):
        tx, ty, tz = _cols(tests.detach().double())
        # with aux (the centring, which always takes the test gradient next) the gradient
        # and w come from the same launch as S: one pass over the sources per step
        fused = aux is not None and ctx.needs_input_grad[0]
        r = engine.pot_soft_batched(src_off, None, None, None, None, tst_off, tx, ty, tz,
                                    eps, pot=True, acc=fused, wsum=fused,
                                    packed=_src_packed(src_pos, src_m))
        ctx.save_for_backward(tests, src_pos, src_m)
        ctx.src_off, ctx.tst_off, ctx.eps, ctx.aux = src_off, tst_off, float(eps), aux
        ctx.fused = (r["acc"], r["wsum"]) if fused else None
        return r["pot"]

    @staticmethod
    def backward(ctx, g):
        tests, src_pos, src_m = ctx.saved_tensors
        g = g.detach().double().contiguous()
        gt = gp = gm = None
        tx, ty, tz = _cols(tests.detach().double())
        if ctx.needs_input_grad[0]:
            if ctx.fused is not None:
                r = {"acc": ctx.fused[0], "wsum": ctx.fused[1]}
            else:
                r = engine.pot_soft_batched(ctx.src_off, None, None, None, None, ctx.tst_off,
                                            tx, ty, tz, ctx.eps, pot=False, acc=True,
                                            wsum=ctx.aux is not None,
                                            packed=_src_packed(src_pos, src_m))
            gt = (g[:, None] * r["acc"]).to(tests.dtype)
            if ctx.aux is not None:
                ctx.aux["wsum"] = r["wsum"]
        if ctx.needs_input_grad[1] or ctx.needs_input_grad[2]:
            sx, sy, sz = _src_cols(src_pos)
            # dS_i/dx_j = m_j (t_i - x_j)(r^2+eps^2)^-3/2 and dS_i/dm_j = (r^2+eps^2)^-1/2:
            # the same sums with the tests as sources (masses g_i) and the sources as tests
            r = engine.pot_soft_batched(ctx.tst_off, g.float(), tx.float(), ty.float(),
                                        tz.float(), ctx.src_off, sx.double(), sy.double(),
                                        sz.double(), ctx.eps, pot=ctx.needs_input_grad[2],
                                        acc=ctx.needs_input_grad[1])
            if ctx.needs_input_grad[1]:
                gp = (src_m.detach().double()[:, None] * r["acc"]).to(src_pos.dtype)
            if ctx.needs_input_grad[2]:
                gm = r["pot"].to(src_m.dtype)
        return gt, gp, gm, None, None, None, None


def soft_potential(tests, src_pos, src_m, src_off, tst_off, eps, aux=None):
    return SoftPotential.apply(tests, src_pos, src_m, np.asarray(src_off, np.int64),
                               np.asarray(tst_off, np.int64), float(eps), aux)


def _halo_sums(v, offsets_dev, offsets):
    """Per-halo FP64 sums of v over the CSR ranges (the library's double-double group
    reduction: a fixed order, no atomics)."""
    n = int(v.numel())
    grp = engine.Groups(torch.arange(max(n, 1), dtype=torch.int32, device=v.device),
                        offsets_dev, offsets)
    z = torch.zeros(max(n, 1), dtype=torch.float32, device=v.device)
    vv = v if n else torch.zeros(1, dtype=torch.float64, device=v.device)
    return engine.group_reduce(grp, z, z, z, vv)[:, 0]


class HaloPotentialEnergy(torch.autograd.Function):
    """W_h = -G/2 sum_{i in h} m_i S_i (one per halo, float64) for members pos (N, 3) and
    masses m (N,) in halo order (CSR host offsets), S_i the softened sum over the other
    members of i's halo.  Forward: halma_pot_soft_self_batched (S); backward: the same
    entry's acceleration kernel (a), pair terms recomputed:
      dW/dx_k = -G m_k a_k,   dW/dm_k = -G S_k.
    With ``fuse`` and positions that require a gradient, the forward launch evaluates the
    acceleration with S in the same pass over the pairs (the differences, distances and
    the coordinate rule are shared) and keeps it (N x 3) for the backward, which then
    launches nothing; without it the backward recomputes every pair term."""

    @staticmethod
    def forward(ctx, pos, m, offsets, eps, G, fuse):
        packed = torch.cat([pos.detach().float(), m.detach().float()[:, None]], 1).contiguous()
        fused = bool(fuse) and ctx.needs_input_grad[0]
        r = engine.pot_soft_self(offsets, packed, eps, pot=True, acc=fused)
        S = r["pot"]
        off_dev = engine.to_dev(offsets, np.int64)
        W = -0.5 * G * _halo_sums(m.detach().double() * S, off_dev, offsets)
        ctx.save_for_backward(m, S)
        ctx.packed, ctx.offsets, ctx.off_dev, ctx.eps, ctx.G = packed, offsets, off_dev, eps, G
        ctx.acc = r["acc"] if fused else None
        ctx.pos_dtype = pos.dtype
        return W

    @staticmethod
    def backward(ctx, gW):
        m, S = ctx.saved_tensors
        seg = torch.repeat_interleave(
            torch.arange(len(ctx.offsets) - 1, device=S.device),
            torch.as_tensor(np.diff(ctx.offsets), device=S.device))
        g = gW.detach().double()[seg]
        gp = gm = None
        if ctx.needs_input_grad[0]:
            a = ctx.acc
            if a is None:
                a = engine.pot_soft_self(ctx.offsets, ctx.packed, ctx.eps, pot=False,
                                         acc=True)["acc"]
            gp = ((-ctx.G * g * m.detach().double())[:, None] * a).to(ctx.pos_dtype)
        if ctx.needs_input_grad[1]:
            gm = (-ctx.G * g * S).to(m.dtype)
        return gp, gm, None, None, None, None


def potential_energy(pos, m, offsets, eps, G=G_CONST, fuse=True):
    """W_h = -G/2 sum_{i in h} m_i S_i (float64, one per halo) for particle positions
    pos (N, 3) in halo order with CSR host offsets; differentiable in pos and m.  The
    pair terms use the FP32 positions and masses (as every pair sum of the library), so
    a member's own pair is dropped by the coordinate rule whatever the input precision.
    ``fuse``: when pos requires a gradient, the forward pass also evaluates the
    acceleration (one pass over the pairs for both); else the backward recomputes it."""
    offsets = np.asarray(offsets, np.int64)
    if offsets.ndim != 1 or len(offsets) < 1 or offsets[0] != 0 or offsets[-1] != len(pos) \
            or len(m) != len(pos) or np.any(np.diff(offsets) < 0):
        raise ValueError("offsets must be CSR offsets from 0 to len(pos) over pos and m")
    return HaloPotentialEnergy.apply(pos, m, offsets, float(eps), float(G), bool(fuse))


class CentringResult:
    def __init__(self, centres, S, iters, steps):
        self.centres = centres  # (H, 3) float64 device tensor
        self.S = S              # softened sum at the centres (Phi = -G S)
        self.iters = iters      # steps taken per halo
        self.steps = steps      # last step length per halo (Mpc)


def refine_centres(x, y, z, m, offsets, seeds, eps, max_iter=200, tol=1e-6):
    """Potential-minimum centre per halo.  Particles x, y, z, m in halo order (CSR
    host offsets, H+1 entries); seeds (H, 3) starting centres; eps the softening
    length (scalar or one per halo is not supported: one value).  Iterates until
    every halo's step is below tol * eps, or max_iter.  Each step is autograd's
    gradient of -sum S at the centres through SoftPotential, scaled by 1 / w."""
    offsets = np.asarray(offsets, np.int64)
    H = len(offsets) - 1
    dev = engine.to_dev(np.zeros(1), np.float64).device
    pos = torch.stack([engine.to_dev(a, np.float32) for a in (x, y, z)], dim=1)
    mt = engine.to_dev(m, np.float32)
    c = torch.as_tensor(np.asarray(seeds, np.float64), device=dev).clone()
    tst_off = np.arange(H + 1, dtype=np.int64)
    iters = torch.zeros(H, dtype=torch.int32, device=dev)
    active = torch.ones(H, dtype=torch.bool, device=dev)
    step_len = torch.full((H,), float("inf"), dtype=torch.float64, device=dev)
    S = None
    for _ in range(max_iter):
        cc = c.detach().requires_grad_(True)
        aux = {}
        S = soft_potential(cc, pos, mt, offsets, tst_off, eps, aux)
        (grad,) = torch.autograd.grad(-S.sum(), cc)   # = -acc(c)
        w = aux["wsum"]
        step = torch.where(w[:, None] > 0, -grad / w[:, None], torch.zeros_like(grad))
        step = torch.where(active[:, None], step, torch.zeros_like(step))
        c = c + step
        step_len = torch.where(active, step.norm(dim=1), step_len)
        iters += active.to(torch.int32)
        active = active & (step_len > tol * eps)
        if not bool(active.any()):
            break
    S = soft_potential(c, pos, mt, offsets, tst_off, eps).detach()
    clear_cache()  # the columns of this call's source array
    return CentringResult(c.detach(), S, iters, step_len)
