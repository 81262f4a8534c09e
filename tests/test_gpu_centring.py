"""Softened potential, its hand-written gradient kernel behind torch.autograd, and the
potential-minimum centring, against the FP64 oracle on the GPU.

Bars: S within the FP32 bound of the FP64 oracle; the gradient and w within the bound
relative to the sum of |terms|; autograd gradients (tests, source positions, source
masses, and the halo potential energy w.r.t. positions) against central finite
differences of the FP64 oracle and against the FP64 acceleration; refined centres are
stationary in FP64, at least as deep as the most-bound particle (brute force over every
particle), and the deepest point of an FP64 grid around them."""
import numpy as np
import pytest
import torch

import oracle

pytestmark = pytest.mark.gpu

EPS = 5e-4


def _haloes(n_haloes=3, seed=5, nscale=0.01):
    from pyhalma_amd import synth

    return synth.hernquist_haloes(seed=seed, n_haloes=n_haloes, scale=1.0, nscale=nscale)


def _terms_abs(m, x, y, z, t, eps):
    """sum_j |m_j d_k (r^2+eps^2)^-3/2| per test and axis (FP64, numpy), and w."""
    d = np.stack([x, y, z], 1).astype(np.float64)[None, :, :] - t[:, None, :]
    keep = np.all(d != 0.0, axis=2)
    r2 = (d * d).sum(2) + eps * eps
    w = np.where(keep, m.astype(np.float64)[None, :] * r2 ** -1.5, 0.0)
    return (np.abs(d) * w[:, :, None]).sum(1), w.sum(1)


def test_soft_sums_vs_oracle(gpu):
    from pyhalma_amd import engine

    w = _haloes()
    off = w["offsets"]
    rng = np.random.default_rng(1)
    H = len(off) - 1
    tests, toff = [], [0]
    for h in range(H):
        c = w["centres"][h]
        tests.append(c + rng.normal(scale=w["a"][h], size=(40, 3)))
        toff.append(toff[-1] + 40)
    t = np.concatenate(tests)
    toff = np.array(toff)
    r = engine.pot_soft_batched(off, w["mass"], w["x"], w["y"], w["z"], toff, t[:, 0], t[:, 1],
                                t[:, 2], EPS, pot=True, acc=True, wsum=True)
    pot, acc, ws = (r[k].cpu().numpy() for k in ("pot", "acc", "wsum"))
    B = engine.pot_error_bound()
    for h in range(H):
        s, e = off[h], off[h + 1]
        a, b = toff[h], toff[h + 1]
        args = (w["mass"][s:e], w["x"][s:e], w["y"][s:e], w["z"][s:e])
        So = oracle.pot_soft_f64(*args, t[a:b, 0], t[a:b, 1], t[a:b, 2], EPS)
        Ao = oracle.acc_soft_f64(*args, t[a:b, 0], t[a:b, 1], t[a:b, 2], EPS)
        mag, wo = _terms_abs(*args, t[a:b], EPS)
        assert np.all(np.abs(pot[a:b] - So) <= B * So), h
        assert np.all(np.abs(acc[a:b] - Ao) <= B * mag + 1e-300), h
        assert np.all(np.abs(ws[a:b] - wo) <= B * wo), h
    print(f"haloes={H} members={off[-1]} tests={len(t)}", flush=True)


def test_autograd_test_gradient_vs_finite_differences(gpu):
    from pyhalma_amd import centring, engine

    w = _haloes(n_haloes=2, seed=9)
    off = w["offsets"]
    H = len(off) - 1
    rng = np.random.default_rng(2)
    t0 = np.concatenate([w["centres"][h] + rng.normal(scale=w["a"][h], size=(8, 3))
                         for h in range(H)])
    toff = np.arange(0, 8 * H + 1, 8)
    dev = torch.device("cuda", 0)
    pos = torch.stack([torch.from_numpy(w[k]).to(dev) for k in "xyz"], 1)
    m = torch.from_numpy(w["mass"]).to(dev)
    coef = torch.from_numpy(rng.uniform(0.5, 2.0, len(t0))).to(dev)
    t = torch.from_numpy(t0).to(dev).requires_grad_(True)
    S = centring.soft_potential(t, pos, m, off, toff, EPS)
    (S * coef).sum().backward()
    g = t.grad.cpu().numpy()
    hstep = 1e-3 * EPS
    fd = np.zeros_like(t0)
    for h in range(H):
        s, e = off[h], off[h + 1]
        a, b = toff[h], toff[h + 1]
        args = (w["mass"][s:e], w["x"][s:e], w["y"][s:e], w["z"][s:e])
        for k in range(3):
            tp, tm = t0[a:b].copy(), t0[a:b].copy()
            tp[:, k] += hstep
            tm[:, k] -= hstep
            fd[a:b, k] = (oracle.pot_soft_f64(*args, tp[:, 0], tp[:, 1], tp[:, 2], EPS)
                          - oracle.pot_soft_f64(*args, tm[:, 0], tm[:, 1], tm[:, 2], EPS)) / (2 * hstep)
        Ao = oracle.acc_soft_f64(*args, t0[a:b, 0], t0[a:b, 1], t0[a:b, 2], EPS)
        mag, _ = _terms_abs(*args, t0[a:b], EPS)
        c = coef.cpu().numpy()[a:b, None]
        assert np.all(np.abs(g[a:b] - c * Ao) <= c * engine.pot_error_bound() * mag)
    np.testing.assert_allclose(g, coef.cpu().numpy()[:, None] * fd, rtol=1e-4,
                               atol=1e-4 * np.abs(g).max())


def test_autograd_source_gradients_vs_finite_differences(gpu):
    from pyhalma_amd import centring

    rng = np.random.default_rng(3)
    n, nt = 60, 12
    xs = rng.normal(scale=2e-3, size=(n, 3)).astype(np.float32)
    ms = rng.uniform(1e4, 1e6, n).astype(np.float32)
    t0 = rng.normal(scale=2e-3, size=(nt, 3))
    coef = rng.uniform(0.5, 2.0, nt)
    dev = torch.device("cuda", 0)
    pos = torch.from_numpy(xs).to(dev).requires_grad_(True)
    m = torch.from_numpy(ms).to(dev).requires_grad_(True)
    t = torch.from_numpy(t0).to(dev)
    S = centring.soft_potential(t, pos, m, [0, n], [0, nt], EPS)
    (S * torch.from_numpy(coef).to(dev)).sum().backward()
    gp = pos.grad.double().cpu().numpy()
    gm = m.grad.double().cpu().numpy()

    def L(xx, mm):
        return float((coef * oracle.pot_soft_f64(mm, xx[:, 0], xx[:, 1], xx[:, 2], t0[:, 0],
                                                 t0[:, 1], t0[:, 2], EPS)).sum())
    # source positions are FP32 in the kernel: differentiate the oracle in FP64 at the
    # same FP32 points by an exact analytic route (masses: S is linear in m_j)
    for j in range(n):
        e = np.zeros(n, np.float32)
        e[j] = 1.0
        d = np.stack([xs[j].astype(np.float64)] * nt) - t0
        keep = np.all(d != 0.0, axis=1)
        r2 = (d * d).sum(1) + EPS * EPS
        want_m = float((coef * np.where(keep, r2 ** -0.5, 0.0)).sum())
        assert abs(gm[j] - want_m) <= 2e-6 * abs(want_m)
        want_p = -(coef[:, None] * np.where(keep[:, None], float(ms[j]) * d * r2[:, None] ** -1.5,
                                            0.0)).sum(0)
        mag = (coef[:, None] * np.abs(float(ms[j]) * d * r2[:, None] ** -1.5)).sum(0)
        assert np.all(np.abs(gp[j] - want_p) <= 2e-6 * mag + 1e-30), j
    # and the mass gradient is the derivative of the oracle's L
    j = 7
    h = 1.0
    mp, mm_ = ms.astype(np.float64).copy(), ms.astype(np.float64).copy()
    mp[j] += h
    mm_[j] -= h
    fd = (L(xs, mp.astype(np.float32)) - L(xs, mm_.astype(np.float32))) / (
        float(np.float32(mp[j])) - float(np.float32(mm_[j])))
    assert abs(gm[j] - fd) <= 1e-6 * abs(fd)


def test_potential_energy_gradient_vs_fp64_acceleration(gpu):
    from pyhalma_amd import centring

    w = _haloes(n_haloes=2, seed=11, nscale=0.004)
    off = w["offsets"]
    dev = torch.device("cuda", 0)
    pos = torch.stack([torch.from_numpy(w[k]).to(dev) for k in "xyz"], 1).double()
    pos.requires_grad_(True)
    m = torch.from_numpy(w["mass"]).to(dev)
    W = centring.potential_energy(pos, m, off, EPS)
    W.sum().backward()
    g = pos.grad.cpu().numpy()
    # the backward that recomputes the acceleration gives the same bits as the fused one
    pos2 = pos.detach().clone().requires_grad_(True)
    W2 = centring.potential_energy(pos2, m, off, EPS, fuse=False)
    W2.sum().backward()
    assert torch.equal(W, W2) and torch.equal(pos.grad, pos2.grad)
    G = centring.G_CONST
    for h in range(len(off) - 1):
        s, e = off[h], off[h + 1]
        args = (w["mass"][s:e], w["x"][s:e], w["y"][s:e], w["z"][s:e])
        t = np.stack([w[k][s:e].astype(np.float64) for k in "xyz"], 1)
        Ao = oracle.acc_soft_f64(*args, t[:, 0], t[:, 1], t[:, 2], EPS)
        mag, _ = _terms_abs(*args, t, EPS)
        want = -G * w["mass"][s:e, None].astype(np.float64) * Ao
        tol = 4e-6 * G * w["mass"][s:e, None] * mag
        assert np.all(np.abs(g[s:e] - want) <= tol), h
    print(f"haloes={len(off) - 1} members={off[-1]}", flush=True)


def test_refine_centres_reaches_fp64_minimum(gpu):
    from pyhalma_amd import centring

    w = _haloes(n_haloes=4, seed=13, nscale=0.02)
    off = w["offsets"]
    res = centring.refine_centres(w["x"], w["y"], w["z"], w["mass"], off, w["seeds"], EPS,
                                  max_iter=300, tol=1e-5)
    c = res.centres.cpu().numpy()
    it = res.iters.cpu().numpy()
    print(f"haloes={len(off) - 1} iters={it.tolist()}", flush=True)
    for h in range(len(off) - 1):
        s, e = off[h], off[h + 1]
        args = (w["mass"][s:e], w["x"][s:e], w["y"][s:e], w["z"][s:e])
        a = w["a"][h]
        assert np.linalg.norm(c[h] - w["centres"][h]) < 0.1 * a
        # stationary in FP64
        Ao = oracle.acc_soft_f64(*args, c[h:h + 1, 0], c[h:h + 1, 1], c[h:h + 1, 2], EPS)[0]
        _, wo = _terms_abs(*args, c[h:h + 1], EPS)
        assert np.linalg.norm(Ao / wo[0]) < 1e-3 * EPS
        Sc = oracle.pot_soft_f64(*args, c[h:h + 1, 0], c[h:h + 1, 1], c[h:h + 1, 2], EPS)[0]
        # at least as deep as the most-bound particle (FP64 brute force over all of them)
        Sp = oracle.pot_soft_f64(*args, *(w[k][s:e].astype(np.float64) for k in "xyz"), EPS)
        assert Sc >= Sp.max()
        # the deepest point of a 7^3 FP64 grid spanning +-EPS/4 around the centre
        g1 = np.linspace(-EPS / 4, EPS / 4, 7)
        gx, gy, gz = np.meshgrid(g1, g1, g1, indexing="ij")
        pts = c[h] + np.stack([gx.ravel(), gy.ravel(), gz.ravel()], 1)
        Sg = oracle.pot_soft_f64(*args, pts[:, 0], pts[:, 1], pts[:, 2], EPS)
        assert Sc >= Sg.max() * (1 - 1e-12)


def test_soft_potential_source_cache_follows_changes(gpu):
    """The FP32 source columns kept between calls are reused only for the unchanged
    array: after an in-place change of the positions, and for a different array, the sums
    equal those of a fresh call."""
    import torch

    from pyhalma_amd import centring

    rng = np.random.default_rng(41)
    n = 3000
    pos = torch.tensor(rng.normal(scale=0.01, size=(n, 3)), dtype=torch.float32,
                       device="cuda")
    m = torch.tensor(rng.uniform(1e4, 1e6, n), dtype=torch.float32, device="cuda")
    off = np.array([0, 1000, n])
    tests = torch.tensor(rng.normal(scale=0.005, size=(2, 3)), dtype=torch.float64,
                         device="cuda")
    t_off = np.array([0, 1, 2])
    eps = 5e-4
    a = centring.soft_potential(tests, pos, m, off, t_off, eps)
    pos[7, 0] += 0.003  # in place: the version counter moves
    b = centring.soft_potential(tests, pos, m, off, t_off, eps)
    centring.clear_cache()
    fresh = centring.soft_potential(tests, pos, m, off, t_off, eps)
    assert torch.equal(b, fresh) and not torch.equal(a, b)
    other = pos.clone()
    other[2000, 1] -= 0.002
    c = centring.soft_potential(tests, other, m, off, t_off, eps)
    centring.clear_cache()
    assert torch.equal(c, centring.soft_potential(tests, other, m, off, t_off, eps))
    centring.clear_cache()


def test_soft_packed_sources_match_separate_arrays(gpu):
    """halma_pot_soft_batched_packed on (x, y, z, m) rows gives the same sums, gradients
    and weights, bit for bit, as halma_pot_soft_batched on the separate arrays (with an
    offset first source, as a sub-range of a larger catalogue)."""
    import torch

    from pyhalma_amd import engine

    rng = np.random.default_rng(43)
    n = 5000
    P = rng.normal(scale=0.01, size=(n, 3)).astype(np.float32)
    m = rng.uniform(1e4, 1e6, n).astype(np.float32)
    src_off = np.array([100, 1800, 4700])
    T = rng.normal(scale=0.004, size=(5, 3))
    tst_off = np.array([0, 2, 5])
    eps = 5e-4
    sep = engine.pot_soft_batched(src_off, m, P[:, 0], P[:, 1], P[:, 2], tst_off, T[:, 0],
                                  T[:, 1], T[:, 2], eps, pot=True, acc=True, wsum=True)
    packed = torch.tensor(np.concatenate([P, m[:, None]], 1), device="cuda")
    pk = engine.pot_soft_batched(src_off, None, None, None, None, tst_off, T[:, 0], T[:, 1],
                                 T[:, 2], eps, pot=True, acc=True, wsum=True, packed=packed)
    for k in ("pot", "acc", "wsum"):
        assert torch.equal(sep[k], pk[k]), k


def test_potential_energy_fp64_positions_drop_self_pair(gpu):
    """Positions given in FP64 that FP32 cannot hold: the energy uses the FP32 values for
    every pair (tests and sources alike), so each member's own pair is dropped and W
    equals the FP64 oracle sum over j != i at the FP32 points, and the gradient is the
    FP64 acceleration at the same points."""
    from pyhalma_amd import centring

    rng = np.random.default_rng(51)
    n = 1500
    p64 = rng.normal(scale=0.01, size=(n, 3)) + 0.123456789012345  # not FP32 values
    assert np.any(p64.astype(np.float32).astype(np.float64) != p64)
    m = rng.uniform(1e4, 1e6, n).astype(np.float32)
    dev = torch.device("cuda", 0)
    pos = torch.from_numpy(p64).to(dev).requires_grad_(True)
    mt = torch.from_numpy(m).to(dev)
    W = centring.potential_energy(pos, mt, [0, n], EPS)
    W.sum().backward()
    p32 = p64.astype(np.float32)
    S = oracle.pot_soft_f64(m, p32[:, 0], p32[:, 1], p32[:, 2], *(p32[:, k].astype(np.float64)
                                                                 for k in range(3)), EPS)
    G = centring.G_CONST
    Wo = -0.5 * G * float(np.sum(m.astype(np.float64) * S))
    B = engine_bound()
    assert abs(W.item() - Wo) <= B * abs(Wo), (W.item(), Wo)
    t = p32.astype(np.float64)
    Ao = oracle.acc_soft_f64(m, p32[:, 0], p32[:, 1], p32[:, 2], t[:, 0], t[:, 1], t[:, 2], EPS)
    mag, _ = _terms_abs(m, p32[:, 0], p32[:, 1], p32[:, 2], t, EPS)
    want = -G * m[:, None].astype(np.float64) * Ao
    g = pos.grad.cpu().numpy()
    assert np.all(np.abs(g - want) <= B * G * m[:, None] * mag + 1e-300)
    print(f"n={n} W={W.item():.6e} oracle={Wo:.6e}", flush=True)


def engine_bound():
    from pyhalma_amd import engine

    return engine.pot_error_bound()


def test_soft_sums_independent_of_other_haloes_and_merge_path(gpu):
    """A halo's softened sums are the same bits whether it is alone in a call (fewer than
    16384 tests: a wavefront merges each test) or shares it with a larger halo (a thread
    per test), and the large call's sums stay within the bound of the FP64 oracle."""
    from pyhalma_amd import engine

    rng = np.random.default_rng(52)
    na, nb = 3000, 17000
    P = rng.normal(scale=0.01, size=(na + nb, 3)).astype(np.float32)
    m = rng.uniform(1e4, 1e6, na + nb).astype(np.float32)
    T = P.astype(np.float64) + rng.normal(scale=1e-4, size=P.shape)
    one = engine.pot_soft_batched([0, na], m, P[:, 0], P[:, 1], P[:, 2], [0, na], T[:na, 0],
                                  T[:na, 1], T[:na, 2], EPS, pot=True, acc=True, wsum=True)
    both = engine.pot_soft_batched([0, na, na + nb], m, P[:, 0], P[:, 1], P[:, 2],
                                   [0, na, na + nb], T[:, 0], T[:, 1], T[:, 2], EPS, pot=True,
                                   acc=True, wsum=True)
    for k in ("pot", "acc", "wsum"):
        assert torch.equal(one[k], both[k][:na]), k
    B = engine.pot_error_bound()
    pot = both["pot"].cpu().numpy()[na:]
    So = oracle.pot_soft_f64(m[na:], P[na:, 0], P[na:, 1], P[na:, 2], T[na:, 0], T[na:, 1],
                             T[na:, 2], EPS)
    assert np.all(np.abs(pot - So) <= B * So)
    # the pot-only and gradient-only calls give the same bits as the combined one
    po = engine.pot_soft_batched([0, na], m, P[:, 0], P[:, 1], P[:, 2], [0, na], T[:na, 0],
                                 T[:na, 1], T[:na, 2], EPS, pot=True)
    ao = engine.pot_soft_batched([0, na], m, P[:, 0], P[:, 1], P[:, 2], [0, na], T[:na, 0],
                                 T[:na, 1], T[:na, 2], EPS, pot=False, acc=True, wsum=True)
    assert torch.equal(po["pot"], one["pot"])
    assert torch.equal(ao["acc"], one["acc"]) and torch.equal(ao["wsum"], one["wsum"])


def test_soft_few_tests_lanes_over_sources(gpu):
    """Haloes with one to eight tests (the wavefront spreads the sources over its lanes)
    and one with nine (64 tests per wavefront), against the FP64 oracle, with a source
    count that leaves a ragged last chunk."""
    from pyhalma_amd import engine

    rng = np.random.default_rng(53)
    ns = [9000, 4097, 1, 0, 12345]
    nt = [1, 8, 3, 2, 9]
    so = np.concatenate([[0], np.cumsum(ns)])
    to = np.concatenate([[0], np.cumsum(nt)])
    P = rng.normal(scale=0.01, size=(so[-1], 3)).astype(np.float32)
    m = rng.uniform(1e4, 1e6, so[-1]).astype(np.float32)
    T = rng.normal(scale=0.01, size=(to[-1], 3))
    T[0] = P[5]  # a test on a source: that pair is dropped
    r = engine.pot_soft_batched(so, m, P[:, 0], P[:, 1], P[:, 2], to, T[:, 0], T[:, 1],
                                T[:, 2], EPS, pot=True, acc=True, wsum=True)
    pot, acc, ws = (r[k].cpu().numpy() for k in ("pot", "acc", "wsum"))
    B = engine.pot_error_bound()
    for h in range(len(ns)):
        a, b = to[h], to[h + 1]
        s, e = so[h], so[h + 1]
        if e == s:
            assert np.all(pot[a:b] == 0) and np.all(acc[a:b] == 0)
            continue
        args = (m[s:e], P[s:e, 0], P[s:e, 1], P[s:e, 2])
        So = oracle.pot_soft_f64(*args, T[a:b, 0], T[a:b, 1], T[a:b, 2], EPS)
        Ao = oracle.acc_soft_f64(*args, T[a:b, 0], T[a:b, 1], T[a:b, 2], EPS)
        mag, wo = _terms_abs(*args, T[a:b], EPS)
        assert np.all(np.abs(pot[a:b] - So) <= B * So + 1e-300), h
        assert np.all(np.abs(acc[a:b] - Ao) <= B * mag + 1e-300), h
        assert np.all(np.abs(ws[a:b] - wo) <= B * wo + 1e-300), h
