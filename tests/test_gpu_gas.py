"""Batched gas energy test (RPS) against the FP64 oracle on the GPU.

Bars: bound masks identical to the FP64 decision (S from the oracle's fixed-order FP64
sum over the halo's gas, dark matter and stars; E = 1/2|v - v_h|^2 - 2 G S); S within
the FP32 bound; per-halo masses within 1e-12; a knife-edge set decided by the FP64
re-test; empty haloes and haloes without dark matter or stars."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def _workload(seed=3, n=10, nmax=12000):
    from pyhalma_amd import synth

    return synth.gas_haloes(seed=seed, n_haloes=n, nmin=300, nmax=nmax, dm_per_gas=0.5,
                            st_per_gas=0.2)


def _sources(w, h):
    parts = [(w["x"], w["y"], w["z"], w["mass"], w["offsets"])] + [w[k] for k in ("dm", "stars")
                                                                  if k in w]
    cols = [[], [], [], []]
    for p in parts:
        a, b = p[4][h], p[4][h + 1]
        for k in range(4):
            cols[k].append(np.asarray(p[k])[a:b])
    return [np.concatenate(c) for c in cols]


def _oracle_masks(w, halo_v, factor=2.0):
    off = w["offsets"]
    bound = np.zeros(off[-1], bool)
    S = np.zeros(off[-1])
    for h in range(len(off) - 1):
        a, b = off[h], off[h + 1]
        if b == a:
            continue
        sx, sy, sz, sm = _sources(w, h)
        S[a:b] = oracle.pot_sum_f64(sm, sx, sy, sz, w["x"][a:b], w["y"][a:b], w["z"][a:b])
        _, ok = oracle.energy_test(S[a:b], w["vx"][a:b], w["vy"][a:b], w["vz"][a:b],
                                   tuple(halo_v[h]), factor)
        bound[a:b] = ok
    return bound, S


def test_energy_test_batched_vs_oracle(gpu):
    from pyhalma_amd import engine
    from pyhalma_amd.halo_gas import _source_csr

    w = _workload()
    off = w["offsets"]
    H = len(off) - 1
    groups = [(w["x"], w["y"], w["z"], w["mass"], off), w["dm"], w["stars"]]
    src_off, (sx, sy, sz, sm) = _source_csr(groups, H)
    bound, S, retested = engine.energy_test_batched(
        src_off, sm, sx, sy, sz, off, w["x"], w["y"], w["z"], w["vx"], w["vy"], w["vz"],
        w["bulk"], 2.0, with_pot=True)
    want, S64 = _oracle_masks(w, w["bulk"])
    got = bound.cpu().numpy().astype(bool)
    print(f"haloes={H} cells={off[-1]} bound={int(want.sum())} retested={retested}",
          flush=True)
    assert np.array_equal(got, want), int((got != want).sum())
    rel = np.abs(S.cpu().numpy() - S64) / S64
    assert rel.max() <= engine.pot_error_bound()


def test_rps_batched_vs_oracle(gpu):
    from pyhalma_amd import halo_gas

    w = _workload(seed=4, n=8)
    off = w["offsets"]
    gas = (w["x"], w["y"], w["z"], w["vx"], w["vy"], w["vz"], w["mass"], w["temp"])
    out, bound = halo_gas.RPS_batched(gas, off, w["bulk"], dm=w["dm"], stars=w["stars"])
    want, _ = _oracle_masks(w, w["bulk"])
    assert np.array_equal(bound.cpu().numpy().astype(bool), want)
    m = w["mass"].astype(np.float64)
    cold = w["temp"] < 5e4
    for h in range(len(off) - 1):
        a, b = off[h], off[h + 1]
        tot = m[a:b].sum()
        cb = m[a:b][cold[a:b] & want[a:b]].sum()
        uc = m[a:b][cold[a:b] & ~want[a:b]].sum()
        uh = m[a:b][~cold[a:b] & ~want[a:b]].sum()
        np.testing.assert_allclose(out[h], [tot, cb / tot, uc, uh], rtol=1e-12, atol=1e-9)
    # the per-halo drop-in (FP32 sums, reference rounding) agrees up to cells whose E is
    # within FP32 rounding of zero: compare its masses to a few cells' worth
    for h in range(3):
        a, b = off[h], off[h + 1]
        d0, d1 = w["dm"][4][h], w["dm"][4][h + 1]
        s0, s1 = w["stars"][4][h], w["stars"][4][h + 1]
        ref = halo_gas.RPS(*(g[a:b] for g in gas), *(np.asarray(c)[d0:d1] for c in w["dm"][:4]),
                           *(np.asarray(c)[s0:s1] for c in w["stars"][:4]), *w["bulk"][h],
                           10 ** 9, 5e5, 1)
        assert abs(ref[0] - out[h][0]) <= 1e-6 * out[h][0]
        assert abs(ref[2] - out[h][2]) <= 3 * 1.5e5 and abs(ref[3] - out[h][3]) <= 3 * 1.5e5


def test_energy_test_knife_edge_and_edges(gpu):
    from pyhalma_amd import engine

    rng = np.random.default_rng(5)
    # halo 0: speeds set so that E sits within 1e-9 relative of zero; halo 1: no tests;
    # halo 2: tests but no sources (E = kin >= 0: bound only at rest)
    n = 2000
    p = rng.normal(scale=0.01, size=(n, 3)).astype(np.float32)
    m = np.full(n, 1e6, np.float32)
    S = oracle.pot_sum_f64(m, p[:, 0], p[:, 1], p[:, 2], p[:, 0], p[:, 1], p[:, 2])
    target = 2.0 * oracle.G_CONST * S * (1 + rng.uniform(-1e-9, 1e-9, n))
    u = rng.normal(size=(n, 3))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    v = u * np.sqrt(2 * target)[:, None]
    t2 = rng.normal(size=(3, 3)).astype(np.float32)
    tx = np.concatenate([p[:, 0], t2[:, 0]])
    ty = np.concatenate([p[:, 1], t2[:, 1]])
    tz = np.concatenate([p[:, 2], t2[:, 2]])
    vv = np.concatenate([v, np.array([[0.0, 0, 0], [1.0, 0, 0], [0, 0, 0]])])
    src_off = np.array([0, n, n, n])
    tst_off = np.array([0, n, n, n + 3])
    bound, _, retested = engine.energy_test_batched(src_off, m, p[:, 0], p[:, 1], p[:, 2],
                                                    tst_off, tx, ty, tz, vv[:, 0], vv[:, 1],
                                                    vv[:, 2], np.zeros((3, 3)), 2.0)
    got = bound.cpu().numpy().astype(bool)
    _, want = oracle.energy_test(S, v[:, 0], v[:, 1], v[:, 2], (0.0, 0.0, 0.0), 2.0)
    print(f"retested={retested} bound={int(want.sum())} of {n}", flush=True)
    assert retested > n // 2
    assert 0 < want.sum() < n
    assert np.array_equal(got[:n], want)
    assert got[n:].tolist() == [True, False, True]


def test_energy_test_self_form_vs_oracle(gpu):
    """The self form (the gas cells are their own first sources: symmetric kernel for the
    gas pairs, chunked kernel for dark matter and stars) decides every cell as the FP64
    oracle does, its S is within the bound of the FP64 sum, and haloes without extra
    sources work."""
    from pyhalma_amd import engine
    from pyhalma_amd.halo_gas import _source_csr

    w = _workload(seed=6, n=8, nmax=40000)
    off = w["offsets"]
    H = len(off) - 1
    ext_off, (ex, ey, ez, em) = _source_csr([w["dm"], w["stars"]], H)
    bound, S, retested = engine.energy_test_self_batched(
        off, w["mass"], w["x"], w["y"], w["z"], w["vx"], w["vy"], w["vz"], w["bulk"], 2.0,
        ext_off=ext_off, em=em, ex=ex, ey=ey, ez=ez, with_pot=True)
    want, S64 = _oracle_masks(w, w["bulk"])
    got = bound.cpu().numpy().astype(bool)
    print(f"haloes={H} cells={off[-1]} largest={np.diff(off).max()} retested={retested}",
          flush=True)
    assert np.array_equal(got, want), int((got != want).sum())
    rel = np.abs(S.cpu().numpy() - S64) / S64
    assert rel.max() <= max(engine.pot_error_bound(), engine.pot_sym_error_bound())
    # no extra sources: the gas alone, against the plain form
    b2, S2, _ = engine.energy_test_self_batched(off, w["mass"], w["x"], w["y"], w["z"], w["vx"],
                                                w["vy"], w["vz"], w["bulk"], 2.0, with_pot=True)
    b3, S3, _ = engine.energy_test_batched(off, w["mass"], w["x"], w["y"], w["z"], off, w["x"],
                                           w["y"], w["z"], w["vx"], w["vy"], w["vz"], w["bulk"],
                                           2.0, with_pot=True)
    assert torch_equal(b2, b3)
    rel = np.abs(S2.cpu().numpy() - S3.cpu().numpy()) / S3.cpu().numpy()
    assert rel.max() <= engine.pot_error_bound() + engine.pot_sym_error_bound()


def torch_equal(a, b):
    import torch

    return bool(torch.equal(a, b))
