"""Batched unbinding on the GPU against the FP64 oracle.

Bars: bound masks and per-halo pass counts identical to oracle.unbind_iterative; v_cm
within 1e-14; the FP32 potentials within the rigorous bound; a knife-edge set where
almost every particle sits inside the FP32 error band (so the FP64 re-test decides) still
matches the oracle bit for bit; graph replay and plain launches agree; the plain FP64
comparator agrees with the oracle."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def _groups(w, min_size=50):
    from pyhalma_amd import engine

    lab = oracle.fof_labels_c(w["x"], w["y"], w["z"], w["ll"])
    return engine.groups_from_labels(lab.astype(np.int32), min_size)


def _oracle(g, w, factor, max_iter=100):
    return oracle.unbind_iterative(g.offsets_host, g.members[: g.n_members].cpu().numpy(),
                                   w["x"], w["y"], w["z"], w["vx"], w["vy"], w["vz"],
                                   w["mass"], factor, max_iter)


@pytest.mark.parametrize("factor", [1.0, 9.0])
def test_unbind_matches_oracle(gpu, factor):
    from pyhalma_amd import engine

    w = __import__("pyhalma_amd.synth", fromlist=["x"]).clumps(seed=12, scale=0.01)
    g = _groups(w)
    print(f"factor={factor} haloes={g.n_groups} members={g.n_members}", flush=True)
    r = engine.unbind(g, w["x"], w["y"], w["z"], w["vx"], w["vy"], w["vz"], w["mass"],
                      factor=factor, with_pot=True)
    bound, iters, S = _oracle(g, w, factor)
    got = r.bound.cpu().numpy().astype(bool)
    assert np.array_equal(got, bound), int((got != bound).sum())
    assert np.array_equal(r.iters.cpu().numpy(), iters)
    pot = r.pot.cpu().numpy()
    ok = ~np.isnan(S)
    rel = np.abs(pot[ok] - S[ok]) / S[ok]
    # every S the unbinding keeps is within 1e-6 of the FP64 direct sum: incremental
    # passes subtract the removed members' sums, and a halo whose accumulated bound would
    # pass 1e-6 is recomputed in full
    assert rel.max() <= 1e-6
    print(f"passes={r.stats['passes']} retested={r.stats['retested']} "
          f"resummed={r.stats['resummed']} unbound={int((~bound).sum())} "
          f"max_rel_pot_err={rel.max():.2e}", flush=True)


def test_unbind_equal_coordinates(gpu):
    """Members sharing a coordinate value with another member of their halo (and
    coincident members), whose pairs the reference's rule skips: masks and pass counts
    equal the oracle's."""
    from pyhalma_amd import engine, synth

    w = synth.clumps(seed=14, scale=0.01)
    g = _groups(w)
    rng = np.random.default_rng(14)
    w = dict(w)
    for k in ("x", "y", "z"):
        w[k] = w[k].copy()
    off = g.offsets_host
    mem = g.members[: g.n_members].cpu().numpy()
    n_set = 0
    for h in range(g.n_groups):
        idx = mem[off[h]:off[h + 1]]
        n = len(idx)
        src = rng.choice(idx, size=n // 8)
        dst = rng.choice(idx, size=n // 8)
        ax = rng.integers(0, 3, size=n // 8)
        for s, d, a in zip(src, dst, ax):
            k = "xyz"[a]
            w[k][d] = w[k][s]
        n_set += n // 8
        for k in ("x", "y", "z"):  # two coincident members
            w[k][idx[1]] = w[k][idx[0]]
    r = engine.unbind(g, w["x"], w["y"], w["z"], w["vx"], w["vy"], w["vz"], w["mass"], 1.0)
    bound, iters, _ = _oracle(g, w, 1.0)
    got = r.bound.cpu().numpy().astype(bool)
    print(f"haloes={g.n_groups} members={g.n_members} copied coordinates={n_set} "
          f"unbound={int((~bound).sum())}", flush=True)
    assert np.array_equal(got, bound), int((got != bound).sum())
    assert np.array_equal(r.iters.cpu().numpy(), iters)


@pytest.mark.parametrize("limit", [0, 700])
def test_unbind_full_passes_on_chunked_kernel(gpu, limit):
    """Haloes above the symmetric-kernel limit (all of them at 0, the larger ones at
    700 members) run their full passes on the chunked kernel; masks and pass counts are
    the oracle's either way."""
    from pyhalma_amd import _lib, engine, synth

    w = synth.clumps(seed=13, scale=0.01)
    g = _groups(w)
    sizes = np.diff(g.offsets_host)
    assert limit == 0 or (sizes.max() > limit and sizes.min() <= limit)
    old = engine.tune_get(_lib.TUNE_SYM_MAX)
    engine.tune_set(_lib.TUNE_SYM_MAX, limit)
    try:
        r = engine.unbind(g, w["x"], w["y"], w["z"], w["vx"], w["vy"], w["vz"], w["mass"],
                          1.0, time_pot=True)
    finally:
        engine.tune_set(_lib.TUNE_SYM_MAX, old)
    bound, iters, _ = _oracle(g, w, 1.0)
    got = r.bound.cpu().numpy().astype(bool)
    print(f"limit={limit} haloes={g.n_groups} sizes {sizes.min()}..{sizes.max()} "
          f"symmetric pairs={r.stats['full_pairs']} all pairs={r.stats['pairs']}", flush=True)
    assert np.array_equal(got, bound), int((got != bound).sum())
    assert np.array_equal(r.iters.cpu().numpy(), iters)
    if limit == 0:
        assert r.stats["full_pairs"] == 0
    else:
        assert 0 < r.stats["full_pairs"] < r.stats["pairs"]


def test_unbind_plummer_background(gpu):
    """The Plummer-plus-background set at its full size (20,000 + 4,000 particles, FP64):
    FoF labels on the GPU equal the oracle's, and the groups built from them unbind to
    the oracle's masks, pass counts and potentials (within 1e-6)."""
    from pyhalma_amd import engine, synth

    w = synth.plummer_with_background()
    assert len(w["x"]) == 24000 and w["x"].dtype == np.float64
    lab = engine.fof_labels(w["x"], w["y"], w["z"], w["ll"]).cpu().numpy()
    want = oracle.fof_labels_c(w["x"], w["y"], w["z"], w["ll"])
    assert np.array_equal(lab, want)
    g = engine.groups_from_labels(lab.astype(np.int32), 50)
    r = engine.unbind(g, w["x"], w["y"], w["z"], w["vx"], w["vy"], w["vz"], w["mass"], 1.0,
                      with_pot=True)
    bound, iters, S = _oracle(g, w, 1.0)
    print(f"groups={g.n_groups} members={g.n_members} bound={int(bound.sum())} "
          f"passes={iters.tolist()}", flush=True)
    assert np.array_equal(r.bound.cpu().numpy().astype(bool), bound)
    assert np.array_equal(r.iters.cpu().numpy(), iters)
    ok = ~np.isnan(S)
    assert np.max(np.abs(r.pot.cpu().numpy()[ok] - S[ok]) / S[ok]) <= 1e-6


def test_unbind_knife_edge_band(gpu):
    """Velocities set so that E = kin - G*S sits within ~1e-8 relative of zero for every
    particle: the FP32 band contains them all and the FP64 re-test must decide exactly as
    the oracle does.  Mirror-symmetric positions and velocities keep v_cm = 0."""
    from pyhalma_amd import engine, synth

    rng = np.random.default_rng(8)
    half = synth.plummer_positions(rng, 1500, 0.005)
    pos = np.concatenate([half, -half]).astype(np.float32).astype(np.float64)
    m = np.full(len(pos), 1e5)
    S = oracle.pot_sum_f64(m, pos[:, 0], pos[:, 1], pos[:, 2], pos[:, 0], pos[:, 1], pos[:, 2])
    target = oracle.G_CONST * S * (1.0 + rng.uniform(-3e-8, 3e-8, len(pos)))
    speed = np.sqrt(2.0 * target[:1500])
    u = rng.normal(size=(1500, 3))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    v = np.concatenate([u * speed[:, None], -u * speed[:, None]])
    # the mirrored half has its own S; set its speed from its own target instead
    sp2 = np.sqrt(2.0 * target[1500:])
    v[1500:] = -u * sp2[:, None]
    # re-balance momentum exactly: equal masses, v_cm must vanish -> scale pairs alike
    vv = v.copy()
    vv[1500:] = -vv[:1500]
    w = dict(x=pos[:, 0], y=pos[:, 1], z=pos[:, 2], vx=vv[:, 0], vy=vv[:, 1], vz=vv[:, 2],
             mass=m, ll=1.0)
    members = np.arange(len(pos), dtype=np.int32)
    import torch
    g = engine.Groups(torch.from_numpy(members).cuda(),
                      torch.tensor([0, len(pos)], dtype=torch.int64).cuda(),
                      np.array([0, len(pos)], np.int64))
    r = engine.unbind(g, w["x"], w["y"], w["z"], w["vx"], w["vy"], w["vz"], m, 1.0, max_iter=1)
    bound, _, _ = oracle.unbind_iterative(g.offsets_host, members, w["x"], w["y"], w["z"],
                                          w["vx"], w["vy"], w["vz"], m, 1.0, max_iter=1)
    got = r.bound.cpu().numpy().astype(bool)
    print(f"retested={r.stats['retested']} of {len(pos)}; bound={int(bound.sum())}", flush=True)
    assert r.stats["retested"] > len(pos) // 2
    assert 0 < bound.sum() < len(pos)
    assert np.array_equal(got, bound)


def test_unbind_graph_vs_launches_and_plain(gpu):
    from pyhalma_amd import engine, synth

    w = synth.clumps(seed=21, scale=0.005)
    g = _groups(w)
    args = (g, w["x"], w["y"], w["z"], w["vx"], w["vy"], w["vz"], w["mass"], 1.0)
    a = engine.unbind(*args, use_graph=True)
    b = engine.unbind(*args, use_graph=False)
    assert np.array_equal(a.bound.cpu().numpy(), b.bound.cpu().numpy())
    assert np.array_equal(a.iters.cpu().numpy(), b.iters.cpu().numpy())
    pb, piters = engine.unbind_plain_f64(*args)
    bound, iters, _ = _oracle(g, w, 1.0)
    diff = int((pb.cpu().numpy().astype(bool) != bound).sum())
    assert diff == 0, diff
    assert np.array_equal(piters, iters)


def test_unbind_empty_and_tiny(gpu):
    from pyhalma_amd import engine
    import torch

    e = engine.Groups(torch.zeros(1, dtype=torch.int32).cuda(),
                      torch.zeros(1, dtype=torch.int64).cuda(), np.zeros(1, np.int64))
    z = np.zeros(3)
    r = engine.unbind(e, z, z, z, z, z, z, z)
    assert r.bound.numel() == 0
    # two bodies at rest: bound; two bodies flying apart: unbound
    x = np.array([0.0, 0.001, 5.0, 5.001])
    y = np.array([0.0, 0.0005, 5.0, 5.0005])
    zz = np.array([0.0, 0.0002, 5.0, 5.0002])
    vx = np.array([0.0, 0.0, -1e4, 1e4])
    m = np.full(4, 1e8)
    members = torch.arange(4, dtype=torch.int32).cuda()
    g = engine.Groups(members, torch.tensor([0, 2, 4], dtype=torch.int64).cuda(),
                      np.array([0, 2, 4], np.int64))
    r = engine.unbind(g, x, y, zz, vx, np.zeros(4), np.zeros(4), m, 1.0)
    assert r.bound.cpu().numpy().tolist() == [1, 1, 0, 0]
