"""Symmetric self-potential kernel against the FP64 oracle: every member within the
rigorous bound, on haloes of every tile/block shape (1 member, a partial tile, exact
tiles and blocks, several blocks), with coincident coordinates (the skip rule on both
sides of a pair), and against the unsymmetric batched kernel."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def test_pot_sym_shapes_vs_oracle(gpu):
    from pyhalma_amd import engine

    rng = np.random.default_rng(17)
    sizes = [1, 2, 37, 128, 129, 511, 512, 513, 1024, 2047, 3000]
    parts = []
    for n in sizes:
        p = rng.normal(scale=0.01, size=(n, 3)).astype(np.float32)
        if n > 10:  # members sharing one coordinate value with another member
            p[: n // 10, 0] = p[n // 10: 2 * (n // 10), 0]
        parts.append(p)
    P = np.concatenate(parts)
    m = rng.uniform(1e4, 1e6, len(P)).astype(np.float32)
    off = np.concatenate([[0], np.cumsum(sizes)])
    got = engine.pot_sum_self(off, m, P[:, 0], P[:, 1], P[:, 2]).cpu().numpy()
    B = engine.pot_sym_error_bound()
    for h, n in enumerate(sizes):
        a, b = off[h], off[h + 1]
        want = oracle.pot_sum_f64(m[a:b], P[a:b, 0], P[a:b, 1], P[a:b, 2], P[a:b, 0],
                                  P[a:b, 1], P[a:b, 2])
        err = np.abs(got[a:b] - want)
        assert np.all(err <= B * want + 1e-300), (n, float((err / np.maximum(want, 1e-300)).max()))
    print(f"haloes={len(sizes)} members={len(P)} bound={B:.2e}", flush=True)


def test_pot_sym_matches_batched_on_clumps(gpu):
    from pyhalma_amd import engine, synth

    w = synth.clumps(seed=8, scale=0.01)
    lab = oracle.fof_labels_c(w["x"], w["y"], w["z"], w["ll"])
    g = engine.groups_from_labels(lab.astype(np.int32), 50)
    md = engine.gather_members(g, w["x"], w["y"], w["z"], w["mass"])
    off = md.groups.offsets_host
    x, y, z, m = (t.cpu().numpy() for t in (md.x, md.y, md.z, md.m))
    s_sym = engine.pot_sum_self(off, m, x, y, z).cpu().numpy()
    s_bat = engine.pot_sum_batched(off, m, x, y, z, off, x, y, z).cpu().numpy()
    rel = np.abs(s_sym - s_bat) / s_bat
    assert rel.max() <= engine.pot_sym_error_bound() + engine.pot_error_bound()
    h = int(np.argmax(np.diff(off)))
    a, b = off[h], off[h + 1]
    want = oracle.pot_sum_f64(m[a:b], x[a:b], y[a:b], z[a:b], x[a:b], y[a:b], z[a:b])
    assert np.all(np.abs(s_sym[a:b] - want) <= engine.pot_sym_error_bound() * want)


def _max_chain(parts):
    worst = 0
    for p in parts:
        for ax in range(3):
            _, c = np.unique(p[:, ax], return_counts=True)
            worst = max(worst, int(c.max()))
    return worst


def _colliding_pairs(p):
    """Pairs of one halo with an equal coordinate (counted once)."""
    n = len(p)
    eq = np.zeros((n, n), bool)
    for ax in range(3):
        eq |= p[:, ax][:, None] == p[:, ax][None, :]
    return int((np.triu(eq, 1)).sum())


def test_pot_sym_collisions_vs_oracle(gpu):
    """Dense equal coordinates (tens of thousands of pairs the reference's rule skips,
    spread over every tile pair) and coincident members: every member within the bound
    of the oracle, which applies the rule to every pair."""
    from pyhalma_amd import engine

    rng = np.random.default_rng(23)
    sizes = [300, 700, 1500, 2600]
    parts = []
    for n in sizes:
        p = rng.normal(scale=0.01, size=(n, 3))
        p = (np.round(p / 1e-4) * 1e-4).astype(np.float32)
        p[:20] = p[20:40]  # coincident members
        parts.append(p)
    assert _max_chain(parts) > 5
    n_coll = sum(_colliding_pairs(p) for p in parts)
    assert n_coll > 10000
    P = np.concatenate(parts)
    m = rng.uniform(1e4, 1e6, len(P)).astype(np.float32)
    off = np.concatenate([[0], np.cumsum(sizes)])
    got = engine.pot_sum_self(off, m, P[:, 0], P[:, 1], P[:, 2]).cpu().numpy()
    B = engine.pot_sym_error_bound()
    for h in range(len(sizes)):
        a, b = off[h], off[h + 1]
        want = oracle.pot_sum_f64(m[a:b], P[a:b, 0], P[a:b, 1], P[a:b, 2], P[a:b, 0],
                                  P[a:b, 1], P[a:b, 2])
        assert np.all(np.isfinite(got[a:b]))
        assert np.all(np.abs(got[a:b] - want) <= B * want), h
    print(f"colliding pairs={n_coll}", flush=True)


def test_pot_sym_shared_plane_vs_oracle(gpu):
    """A halo whose members all share one z value: the rule skips every pair of it (S = 0
    exactly), next to an ordinary halo in the same call."""
    from pyhalma_amd import engine

    rng = np.random.default_rng(29)
    n = 900
    p = rng.normal(scale=0.01, size=(n, 3)).astype(np.float32)
    p[:, 2] = np.float32(0.003)
    q = rng.normal(scale=0.01, size=(500, 3)).astype(np.float32)
    P = np.concatenate([p, q])
    m = rng.uniform(1e4, 1e6, len(P)).astype(np.float32)
    off = np.array([0, n, n + 500])
    got = engine.pot_sum_self(off, m, P[:, 0], P[:, 1], P[:, 2]).cpu().numpy()
    B = engine.pot_sym_error_bound()
    want0 = oracle.pot_sum_f64(m[:n], *(P[:n, k] for k in range(3)), *(P[:n, k] for k in range(3)))
    # every pair of the first halo shares z: the reference's rule skips them all
    assert np.all(want0 == 0.0) and np.all(got[:n] == 0.0)
    want1 = oracle.pot_sum_f64(m[n:], *(P[n:, k] for k in range(3)), *(P[n:, k] for k in range(3)))
    assert np.all(np.abs(got[n:] - want1) <= B * want1)


def test_pair_kernels_nan_coordinates(gpu):
    """A member with a NaN coordinate: the reference's /= is true for NaN, so its pairs
    count and every sum of its halo is NaN, also when all three coordinates are NaN (no
    difference is zero).  The symmetric kernel and the chunked batched kernel give NaN
    exactly where the oracle does, and the clean halo of the same call within the bound."""
    from pyhalma_amd import engine

    rng = np.random.default_rng(31)
    sizes = [700, 600, 500]
    P = rng.normal(scale=0.01, size=(sum(sizes), 3)).astype(np.float32)
    off = np.concatenate([[0], np.cumsum(sizes)])
    P[off[1] + 5, 0] = np.nan           # one coordinate of a member of halo 1
    P[off[2] + 9, :] = np.nan           # all three of a member of halo 2
    m = rng.uniform(1e4, 1e6, len(P)).astype(np.float32)
    want = np.concatenate([
        oracle.pot_sum_f64(m[a:b], *(P[a:b, k] for k in range(3)), *(P[a:b, k] for k in range(3)))
        for a, b in zip(off[:-1], off[1:])])
    assert np.all(np.isfinite(want[:off[1]])) and np.all(np.isnan(want[off[1]:]))
    sym = engine.pot_sum_self(off, m, P[:, 0], P[:, 1], P[:, 2]).cpu().numpy()
    bat = engine.pot_sum_batched(off, m, P[:, 0], P[:, 1], P[:, 2], off, P[:, 0], P[:, 1],
                                 P[:, 2])
    bat = bat.cpu().numpy() if hasattr(bat, "cpu") else np.asarray(bat)
    for got, bound in ((sym, engine.pot_sym_error_bound()), (bat, engine.pot_error_bound())):
        assert np.array_equal(np.isnan(got), np.isnan(want))
        a, b = off[0], off[1]
        assert np.all(np.abs(got[a:b] - want[a:b]) <= bound * want[a:b])


def test_pot_sym_banded_large_haloes(gpu):
    """Haloes of more than 64 blocks of 512 members run their block pairs tile by tile
    (bands of 64 blocks, partial sums merged between tiles): sampled members of haloes of
    40,000 (two bands), 70,000 (three bands) and 32,768 members (exactly one band) are
    within the bound of the FP64 oracle, and the small haloes of the call have the same
    bits as in a call of their own."""
    from pyhalma_amd import engine, synth

    rng = np.random.default_rng(23)
    sizes = [40000, 500, 70000, 3000, 32768]
    parts = [synth.plummer_positions(rng, n, 0.01).astype(np.float32) for n in sizes]
    P = np.concatenate(parts)
    m = rng.uniform(1e4, 1e6, len(P)).astype(np.float32)
    off = np.concatenate([[0], np.cumsum(sizes)])
    got = engine.pot_sum_self(off, m, P[:, 0], P[:, 1], P[:, 2]).cpu().numpy()
    B = engine.pot_sym_error_bound()
    for h, n in enumerate(sizes):
        a, b = off[h], off[h + 1]
        pick = np.sort(rng.choice(n, min(n, 1500), replace=False)) + a
        want = oracle.pot_sum_f64(m[a:b], P[a:b, 0], P[a:b, 1], P[a:b, 2], P[pick, 0],
                                  P[pick, 1], P[pick, 2])
        err = np.abs(got[pick] - want) / want
        print(f"n={n} max rel err {err.max():.2e}", flush=True)
        assert err.max() <= B, n
    # halo 1 (500) and halo 3 (3000) alone: the same bits
    for h in (1, 3):
        a, b = off[h], off[h + 1]
        alone = engine.pot_sum_self([0, b - a], m[a:b], P[a:b, 0], P[a:b, 1],
                                    P[a:b, 2]).cpu().numpy()
        assert np.array_equal(alone, got[a:b])


def test_pot_sym_forms_same_bits(gpu):
    """A halo's sums do not depend on how its launch is split: a 3000-member halo alone
    (21 block pairs, each taken by a 1024-thread workgroup, one tile pair per wavefront)
    and beside a 23,040-member halo (1056 block pairs, each taken by a 256-thread
    workgroup) give the same bits; the large halo is within the bound of the FP64 oracle
    on a sample of its members."""
    from pyhalma_amd import engine, synth

    rng = np.random.default_rng(29)
    sizes = [23040, 3000]
    parts = [synth.plummer_positions(rng, n, 0.01).astype(np.float32) for n in sizes]
    P = np.concatenate(parts)
    m = rng.uniform(1e4, 1e6, len(P)).astype(np.float32)
    off = np.concatenate([[0], np.cumsum(sizes)])
    got = engine.pot_sum_self(off, m, P[:, 0], P[:, 1], P[:, 2]).cpu().numpy()
    for h in (0, 1):
        a, b = off[h], off[h + 1]
        alone = engine.pot_sum_self([0, b - a], m[a:b], P[a:b, 0], P[a:b, 1],
                                    P[a:b, 2]).cpu().numpy()
        print(f"halo {h} ({b - a} members): alone == joint", flush=True)
        assert np.array_equal(alone, got[a:b])
    pick = np.sort(rng.choice(sizes[0], 1000, replace=False))
    want = oracle.pot_sum_f64(m[:sizes[0]], P[:sizes[0], 0], P[:sizes[0], 1], P[:sizes[0], 2],
                              P[pick, 0], P[pick, 1], P[pick, 2])
    assert np.max(np.abs(got[pick] - want) / want) <= engine.pot_sym_error_bound()


def test_pair_kernels_equal_infinite_coordinates(gpu):
    """Two members whose x is the same infinity.  The reference's x_j /= x_i is false for
    them, so it drops their pair; the kernels test the smallest |difference| against zero
    and inf - inf is NaN, so they count it and those two members' sums are NaN (a
    documented divergence: infinite coordinates are outside parity).  Every other member
    of the halo is unaffected: its terms with the infinite members are m / inf = 0, as in
    the reference, and its sum is within the bound of the oracle's."""
    from pyhalma_amd import engine

    rng = np.random.default_rng(37)
    n = 600
    P = rng.normal(scale=0.01, size=(n, 3)).astype(np.float32)
    P[10, 0] = P[20, 0] = np.inf
    m = rng.uniform(1e4, 1e6, n).astype(np.float32)
    want = oracle.pot_sum_f64(m, *(P[:, k] for k in range(3)), *(P[:, k] for k in range(3)))
    off = np.array([0, n])
    sym = engine.pot_sum_self(off, m, P[:, 0], P[:, 1], P[:, 2]).cpu().numpy()
    finite = np.ones(n, bool)
    finite[[10, 20]] = False
    assert np.all(np.isfinite(want))
    assert np.all(np.isnan(sym[~finite]))
    assert np.all(np.abs(sym[finite] - want[finite]) <= engine.pot_sym_error_bound() * want[finite])
