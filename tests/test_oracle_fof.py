"""Pins the friends-of-friends oracles (CPU only).

* the BFS restatement of fof.py:18-79, scipy's connected components, an O(N^2) brute
  force with the FP64 predicate written out, and the C cell-list oracle agree on random,
  clustered, duplicated and tie-heavy inputs;
* the golden fixture (scipy labels) holds exact-ll ties both ways of the rounding of
  ((dx*dx + dy*dy) + dz*dz) <= ll*ll;
* invariants: labels are minimum member indices, groups come out sorted and ordered by
  minimum index, singletons are excluded, permutation equivariance.
"""
import numpy as np
import pytest

import oracle


def _rand_cases():
    rng = np.random.default_rng(42)
    out = []
    for n, ll in ((1, 0.1), (2, 0.1), (50, 0.15), (300, 0.08), (700, 0.05)):
        out.append((rng.uniform(0, 1, (n, 3)), ll))
    c = np.concatenate([rng.normal(0, 0.01, (200, 3)), rng.normal(0.3, 0.02, (150, 3)),
                        rng.uniform(-0.5, 0.5, (100, 3))])
    out.append((c, 0.01))
    d = rng.uniform(0, 1, (100, 3))
    out.append((np.concatenate([d, d[:40], d[:10]]), 0.05))  # duplicates
    g = np.arange(12) * 0.25
    X, Y, Z = np.meshgrid(g, g, g[:3], indexing="ij")
    out.append((np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1), 0.25))  # lattice at ll
    return out


@pytest.mark.parametrize("k", range(8))
def test_oracles_agree(k):
    data, ll = _rand_cases()[k]
    x, y, z = data[:, 0], data[:, 1], data[:, 2]
    brute = oracle.fof_labels_brute(x, y, z, ll)
    sp = oracle.fof_labels_scipy(x, y, z, ll)
    cc = oracle.fof_labels_c(x, y, z, ll)
    cells = oracle.fof_labels_cells(x, y, z, ll)
    assert np.array_equal(brute, sp)
    assert np.array_equal(brute, cc)
    assert np.array_equal(brute, cells)
    assert oracle.fof_bfs(data, ll) == oracle.labels_to_groups(brute)


def test_golden_fixture(golden_dir):
    g = np.load(f"{golden_dir}/fof_ref.npz")
    for name in ("clumps_f64", "clumps_f32", "ties"):
        x = g[f"{name}_x"]
        ll = float(g[f"{name}_ll"])
        want = g[f"{name}_labels"]
        for fn in (oracle.fof_labels_brute, oracle.fof_labels_c, oracle.fof_labels_scipy,
                   oracle.fof_labels_cells):
            assert np.array_equal(fn(x[:, 0], x[:, 1], x[:, 2], ll), want), (name, fn)
        assert oracle.fof_bfs(x.astype(np.float64), ll) == oracle.labels_to_groups(want)


def test_tie_rounding_both_ways():
    ll = 0.5
    # (1.3-1.0)^2 + 0.4^2 rounds above 0.25; (7.3-7.0)^2 + 0.4^2 rounds below
    p = np.array([[1.0, 0, 0], [1.3, 0.4, 0], [7.0, 0, 0], [7.3, 0.4, 0],
                  [3.0, 3.0, 3.0], [3.0, 3.5, 3.0]])
    lab = oracle.fof_labels_brute(p[:, 0], p[:, 1], p[:, 2], ll)
    assert lab.tolist() == [0, 1, 2, 2, 4, 4]
    assert np.array_equal(oracle.fof_labels_c(p[:, 0], p[:, 1], p[:, 2], ll), lab)
    assert np.array_equal(oracle.fof_labels_cells(p[:, 0], p[:, 1], p[:, 2], ll), lab)


def test_cells_oracle_on_dense_clumps():
    """The sub-cell oracle (cells of side ll/sqrt(3) linked whole) against the plain
    cell-list oracle on dense clumps (thousands of particles per sub-cell), pairs of
    clumps a little more and a little less than ll apart, lattices at spacings ll and
    ll/sqrt(3) (points on sub-cell faces), and a wide box that needs the fallback."""
    rng = np.random.default_rng(17)
    ll = 0.01
    dense = np.concatenate([rng.normal(c, 0.002, (3000, 3))
                            for c in rng.uniform(-0.3, 0.3, (6, 3))])
    pair = np.concatenate([rng.normal(0, 0.0004, (200, 3)),
                           rng.normal(0, 0.0004, (200, 3)) + [0.0105, 0, 0],
                           rng.normal(0, 0.0004, (200, 3)) + [0.5, 0, 0],
                           rng.normal(0, 0.0004, (200, 3)) + [0.5 + 0.0125, 0, 0]])
    g = np.arange(9) * ll
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    lat = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1) + 3.0
    g2 = np.arange(9) * (ll / np.sqrt(3.0))
    X, Y, Z = np.meshgrid(g2, g2, g2, indexing="ij")
    lat2 = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1) - 3.0
    wide = np.concatenate([rng.uniform(0, 1, (300, 3)), [[1e5, 0, 0], [1e5 + ll, 0, 0]]])
    for data in (dense, pair, lat, lat2, np.concatenate([dense, pair, lat, lat2]), wide):
        data = data.astype(np.float32).astype(np.float64)
        x, y, z = data[:, 0], data[:, 1], data[:, 2]
        assert np.array_equal(oracle.fof_labels_cells(x, y, z, ll),
                              oracle.fof_labels_c(x, y, z, ll))


def test_groups_order_and_permutation():
    rng = np.random.default_rng(9)
    data = np.concatenate([rng.normal(c, 0.02, (40, 3)) for c in (0.0, 0.5, 1.0)]
                          + [rng.uniform(-1, 2, (60, 3))])
    ll = 0.03
    lab = oracle.fof_labels_c(data[:, 0], data[:, 1], data[:, 2], ll)
    groups = oracle.labels_to_groups(lab)
    assert all(g == sorted(g) and len(g) >= 2 for g in groups)
    assert [g[0] for g in groups] == sorted(g[0] for g in groups)
    assert all(lab[g[0]] == g[0] for g in groups)
    perm = rng.permutation(len(data))
    lab_p = oracle.fof_labels_c(data[perm, 0], data[perm, 1], data[perm, 2], ll)
    # same partition: particles i, j share a label iff perm-images do
    inv = np.empty_like(perm)
    inv[perm] = np.arange(len(perm))
    same = lab[:, None] == lab[None, :]
    same_p = lab_p[inv][:, None] == lab_p[inv][None, :]
    assert np.array_equal(same, same_p)


def test_parallel_restatement_exact_when_buffer_suffices():
    rng = np.random.default_rng(4)
    cen = rng.uniform(-4, 4, (20, 3))
    data = np.concatenate([rng.normal(c, 0.005, (60, 3)) for c in cen])
    m = np.ones(len(data))
    ll = 0.004
    par = oracle.fof_parallel(data[:, 0], data[:, 1], data[:, 2], ll, 10.0, 20, m, 8)
    lab = oracle.fof_labels_c(data[:, 0], data[:, 1], data[:, 2], ll)
    ser = [g for g in oracle.labels_to_groups(lab) if len(g) > 20]
    assert sorted(map(tuple, (sorted(g.tolist()) for g in par))) == sorted(map(tuple, ser))
