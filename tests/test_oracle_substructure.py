"""Oracle of the substructure splitting (pyHALMA.py:703-766) and of half_mass_radius
(halo_properties.py:366-392), pinned by hand-built cases whose answers follow from the
reference's text alone."""
import numpy as np

import oracle


def _axis_points(dists):
    """Points on the x axis at the given distances from the origin (exact distances)."""
    x = np.asarray(dists, np.float64)
    return x, np.zeros_like(x), np.zeros_like(x)


def test_half_mass_radius_closed_forms():
    x, y, z = _axis_points([1.0, 2.0, 3.0, 4.0])
    idx = np.arange(4)
    # equal masses, M = 4: running mass 1, 2 (not > 2), 3 > 2 -> the member before: 2
    assert oracle.half_mass_radius(0, 0, 0, 4.0, idx, x, y, z, np.ones(4)) == 2.0
    # the first member alone holds half the mass (3 of 6): 3 is not > 3, then 4 > 3
    assert oracle.half_mass_radius(0, 0, 0, 6.0, idx, x, y, z, np.array([3.0, 1, 1, 1])) == 1.0
    # the first member exceeds half the mass at once: the smallest distance
    assert oracle.half_mass_radius(0, 0, 0, 8.0, idx, x, y, z, np.array([5.0, 1, 1, 1])) == 1.0
    # M larger than the members' mass: never exceeded, the largest distance
    assert oracle.half_mass_radius(0, 0, 0, 100.0, idx, x, y, z, np.ones(4)) == 4.0
    # the order is by distance, not by member index
    xs, ys, zs = _axis_points([4.0, 1.0, 3.0, 2.0])
    assert oracle.half_mass_radius(0, 0, 0, 4.0, idx, xs, ys, zs, np.ones(4)) == 2.0
    # 3-4-5 distances about a centre off the origin
    px = np.array([3.0, 0.0]) + 1.0
    py = np.array([4.0, 0.0])
    pz = np.array([0.0, 5.0])
    assert oracle.half_mass_radius(1.0, 0, 0, 2.0, [0, 1], px, py, pz, np.ones(2)) == 5.0


def _blob(rng, c, n, r):
    return c + rng.uniform(-r, r, size=(n, 3))


def _two_blobs_and_bridge(seed=1, n=400, nbridge=9):
    """Two dense cubes (half-side 0.6) 10 apart joined by a line of points 0.9 apart that
    starts and ends 0.8 from the cube faces: linked at ll = 1, not at ll = 0.5 (the bridge
    breaks, each cube stays whole)."""
    rng = np.random.default_rng(seed)
    a = _blob(rng, np.array([0.0, 0.0, 0.0]), n, 0.6)
    b = _blob(rng, np.array([10.0, 0.0, 0.0]), n, 0.6)
    t = np.linspace(1.4, 8.6, nbridge)
    bridge = np.stack([t, np.zeros_like(t), np.zeros_like(t)], 1)
    p = np.concatenate([a, b, bridge])
    return p[:, 0], p[:, 1], p[:, 2], n


def test_split_two_blobs_joined_by_a_bridge():
    x, y, z, n = _two_blobs_and_bridge()
    m = np.ones(len(x))
    g = oracle.labels_to_groups(oracle.fof_labels_c(x, y, z, 1.0), 2)
    assert len(g) == 1 and len(g[0]) == len(x)  # one FoF group at ll = 1
    out, count = oracle.split_substructure(g, x, y, z, m, 1.0, minp_sub=25, dec_factor=2.0,
                                           min_ll=0.3, dist_factor=0.75, minp=50)
    # at ll = 0.5 the cubes (n each: 49 % of the group, CM 5 from the centre, half-mass
    # radius about 5) split off; the 9 bridge points left are under 30 %: the group is
    # emptied; at ll = 0.25 nothing is left.  Output: the two cubes, in discovery order.
    assert count == 2
    assert [sorted(o) for o in out] == [list(range(n)), list(range(n, 2 * n))]


def test_no_split_for_a_single_blob():
    rng = np.random.default_rng(2)
    p = _blob(rng, np.zeros(3), 500, 0.6)
    x, y, z = p[:, 0], p[:, 1], p[:, 2]
    g = [list(range(500))]
    out, count = oracle.split_substructure(g, x, y, z, np.ones(500), 1.0, 25, 2.0, 0.3, 0.75, 50)
    # at ll = 0.5 the blob (mean spacing 0.14) stays whole: its one sub-group holds 100 %
    # of the members (not < 60 %), so nothing splits
    assert count == 0 and out == g


def test_split_keeps_group_above_thirty_percent():
    # a big cube with a small cube far off joined by a bridge: the small one (10 %) splits,
    # the big one keeps 87 % of its members and stays first in the list
    rng = np.random.default_rng(3)
    big = _blob(rng, np.zeros(3), 900, 1.2)
    small = _blob(rng, np.array([12.0, 0.0, 0.0]), 100, 0.3)
    t = np.linspace(1.75, 11.1, 12)  # 0.55 from the big cube's face, 0.6 from the small one
    bridge = np.stack([t, np.zeros_like(t), np.zeros_like(t)], 1)
    p = np.concatenate([big, small, bridge])
    x, y, z = p[:, 0], p[:, 1], p[:, 2]
    g = oracle.labels_to_groups(oracle.fof_labels_c(x, y, z, 1.0), 2)
    assert len(g) == 1
    out, count = oracle.split_substructure(g, x, y, z, np.ones(len(x)), 1.0, 25, 2.0, 0.3,
                                           0.75, 50)
    assert count == 1
    assert sorted(out[1]) == list(range(900, 1000))
    assert set(out[0]) == set(range(len(x))) - set(range(900, 1000))
