"""Source gathering (halo_gas.py:9-277, pyHALMA.py:1172-1185) on the GPU against the
oracle: sphere contents identical to scipy's KD-tree queries, gas particles identical to
the restated AMR conversion."""
import math

import numpy as np
import pytest

from oracle import gathering

pytestmark = pytest.mark.gpu


def _spheres(rng, H, lo, hi, rmax):
    c = rng.uniform(lo, hi, size=(H, 3))
    r = rng.uniform(0.0, rmax, H)
    r[:3] = [0.0, rmax * 3, 1e-9]  # empty, large, tiny
    c[3] = [hi + 5 * rmax, 0, 0]   # outside the points' box
    return c, r


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("cell", [0.0, 0.05, 1.0])
def test_ball_query_vs_kdtree(gpu, dtype, cell):
    from pyhalma_amd import engine
    rng = np.random.default_rng(5)
    p = np.concatenate([rng.normal(size=(20000, 3)) * 0.3, rng.uniform(-2, 2, (10000, 3))])
    p = p.astype(dtype)
    m = rng.uniform(0.5, 2.0, len(p))
    c, r = _spheres(rng, 200, -1.5, 1.5, 0.4)
    off, idx, sums = engine.ball_query(p[:, 0], p[:, 1], p[:, 2], c, r, m=m, cell=cell)
    idx = idx.cpu().numpy()
    for h in range(len(r)):
        want = gathering.ball(p[:, 0], p[:, 1], p[:, 2], *c[h], r[h])
        got = idx[off[h]:off[h + 1]]
        assert got.tolist() == want.tolist(), h
        ref = math.fsum(m[want].tolist())
        assert abs(sums[h] - ref) <= 1e-15 * abs(ref), h
    # strict predicate, unsorted output: same sets as parallel_inside
    off2, idx2, _ = engine.ball_query(p[:, 0], p[:, 1], p[:, 2], c, r, strict=True, cell=cell,
                                      sorted=False)
    idx2 = idx2.cpu().numpy()
    for h in range(len(r)):
        want = np.nonzero(gathering.inside(p[:, 0].astype(np.float64), p[:, 1].astype(np.float64),
                                           p[:, 2].astype(np.float64), r[h], *c[h]))[0]
        assert sorted(idx2[off2[h]:off2[h + 1]].tolist()) == want.tolist(), h


def test_ball_query_boundaries_and_overflow(gpu):
    from pyhalma_amd import engine
    # integer points at distance exactly 5 (3-4-5): inside for <=, outside for <
    q = np.array([[3.0, 4.0, 0.0], [0.0, 0.0, 5.0], [5.0, 0.1, 0.0], [0.0, 0.0, 0.0]])
    off, idx, _ = engine.ball_query(q[:, 0], q[:, 1], q[:, 2], [[0, 0, 0]], [5.0])
    assert idx.cpu().tolist() == [0, 1, 3]
    off, idx, _ = engine.ball_query(q[:, 0], q[:, 1], q[:, 2], [[0, 0, 0]], [5.0], strict=True)
    assert idx.cpu().tolist() == [3]
    # overlapping spheres: more hits than points (the output is sized again)
    rng = np.random.default_rng(6)
    p = rng.uniform(0, 1, size=(500, 3))
    off, idx, _ = engine.ball_query(p[:, 0], p[:, 1], p[:, 2], np.full((5, 3), 0.5), [2.0] * 5)
    assert off.tolist() == [0, 500, 1000, 1500, 2000, 2500]
    assert idx.cpu().numpy().reshape(5, 500).tolist() == [list(range(500))] * 5
    off, idx, _ = engine.ball_query(p[:, 0], p[:, 1], p[:, 2], np.full((5, 3), 0.5), [2.0] * 5,
                                    count_only=True)
    assert idx is None and off[-1] == 2500
    # no points
    off, idx, s = engine.ball_query(np.zeros(0), np.zeros(0), np.zeros(0), [[0, 0, 0]], [1.0],
                                    m=np.zeros(0))
    assert off.tolist() == [0, 0] and s.tolist() == [0.0]
