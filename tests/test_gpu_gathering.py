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


def _amr_snapshot(seed, n_patches=40, dtype=np.float32):
    """A base grid (level 0) and n_patches level 1-3 patches of 2-12 cells per axis, about
    20 % of the cells refined or overlapped."""
    rng = np.random.default_rng(seed)
    L, ncoarse = 16.0, 8
    levels = [0] + list(rng.integers(1, 4, n_patches))
    shapes = [(ncoarse, ncoarse, ncoarse)] + [tuple(rng.integers(2, 13, 3)) for _ in range(n_patches)]
    res = [L / ncoarse / 2 ** int(lev) for lev in levels]
    rx, ry, rz = (np.array([0.0] + list(rng.uniform(-6, 6, n_patches))) for _ in range(3))
    for p in range(1, n_patches + 1):  # the first parent cell's centre on the parent grid
        for a in (rx, ry, rz):
            a[p] = np.floor(a[p] / (2 * res[p])) * 2 * res[p] + res[p]
    cells = [[], [], [], [], [], [], []]
    for shp in shapes:
        cells[0].append(rng.uniform(-0.9, 5.0, shp).astype(dtype))
        for k in (1, 2, 3):
            cells[k].append(rng.normal(size=shp).astype(dtype) * 1e-3)
        cells[4].append(rng.uniform(1e3, 1e7, shp).astype(dtype))
        cells[5].append(rng.random(shp) > 0.2)
        cells[6].append(rng.random(shp) > 0.2)
    return L, ncoarse, np.array(levels), rx, ry, rz, cells


def test_amr_particles_vs_oracle(gpu):
    from pyhalma_amd import gathering as ga
    L, nc, levels, rx, ry, rz, cells = _amr_snapshot(11)
    grid = ga.AmrGrid(L, nc, levels, [c.shape[0] for c in cells[0]],
                      [c.shape[1] for c in cells[0]], [c.shape[2] for c in cells[0]],
                      rx, ry, rz, cells)
    rng = np.random.default_rng(12)
    c = rng.uniform(-5, 5, size=(30, 3))
    r = rng.uniform(0.2, 3.0, 30)
    r[0] = 0.0
    rete, rho_B = 0.37, 2.5
    boxes = np.stack([c[:, 0] - r, c[:, 0] + r, c[:, 1] - r, c[:, 1] + r, c[:, 2] - r,
                      c[:, 2] + r], 1)
    off, g = ga.amr_particles(grid, rho_B, boxes, c, r, mass_scale=rete ** 3)
    offb, gb = ga.amr_particles(grid, rho_B, boxes)  # AMRgrid_to_particles: box only
    names = ("x", "y", "z", "vx", "vy", "vz", "mass", "temp")
    got = {k: g[k].cpu().numpy() for k in names}
    gotb = {k: gb[k].cpu().numpy() for k in names}
    total = 0
    for h in range(len(r)):
        want = gathering.amr_to_particles(L, nc, levels, rx, ry, rz, cells, r[h], *c[h], rho_B)
        for q, k in enumerate(names):
            assert gotb[k][offb[h]:offb[h + 1]].tolist() == want[q].tolist(), (h, k)
        keep = gathering.inside(want[0], want[1], want[2], r[h], *c[h])
        for q, k in enumerate(names):
            w = want[q][keep] * (rete ** 3 if k == "mass" else 1.0)
            assert got[k][off[h]:off[h + 1]].tolist() == w.tolist(), (h, k)
        total += int(keep.sum())
    assert total > 1000 and off[1] == 0


def test_drop_in_source_gathering(gpu):
    from pyhalma_amd import gathering as ga
    from pyhalma_amd import halo_gas as hg
    L, nc, levels, rx, ry, rz, cells = _amr_snapshot(13)
    npatch = [1] + [int((levels == lev).sum()) for lev in (1, 2, 3)]
    order = np.argsort(levels, kind="stable")  # masclet numbers patches level by level
    levels, rx, ry, rz = levels[order], rx[order], ry[order], rz[order]
    cells = [[lst[i] for i in order] for lst in cells]
    shapes = [a.shape for a in cells[0]]
    grid_data = [None] * 15
    grid_data[5] = npatch
    for k, d in ((6, 0), (7, 1), (8, 2)):
        grid_data[k] = np.array([s[d] for s in shapes])
    grid_data[12], grid_data[13], grid_data[14] = rx, ry, rz
    gas_data = cells
    rng = np.random.default_rng(14)
    dm = [rng.uniform(-6, 6, 20000).astype(np.float32) for _ in range(3)] + \
        [rng.uniform(1e-12, 1e-11, 20000).astype(np.float32)]
    st = [rng.uniform(-6, 6, 5000).astype(np.float32) for _ in range(3)]
    st_data = st + [None, None, None, rng.uniform(1e-13, 1e-12, 5000).astype(np.float32),
                    None, None, rng.permutation(5000).astype(np.int32) + 7]
    for (cx, cy, cz, R) in ((0.3, -1.0, 2.0, 1.5), (4.0, 4.0, -4.0, 2.5), (0.0, 0.0, 0.0, 0.0)):
        got = hg.st_gas_dm_particles_inside(0.5, L, nc, grid_data, gas_data, dm, st_data, None,
                                            None, cx, cy, cz, R, 1.7)
        want = gathering.gather_sources(0.5, L, nc, levels, rx, ry, rz, gas_data, dm,
                                        (st[0], st[1], st[2], st_data[6], st_data[9]),
                                        cx, cy, cz, R, 1.7, ga.MASS_TO_SUN)
        assert len(got) == len(want) == 17
        for a, b in zip(got, want):
            assert np.asarray(a).tolist() == np.asarray(b).tolist()
        g1 = hg.AMRgrid_to_particles(L, nc, grid_data, gas_data, R, cx, cy, cz, 1.7)
        w1 = gathering.amr_to_particles(L, nc, levels, rx, ry, rz, gas_data, R, cx, cy, cz, 1.7)
        assert all(np.asarray(a).tolist() == np.asarray(b).tolist() for a, b in zip(g1, w1))
    # patch_to_particles of one patch
    p = 5
    res = L / nc / 2 ** int(levels[p])
    box = np.array([-3.0, 3.0, -2.0, 4.0, -5.0, 5.0])
    args = (res, rx[p], ry[p], rz[p], *shapes[p], cells[0][p], cells[5][p], cells[6][p],
            cells[1][p], cells[2][p], cells[3][p], cells[4][p], box, 1.7)
    got = hg.patch_to_particles(*args)
    want = gathering.patch_to_particles(*args)
    assert all(np.asarray(a).tolist() == np.asarray(b).tolist() for a, b in zip(got, want))


def test_dark_matter_mass(gpu):
    from pyhalma_amd import gathering as ga
    rng = np.random.default_rng(15)
    p = rng.uniform(-1, 1, size=(50000, 3)).astype(np.float32)
    m = rng.uniform(1e-12, 3e-12, 50000).astype(np.float32)
    c = rng.uniform(-1, 1, size=(40, 3))
    r = rng.uniform(0.0, 0.5, 40)
    got = ga.dark_matter_mass(p[:, 0], p[:, 1], p[:, 2], m, c, r)
    scaled = m * ga.MASS_TO_SUN  # float32, as pyHALMA.py:1176 scales them
    for h in range(40):
        idx = gathering.ball(p[:, 0], p[:, 1], p[:, 2], *c[h], r[h])
        exact = math.fsum(scaled[idx].astype(np.float64).tolist())
        assert abs(got[h] - exact) <= 1e-15 * exact, h           # correctly rounded sum
        ref = float(np.sum(scaled[idx]))                           # the reference's float32 sum
        assert abs(ref - exact) <= 1e-6 * exact + 1e-30, h
