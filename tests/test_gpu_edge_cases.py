"""Empty and degenerate inputs of the widened rows' entry points: no groups, empty groups,
no points, no patches, zero radii -- the results the reference gives for them."""
import numpy as np
import pytest
import torch

import oracle

pytestmark = pytest.mark.gpu


def _groups(lists):
    from pyhalma_amd import engine
    off = np.concatenate([[0], np.cumsum([len(g) for g in lists])]).astype(np.int64)
    mem = (np.concatenate(lists) if off[-1] else np.zeros(1)).astype(np.int32)
    return engine.Groups(torch.from_numpy(mem).cuda(), torch.from_numpy(off).cuda(), off)


def test_cleaning_degenerate_groups(gpu):
    from pyhalma_amd import engine
    rng = np.random.default_rng(1)
    p = rng.uniform(0, 1, size=(300, 3))
    v = rng.normal(size=(300, 3))
    m = np.ones(300)
    args = (p[:, 0], p[:, 1], p[:, 2], v[:, 0], v[:, 1], v[:, 2], m)
    # no groups
    res = engine.phase_space_clean(_groups([]), *args, 0.1, 1.0, 3.0)
    assert res.kept.n_groups == 0 and len(res.n_cell) == 0
    # an empty group, a one-member group, two members, and a normal one
    lists = [np.zeros(0, np.int64), np.array([5]), np.array([7, 9]), np.arange(100, 300)]
    res = engine.phase_space_clean(_groups(lists), *args, 0.1, 1.0, 3.0, min_part=0, minp=1)
    per, kept = oracle.phase_space_clean(lists, *args, 0.1, 1.0, 3.0, min_part=0, minp=1)
    for g, want in enumerate(per):
        a, b = int(res.offsets_host[g]), int(res.offsets_host[g + 1])
        assert res.control_clean.cpu().numpy()[a:b].tolist() == want["control_clean"].tolist()
        assert res.rounds[g] == want["rounds"] and res.n_cell[g] == want["n_cell"]
    assert res.kept.n_groups == len(kept)


def test_progenitors_and_queries_without_data(gpu):
    from pyhalma_amd import engine, gathering
    # new haloes that are all empty
    g = _groups([np.zeros(0, np.int64), np.zeros(0, np.int64)])
    out = engine.progenitors(g, np.zeros(5, np.int64), [1.0, 2.0], [2], np.array([1, 2]), [3.0])
    assert all((a == 0).all() for a in out)
    # no new haloes at all
    out = engine.progenitors(_groups([]), np.zeros(5, np.int64), [], [2], np.array([1, 2]), [3.0])
    assert all(a.size == 0 for a in out)
    # sphere queries: no spheres; zero and negative radii
    x = np.random.default_rng(2).uniform(0, 1, (100, 3))
    off, idx, _ = engine.ball_query(x[:, 0], x[:, 1], x[:, 2], np.zeros((0, 3)), [])
    assert off.tolist() == [0] and idx.numel() == 0
    off, idx, _ = engine.ball_query(x[:, 0], x[:, 1], x[:, 2], [[0.5, 0.5, 0.5]] * 2,
                                    [0.0, -1.0])
    assert off.tolist() == [0, 0, 0]
    # AMR: no refined patches (only the base grid)
    cells = [[np.zeros((2, 2, 2), np.float32)] for _ in range(5)] + \
        [[np.ones((2, 2, 2), bool)] for _ in range(2)]
    grid = gathering.AmrGrid(8.0, 4, [0], [2], [2], [2], [0.0], [0.0], [0.0], cells)
    off, gas = gathering.amr_particles(grid, 1.0, [[-9, 9, -9, 9, -9, 9]])
    assert off.tolist() == [0, 0] and gas["x"].numel() == 0


def test_half_mass_radius_and_shapes_single_members(gpu):
    from pyhalma_amd import engine
    x = np.array([1.0, 2.0, 3.0])
    zero = np.zeros(3)
    m = np.ones(3)
    g = _groups([np.array([0]), np.array([1, 2])])
    r = engine.half_mass_radius(g, x, zero, zero, m).cpu().numpy()
    want = [oracle.half_mass_radius(*oracle.center_of_mass(mem, x, zero, zero, m), mem, x, zero,
                                    zero, m) for mem in g.to_arrays()]
    assert r.tolist() == want
    s = engine.halo_shapes(np.array([0, 1]), np.array([0.0], np.float32),
                           np.array([0.0], np.float32), np.array([0.0], np.float32),
                           np.array([1.0], np.float32)).cpu().numpy()
    assert s.tolist() == [[0.0, 0.0, 0.0]]
