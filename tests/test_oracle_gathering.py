"""Oracle of the source gathering (halo_gas.py:9-277), pinned by cells and spheres whose
contents follow by hand from the reference's text."""
import numpy as np

from oracle import gathering


def _patch(nx, ny, nz, rng):
    shape = (nx, ny, nz)
    return (rng.uniform(-0.5, 3.0, shape), rng.normal(size=shape), rng.normal(size=shape),
            rng.normal(size=shape), rng.uniform(1e4, 1e6, shape),
            np.ones(shape, bool), np.ones(shape, bool))


def test_patch_cells_by_hand():
    rng = np.random.default_rng(1)
    d, vx, vy, vz, t, cr, sol = _patch(2, 2, 2, rng)
    cr[1, 0, 0] = False   # refined: no particle
    sol[0, 1, 1] = False  # overlapped: no particle
    # res 1, first cell centre at rx - 0.5 = 0: centres 0 and 1 per axis; the open box
    # (-0.5, 0.5) x (-1, 2) x (-1, 2) keeps ix = 0 only
    out = gathering.patch_to_particles(1.0, 0.5, 0.5, 0.5, 2, 2, 2, d, cr, sol, vx, vy, vz, t,
                                       np.array([-0.5, 0.5, -1, 2, -1, 2]), 2.0)
    x, y, z, ovx, ovy, ovz, m, ot = out
    assert x.tolist() == [0.0, 0.0, 0.0] and y.tolist() == [0.0, 0.0, 1.0]
    assert z.tolist() == [0.0, 1.0, 0.0]  # (0, 1, 1) overlapped, iz fastest
    assert m.tolist() == [(1 + d[0, 0, 0]) * 2.0, (1 + d[0, 0, 1]) * 2.0, (1 + d[0, 1, 0]) * 2.0]
    assert ovx.tolist() == [vx[0, 0, 0], vx[0, 0, 1], vx[0, 1, 0]]
    # a centre exactly on the box edge is outside (strict inequalities)
    out = gathering.patch_to_particles(1.0, 0.5, 0.5, 0.5, 2, 2, 2, d, cr, sol, vx, vy, vz, t,
                                       np.array([0.0, 2.0, -1, 2, -1, 2]), 2.0)
    assert out[0].tolist() == [1.0, 1.0, 1.0] and out[1].tolist() == [0.0, 1.0, 1.0]


def test_amr_levels_and_units():
    rng = np.random.default_rng(2)
    gas = [[], [], [], [], [], [], []]
    for shape in ((2, 2, 2), (2, 2, 2), (4, 2, 2)):
        for lst, a in zip(gas, _patch(*shape, rng)):
            lst.append(a)
    # patch 0 is level 0 (skipped), patch 1 level 1 (res L/ncoarse/2 = 1), patch 2 level 2
    out = gathering.amr_to_particles(8.0, 4, [0, 1, 2], [0.5, 0.5, 0.25], [0.5, 0.5, 0.25],
                                     [0.5, 0.5, 0.25], gas, 10.0, 0.0, 0.0, 0.0, 3.0)
    x, vx, m = out[0], out[3], out[6]
    assert x.size == 8 + 16
    assert x[:8].tolist() == [0, 0, 0, 0, 1, 1, 1, 1]
    assert x[8:].tolist() == [0.0] * 4 + [0.5] * 4 + [1.0] * 4 + [1.5] * 4
    assert vx[0] == gas[1][1][0, 0, 0] * 3e5
    assert m[8] == (1 + gas[0][2][0, 0, 0]) * 3.0 * 0.125


def test_sphere_queries():
    rng = np.random.default_rng(3)
    p = rng.uniform(-1, 1, size=(3000, 3))
    for R in (0.1, 0.5, 1.3):
        a = gathering.ball(p[:, 0], p[:, 1], p[:, 2], 0.1, -0.2, 0.3, R)
        b = gathering.ball_brute(p[:, 0], p[:, 1], p[:, 2], 0.1, -0.2, 0.3, R)
        assert a.tolist() == b.tolist() and a.size > 0
    # integer points at distance exactly 5 (3-4-5) are inside (<= R), strict for the gas
    q = np.array([[3.0, 4.0, 0.0], [0.0, 0.0, 5.0], [5.0, 0.1, 0.0]])
    assert gathering.ball(q[:, 0], q[:, 1], q[:, 2], 0, 0, 0, 5.0).tolist() == [0, 1]
    assert gathering.inside(q[:, 0], q[:, 1], q[:, 2], 5.0, 0, 0, 0).tolist() == [False] * 3
