"""Oracle of the phase-space cleaning (halo_properties.py:113-278, pyHALMA.py:789-843),
pinned by constructed groups whose cell statistics and decisions follow by hand from
the reference's text."""
import math

import numpy as np

import oracle
from oracle import cleaning


def _clump(n, rng, scale=1e-3):
    return rng.uniform(-scale, scale, size=(n, 3))


def test_calc_cell_statistics_by_hand():
    # four members in one cell: masses 1, 1, 2, 4; x velocities 1, 3, 0, 0 (others 0)
    p = np.zeros((4, 3))
    p[:, 0] = [0.0, 0.01, 0.02, 0.03]
    v = np.zeros((4, 3))
    v[:, 0] = [1.0, 3.0, 0.0, 0.0]
    m = np.array([1.0, 1.0, 2.0, 4.0])
    cells, which = cleaning.calc_cell(0.0, 0.0, 0.0, np.array([-1.0, 1.0]) - 0.5, np.arange(4),
                                      p[:, 0], p[:, 1], p[:, 2], v[:, 0], v[:, 1], v[:, 2], m)
    # nodes -1.5 and 0.5: x = 0..0.03 and y = z = 0 are all nearest to 0.5 (index 1)
    assert (which == 1).all() and which.shape == (4, 3)
    (vc, mass, q, sig), = cells.values()
    assert mass == 8.0 and q == 4.0
    assert vc == [0.5, 0.0, 0.0]  # (1 + 3) / 8
    # deviations 0.5, 2.5, 0.5, 0.5 -> sum of squares 7 -> sqrt(7 / 3 / 4)
    assert sig == math.sqrt(7.0 / 3.0 / 4.0)


def test_outliers_leave_in_one_round():
    # 95 members at rest and 5 moving at a in one tiny clump: v_cell = 0.05 a,
    # sigma = sqrt((5 * 0.95^2 + 95 * 0.05^2) / 3 / 100) a = 0.1258 a: the movers score
    # 7.55 >= 2 * 3, the others 0.40; 5 removed of 100, 5 / 95 <= 0.1: one round
    rng = np.random.default_rng(1)
    p = _clump(100, rng)
    v = np.zeros((100, 3))
    v[[3, 17, 40, 41, 99], 0] = 200.0
    m = np.ones(100)
    per, kept = oracle.phase_space_clean([np.arange(100)], p[:, 0], p[:, 1], p[:, 2], v[:, 0],
                                         v[:, 1], v[:, 2], m, 1.0, 1.0, 3.0, min_part=50, minp=50)
    want = np.ones(100, bool)
    want[[3, 17, 40, 41, 99]] = False
    assert (per[0]["control"] == want).all() and per[0]["rounds"] == 1
    assert len(kept) == 1 and kept[0]["members"].tolist() == np.nonzero(want)[0].tolist()


def test_control_index_quirk_counts_down_to_zero():
    # member 0 sits alone in its cell (count 1 <= q_fil = 1): removed in the first round,
    # 1 / 9 > 0.1 so the loop runs.  Member 1 moves at a, the other 8 of the clump rest:
    # v_cell = a / 9, sigma = sqrt(72 / 81 / 27) a, member 1 scores (8 / 9) / sigma = 4.90:
    # kept at 2 * 3, removed at 1.5 * 3, 1.125 * 3 and 1 * 3.  In every later round it is
    # remaining member 0, so the reference clears control[0] (already 0): the set never
    # changes while `remaining` counts down 8, 7, ..., 0 and the group returns zeros.
    rng = np.random.default_rng(2)
    p = np.concatenate([[[10.5, 0.0, 0.0]], _clump(9, rng)])
    v = np.zeros((10, 3))
    v[1, 0] = 50.0
    m = np.ones(10)
    per, _ = oracle.phase_space_clean([np.arange(10)], p[:, 0], p[:, 1], p[:, 2], v[:, 0],
                                      v[:, 1], v[:, 2], m, 1.0, 1.0, 3.0, min_part=0, minp=1)
    g = per[0]
    assert g["control_clean"].tolist() == [False] + [True] * 9
    assert g["clean"] == (0.0, 0.0, 0.0, 0.0, 0.0)
    assert g["rounds"] == 10
    # min_part above the group size: the driver keeps every member
    per2, kept2 = oracle.phase_space_clean([np.arange(10)], p[:, 0], p[:, 1], p[:, 2], v[:, 0],
                                           v[:, 1], v[:, 2], m, 1.0, 1.0, 3.0, min_part=50,
                                           minp=1)
    assert per2[0]["control"].all() and kept2[0]["members"].tolist() == list(range(10))


def test_grid_is_numpy_arange():
    for R, ll in ((0.0, 0.5), (3.7, 0.2), (1e-3, 3e-4)):
        g = cleaning.grid_of(R, ll)
        assert g.tolist() == np.arange(-(R + ll), R + ll, 2 * ll).tolist()
