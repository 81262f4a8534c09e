"""Oracle of the progenitor matching (pyHALMA.py:1211-1266), pinned by catalogues whose
progenitors and merger types follow by hand from the reference's text."""
import numpy as np

import oracle


def _run(new, masses, old, omm):
    ids = np.concatenate(new) if new else np.zeros(0, np.int64)
    groups, k = [], 0
    for h in new:
        groups.append(np.arange(k, k + len(h)))
        k += len(h)
    return [a.tolist() for a in oracle.progenitors(groups, ids, masses, old, omm)]


def test_hand_built_catalogue():
    old = [np.array([1, 2, 3]), np.array([10, 11]), np.array([20, 21, 22]),
           np.array([30]), np.array([40, 41])]
    omm = [100.0, 10.0, 40.0, 1.0, 50.0]
    new = [
        np.array([1, 2, 99]),        # only old 0: pro1 = 1, mass 90 -> 1.2 * 90 >= 100, one
                                     # intersection: type 0
        np.array([3, 10, 5]),        # old 0 (100) and old 1 (10): f = 0.1 -> minor (2)
        np.array([20, 40]),          # old 2 (40) and old 4 (50): f = 0.8 -> the major (1) is
                                     # overwritten by the else branch: 3
        np.array([21, 30]),          # old 2 (40) and old 3 (1): f = 0.025 -> 3
        np.array([41]),              # old 4 (50), new mass 10: 12 < 50 -> breaking apart (-1)
        np.array([7, 8]),            # no progenitor
        np.array([11, 22, 2]),       # old 0, 1, 2: first 100, second 40: f = 0.4 -> 3
    ]
    masses = [90.0, 95.0, 60.0, 45.0, 10.0, 5.0, 200.0]
    pro1, pro2, nm, t = _run(new, masses, old, omm)
    assert pro1 == [1, 1, 5, 3, 5, 0, 1]
    assert pro2 == [0, 2, 3, 4, 0, 0, 3]
    assert nm == [1, 2, 2, 2, 1, 0, 3]
    assert t == [0, 2, 3, 3, -1, 0, 3]


def test_breakup_takes_precedence_and_fraction_edges():
    old = [np.array([1]), np.array([2]), np.array([3])]
    # second / first exactly 1/20 and exactly 1/3 (as FP64 quotients) are not strictly
    # inside the minor range: type 3
    omm = [60.0, 3.0, 20.0]
    new = [np.array([1, 2]), np.array([1, 3]), np.array([1, 2, 3])]
    pro1, pro2, nm, t = _run(new, [60.0, 60.0, 10.0], old, omm)
    assert (3.0 / 60.0) == 1 / 20 and (20.0 / 60.0) == 1 / 3
    assert t[:2] == [3, 3] and pro2[:2] == [2, 3]
    # 1.2 * 10 < 60: breaking apart, pro2 stays 0 even with three intersections
    assert (pro1[2], pro2[2], nm[2], t[2]) == (1, 0, 3, -1)
