"""Progenitor matching (pyHALMA.py:1211-1266) on the GPU against the oracle."""
import numpy as np
import pytest
import torch

import oracle

pytestmark = pytest.mark.gpu


def _groups(sizes):
    from pyhalma_amd import engine
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    mem = np.arange(max(off[-1], 1), dtype=np.int32)
    return engine.Groups(torch.from_numpy(mem).cuda(), torch.from_numpy(off).cuda(), off)


def _both(new, masses, old, omm):
    from pyhalma_amd import engine
    ids = np.concatenate(new).astype(np.int64) if new else np.zeros(0, np.int64)
    g = _groups([len(h) for h in new])
    got = engine.progenitors(g, ids, masses, [len(o) for o in old],
                             np.concatenate(old) if old else np.zeros(0, np.int64), omm)
    want = oracle.progenitors(g.to_arrays(), ids, masses, old, omm)
    return [a.tolist() for a in got], [a.tolist() for a in want]


def test_hand_built_catalogue(gpu):
    old = [np.array([1, 2, 3]), np.array([10, 11]), np.array([20, 21, 22]),
           np.array([30]), np.array([40, 41])]
    omm = [100.0, 10.0, 40.0, 1.0, 50.0]
    new = [np.array([1, 2, 99]), np.array([3, 10, 5]), np.array([20, 40]),
           np.array([21, 30]), np.array([41]), np.array([7, 8]), np.array([11, 22, 2])]
    got, want = _both(new, [90.0, 95.0, 60.0, 45.0, 10.0, 5.0, 200.0], old, omm)
    assert got == want
    assert got[3] == [0, 2, 3, 3, -1, 0, 3]


@pytest.mark.parametrize("spread", [10 ** 6, 2 ** 62])
def test_random_catalogues_vs_oracle(gpu, spread):
    rng = np.random.default_rng(spread % 1000)
    pool = np.unique(rng.integers(-spread, spread, size=60000))
    rng.shuffle(pool)
    # old: 250 disjoint haloes; new: 300 haloes drawing from 1-4 old haloes and new IDs
    cuts = np.sort(rng.choice(np.arange(1, 30000), 249, replace=False))
    old = np.split(pool[:30000], cuts)
    omm = rng.uniform(1.0, 100.0, len(old))
    new = []
    fresh = np.array_split(pool[30000:], 300)  # IDs no old halo holds
    for i in range(300):
        parts = [fresh[i][: rng.integers(0, 40)]]
        for oh in rng.choice(len(old), rng.integers(0, 5), replace=False):
            o = old[oh]
            parts.append(rng.choice(o, min(len(o), rng.integers(1, 30)), replace=False))
        h = np.unique(np.concatenate(parts))
        if h.size == 0:
            h = fresh[i][:1]
        new.append(rng.permutation(h))
    masses = rng.uniform(1.0, 150.0, len(new))
    got, want = _both(new, masses, old, omm)
    assert got == want
    assert max(want[2]) >= 3 and {-1, 2, 3} <= set(want[3])


def test_ids_repeated_across_old_haloes(gpu):
    # the old catalogue lists some IDs in two haloes (as overlapping domains can): each
    # such ID links the new halo to both
    old = [np.array([1, 2, 3, 4]), np.array([1, 2, 3, 4, 5]), np.array([6, 7])]
    got, want = _both([np.array([1, 2, 3, 4, 6])], [10.0], old, [5.0, 5.0, 1.0])
    assert got == want
    # equal masses 5 and 5: the higher old index ranks first (pro1 = 2, pro2 = 1)
    assert got[0] == [2] and got[1] == [1] and got[2] == [3]


def test_many_haloes(gpu):
    # 2^17 one-member new haloes, each the continuation of one of 2^16 + 10 old haloes
    from pyhalma_amd import engine
    n_old = 2 ** 16 + 10
    n_new = 2 ** 17
    ids = np.arange(n_new, dtype=np.int64) * 3 + 7
    old_of_new = np.arange(n_new) % n_old
    order = np.argsort(old_of_new, kind="stable")
    old_npart = np.bincount(old_of_new, minlength=n_old)
    omm = 1.0 + np.arange(n_old, dtype=np.float64)
    g = _groups(np.ones(n_new, np.int64))
    pro1, pro2, nm, t = engine.progenitors(g, ids, np.full(n_new, 1e9), old_npart, ids[order],
                                           omm)
    assert (pro1 == old_of_new + 1).all() and (pro2 == 0).all()
    assert (nm == 1).all() and (t == 0).all()


def test_empty_catalogues(gpu):
    from pyhalma_amd import engine
    g = _groups([2, 1])
    out = engine.progenitors(g, np.array([4, 5, 6]), [1.0, 2.0], [], np.zeros(0), [])
    assert all((a == 0).all() for a in out)


def test_more_progenitors_than_the_lds_set(gpu):
    # one new halo made of 3000 one-member old haloes (beyond the 2048 distinct old haloes
    # a block keeps in LDS), beside ordinary ones
    rng = np.random.default_rng(9)
    old = [np.array([i]) for i in range(3000)] + [np.array([5000, 5001]), np.array([6000])]
    omm = rng.uniform(1.0, 10.0, len(old))
    new = [np.arange(3000)[::-1].copy(), np.array([5000, 6000]), np.array([7000])]
    got, want = _both(new, [20000.0, 3.0, 1.0], old, omm)
    assert got == want and got[2][0] == 3000
