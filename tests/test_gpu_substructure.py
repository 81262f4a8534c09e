"""Substructure splitting (pyHALMA.py:703-766) and the batched half-mass radius on the
GPU against the oracle: same groups, same order, same count of substructures."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def _groups(x, y, z, ll, minp):
    from pyhalma_amd import engine

    lab = oracle.fof_labels_c(x, y, z, ll)
    return engine.groups_from_labels(lab.astype(np.int32), minp)


def test_half_mass_radius_vs_oracle(gpu):
    from pyhalma_amd import engine, synth

    w = synth.clumps(seed=41, scale=0.01)
    g = _groups(w["x"], w["y"], w["z"], w["ll"], 50)
    red = engine.group_reduce(g, w["x"], w["y"], w["z"], w["mass"])
    got = engine.half_mass_radius(g, w["x"], w["y"], w["z"], w["mass"], red).cpu().numpy()
    for h, mem in enumerate(g.to_arrays()):
        cx, cy, cz, M = oracle.center_of_mass(mem, w["x"], w["y"], w["z"], w["mass"])
        want = oracle.half_mass_radius(cx, cy, cz, M, mem, w["x"], w["y"], w["z"], w["mass"])
        assert abs(got[h] - want) <= 1e-12 * want, (h, got[h], want)
    # closed forms: points on an axis, centre at the origin given explicitly
    import torch
    x = np.array([1.0, 2.0, 3.0, 4.0, 4.0, 1.0, 3.0, 2.0])
    zero = np.zeros(8)
    m = np.array([1.0, 1, 1, 1, 1, 1, 1, 1])
    gg = engine.Groups(torch.arange(8, dtype=torch.int32).cuda(),
                       torch.tensor([0, 4, 8], dtype=torch.int64).cuda(), np.array([0, 4, 8]))
    red = torch.zeros((2, 16), dtype=torch.float64).cuda()
    red[:, 0] = 4.0
    assert engine.half_mass_radius(gg, x, zero, zero, m, red).cpu().tolist() == [2.0, 2.0]


def _compare(sub, want, count):
    got = [sorted(g) for g in sub.to_lists()]
    assert sub.n_found == count
    assert len(got) == len(want), (len(got), len(want))
    for a, b in zip(got, want):
        assert a == sorted(b)


def test_split_two_blobs_joined_by_a_bridge(gpu):
    from pyhalma_amd import engine

    rng = np.random.default_rng(1)
    a = np.array([0.0, 0, 0]) + rng.uniform(-0.6, 0.6, size=(400, 3))
    b = np.array([10.0, 0, 0]) + rng.uniform(-0.6, 0.6, size=(400, 3))
    t = np.linspace(1.4, 8.6, 9)
    p = np.concatenate([a, b, np.stack([t, 0 * t, 0 * t], 1)])
    x, y, z = p[:, 0], p[:, 1], p[:, 2]
    m = np.ones(len(x))
    g = _groups(x, y, z, 1.0, 50)
    sub = engine.split_substructure(g, x, y, z, m, 1.0, minp_sub=25, dec_factor=2.0,
                                    min_ll=0.3, dist_factor=0.75, minp=50)
    _compare(sub, [list(range(400)), list(range(400, 800))], 2)
    assert sub.is_new.tolist() == [True, True] and sub.parent.tolist() == [0, 0]


@pytest.mark.parametrize("levels", [2, 3])
def test_split_bridged_pairs_vs_oracle(gpu, levels):
    from pyhalma_amd import engine, synth

    w = synth.bridged_pairs(seed=5, scale=0.006, n_per_pair=1500)
    ll = float(w["ll"])
    min_ll = ll / (2.0 ** levels) * 1.01
    g = _groups(w["x"], w["y"], w["z"], ll, 50)
    sub = engine.split_substructure(g, w["x"], w["y"], w["z"], w["mass"], ll, minp_sub=25,
                                    dec_factor=2.0, min_ll=min_ll, dist_factor=0.5, minp=50)
    want, count = oracle.split_substructure(g.to_arrays(), w["x"], w["y"], w["z"], w["mass"],
                                            ll, 25, 2.0, min_ll, 0.5, 50)
    print(f"levels={levels} groups in={g.n_groups} out={len(want)} found={count}", flush=True)
    assert count > 0
    _compare(sub, want, count)


def test_split_clumps_vs_oracle(gpu):
    from pyhalma_amd import engine, synth

    w = synth.clumps(seed=43, scale=0.01)
    ll = float(w["ll"])
    g = _groups(w["x"], w["y"], w["z"], ll, 50)
    sub = engine.split_substructure(g, w["x"], w["y"], w["z"], w["mass"], ll, minp_sub=25,
                                    dec_factor=2.0, min_ll=ll / 9, dist_factor=0.75, minp=50)
    want, count = oracle.split_substructure(g.to_arrays(), w["x"], w["y"], w["z"], w["mass"],
                                            ll, 25, 2.0, ll / 9, 0.75, 50)
    print(f"groups in={g.n_groups} out={len(want)} found={count}", flush=True)
    _compare(sub, want, count)


def test_half_mass_radius_drop_in(gpu):
    from pyhalma_amd import halo_properties as hp

    rng = np.random.default_rng(44)
    p = rng.normal(size=(3000, 3))
    m = rng.uniform(1, 2, 3000)
    idx = np.sort(rng.choice(3000, 1200, replace=False))
    cx, cy, cz, M = oracle.center_of_mass(idx, p[:, 0], p[:, 1], p[:, 2], m)
    want = oracle.half_mass_radius(cx, cy, cz, M, idx, p[:, 0], p[:, 1], p[:, 2], m)
    got = hp.half_mass_radius(cx, cy, cz, M, idx, p[:, 0], p[:, 1], p[:, 2], m)
    assert got == want
    with pytest.raises(ValueError):
        hp.half_mass_radius(0.0, 0.0, 0.0, 1.0, [], p[:, 0], p[:, 1], p[:, 2], m)
