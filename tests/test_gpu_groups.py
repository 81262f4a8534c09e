"""Groups, per-group reductions and density peaks on the GPU against the oracle.

Bars: group lists identical (order, membership); sums within 1e-14 relative of the
correctly rounded oracle (double-double accumulation); furthest particle within 1e-14;
density peaks bit-identical (equal masses, FP32 masses, FP64 masses with near ties, a
group too large for the LDS table)."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def _clumps(scale=0.02, seed=7):
    from pyhalma_amd import synth

    return synth.clumps(seed=seed, scale=scale)


def test_groups_from_labels_matches_oracle(gpu):
    from pyhalma_amd import engine

    w = _clumps()
    lab = oracle.fof_labels_c(w["x"], w["y"], w["z"], w["ll"])
    for min_size in (1, 2, 50):
        g = engine.groups_from_labels(lab.astype(np.int32), min_size)
        want = oracle.labels_to_groups(lab, min_size)
        print(f"min_size={min_size} groups={len(want)}", flush=True)
        assert g.to_lists() == want


def test_groups_edge_cases(gpu):
    from pyhalma_amd import engine

    print("case=empty", flush=True)
    assert engine.groups_from_labels(np.zeros(0, np.int32)).n_groups == 0
    lab = np.arange(10, dtype=np.int32)  # all singletons
    print("case=singletons", flush=True)
    assert engine.groups_from_labels(lab, 2).n_groups == 0
    assert engine.groups_from_labels(lab, 1).to_lists() == [[i] for i in range(10)]
    print("case=one-group", flush=True)
    lab = np.zeros(300000, np.int32)  # one big group
    g = engine.groups_from_labels(lab, 2)
    assert g.n_groups == 1 and g.to_lists()[0] == list(range(300000))


def test_groups_from_labels_large(gpu):
    """9M labels in 200,000 groups (FoF-style: a group's label is its minimum index): the
    sort and the prefix sums then take more than 4096 tiles of 2048 elements, the sizes
    where the prefix sum scans its block sums in a kernel of its own."""
    from pyhalma_amd import engine

    n, n_grp, min_size = 9_000_000, 200_000, 50
    rng = np.random.default_rng(11)
    assign = rng.integers(0, n_grp, n)
    _, first, inv = np.unique(assign, return_index=True, return_inverse=True)
    lab = first[inv].astype(np.int32)
    print(f"case=large n={n} groups={len(first)}", flush=True)
    g = engine.groups_from_labels(lab, min_size)
    want = oracle.labels_to_groups(lab, min_size)
    assert g.n_groups == len(want)
    np.testing.assert_array_equal(g.sizes(), [len(w) for w in want])
    got = g.members[: g.n_members].cpu().numpy()
    np.testing.assert_array_equal(got, np.concatenate(want))


@pytest.mark.parametrize("pos_dt", [np.float32, np.float64])
def test_group_reduce(gpu, pos_dt):
    from pyhalma_amd import engine

    w = _clumps(0.01, seed=2)
    x, y, z = (w[k].astype(pos_dt) for k in ("x", "y", "z"))
    lab = oracle.fof_labels_c(x, y, z, w["ll"])
    g = engine.groups_from_labels(lab.astype(np.int32), 20)
    out = engine.group_reduce(g, x, y, z, w["mass"], w["vx"], w["vy"], w["vz"]).cpu().numpy()
    for k, grp in enumerate(g.to_lists()):
        cx, cy, cz, M = oracle.center_of_mass(grp, x, y, z, w["mass"])
        v = oracle.cm_velocity(M, grp, w["vx"], w["vy"], w["vz"], w["mass"])
        r = oracle.furthest_particle(out[k, 1], out[k, 2], out[k, 3], grp, x, y, z)
        np.testing.assert_allclose(out[k, [1, 2, 3, 0]], [cx, cy, cz, M], rtol=1e-14, atol=1e-300)
        np.testing.assert_allclose(out[k, 4:7], v, rtol=1e-13, atol=1e-12)
        assert out[k, 7] == pytest.approx(r, rel=1e-14)
        assert out[k, 15] == 1.0  # FP32 masses of one group sum exactly in FP64


@pytest.mark.parametrize("masses", ["equal", "f32", "f64_ties"])
def test_density_peak_bit_exact(gpu, masses):
    from pyhalma_amd import engine

    w = _clumps(0.01, seed=4)
    x, y, z = w["x"], w["y"], w["z"]
    rng = np.random.default_rng(1)
    if masses == "equal":
        m = np.full(len(x), 1e5)
    elif masses == "f32":
        m = w["mass"]
    else:  # full-mantissa FP64 masses, two values that nearly tie
        m = rng.uniform(1.0, 1.0 + 1e-12, len(x)) * 3.3
    lab = oracle.fof_labels_c(x, y, z, w["ll"])
    g = engine.groups_from_labels(lab.astype(np.int32), 20)
    red = engine.group_reduce(g, x, y, z, m)
    got = engine.density_peak(g, x, y, z, m, w["ll"], reduced=red).cpu().numpy()
    print(f"masses={masses} groups={g.n_groups}", flush=True)
    for k, grp in enumerate(g.to_lists()):
        want = oracle.density_peak(grp, x, y, z, m, w["ll"])
        assert tuple(got[k]) == want, (k, got[k], want)


def test_density_peak_dense_fallback_and_ties(gpu):
    from pyhalma_amd import engine

    rng = np.random.default_rng(3)
    # 20000 uniform particles in a cube: ~all 50^3 nodes distinct -> dense path
    n = 20000
    x, y, z = (rng.uniform(0, 1, n) for _ in range(3))
    m = np.ones(n)
    # a second group where two nodes tie exactly (first in C order must win)
    x2 = np.array([0.0, 0.0, 1.0, 1.0, 0.5])
    y2 = np.array([0.0, 0.0, 1.0, 1.0, 0.5])
    z2 = np.array([0.0, 0.0, 1.0, 1.0, 0.5])
    X, Y, Z = np.concatenate([x, x2]), np.concatenate([y, y2]), np.concatenate([z, z2])
    M = np.ones(len(X))
    lab = np.concatenate([np.zeros(n), np.full(5, n)]).astype(np.int32)
    g = engine.groups_from_labels(lab, 2)
    got = engine.density_peak(g, X, Y, Z, M, 0.01).cpu().numpy()
    for k, grp in enumerate(g.to_lists()):
        assert tuple(got[k]) == oracle.density_peak(grp, X, Y, Z, M, 0.01)


def test_dropin_reductions(gpu):
    from pyhalma_amd import halo_properties as hp

    w = _clumps(0.005, seed=9)
    x, y, z, m = w["x"], w["y"], w["z"], w["mass"]
    grp = np.arange(0, len(x), 7)
    cx, cy, cz, M = hp.center_of_mass(grp, x, y, z, m)
    want = oracle.center_of_mass(grp, x, y, z, m)
    np.testing.assert_allclose([cx, cy, cz, M], want, rtol=1e-14)
    assert hp.total_mass(grp, m) == pytest.approx(want[3], rel=1e-14)
    v = hp.CM_velocity(M, grp, w["vx"], w["vy"], w["vz"], m)
    np.testing.assert_allclose(v, oracle.cm_velocity(M, grp, w["vx"], w["vy"], w["vz"], m),
                               rtol=1e-13, atol=1e-12)
    r = hp.furthest_particle(cx, cy, cz, grp, x, y, z)
    assert r == pytest.approx(oracle.furthest_particle(cx, cy, cz, grp, x, y, z), rel=1e-14)
    assert hp.density_peak(grp, x, y, z, m, w["ll"]) == oracle.density_peak(grp, x, y, z, m,
                                                                           w["ll"])
    assert hp.center_of_mass([], x, y, z, m) == (0.0, 0.0, 0.0, 0.0)
    with pytest.raises(ValueError):
        hp.density_peak([], x, y, z, m, 0.1)


def test_gather_members_then_reduce(gpu):
    """Member-ordered copies give the same reductions, peaks and unbinding as the
    particle-ordered arrays."""
    from pyhalma_amd import engine

    w = _clumps(0.01, seed=5)
    lab = oracle.fof_labels_c(w["x"], w["y"], w["z"], w["ll"])
    g = engine.groups_from_labels(lab.astype(np.int32), 20)
    md = engine.gather_members(g, w["x"], w["y"], w["z"], w["mass"], w["vx"], w["vy"], w["vz"])
    mem = g.members[: g.n_members].cpu().numpy()
    assert np.array_equal(md.x.cpu().numpy(), w["x"][mem])
    assert np.array_equal(md.vz.cpu().numpy(), w["vz"][mem])
    a = engine.group_reduce(g, w["x"], w["y"], w["z"], w["mass"], w["vx"], w["vy"], w["vz"])
    b = engine.group_reduce(md.groups, md.x, md.y, md.z, md.m, md.vx, md.vy, md.vz)
    assert torch_equal(a, b)
    pa = engine.density_peak(g, w["x"], w["y"], w["z"], w["mass"], w["ll"])
    pb = engine.density_peak(md.groups, md.x, md.y, md.z, md.m, w["ll"])
    assert torch_equal(pa, pb)
    ua = engine.unbind(g, w["x"], w["y"], w["z"], w["vx"], w["vy"], w["vz"], w["mass"])
    ub = engine.unbind(md.groups, md.x, md.y, md.z, md.vx, md.vy, md.vz, md.m)
    assert torch_equal(ua.bound, ub.bound)


def torch_equal(a, b):
    import torch

    return bool(torch.equal(a, b))
