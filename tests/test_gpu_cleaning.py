"""Phase-space cleaning (calc_cell / clean_cell, halo_properties.py:113-278, driver
pyHALMA.py:789-843) on the GPU against the oracle: same control arrays, rounds, grid
sizes and node indices; centres, masses and radii to rounding."""
import numpy as np
import pytest
import torch

import oracle

pytestmark = pytest.mark.gpu


def _contaminated(seed, k=40, dtype=np.float64):
    """k separate gaussian groups (40..2500 members) whose first 0..45 % of members are
    interlopers moving at about 300 against a dispersion of 50 per axis."""
    rng = np.random.default_rng(seed)
    P, V, M, groups = [], [], [], []
    base = 0
    for h in range(k):
        n = int(rng.integers(40, 2500))
        f = rng.uniform(0, 0.45)
        p = np.array([h * 100.0, 0, 0]) + rng.normal(size=(n, 3)) * rng.uniform(0.5, 2.0)
        v = rng.normal(size=(n, 3)) * 50
        ni = int(f * n)
        d = rng.normal(size=(ni, 3))
        d /= np.linalg.norm(d, axis=1)[:, None]
        v[:ni] += d * rng.normal(300, 100, size=(ni, 1))
        P.append(p)
        V.append(v)
        M.append(rng.uniform(0.5, 2.0, n))
        groups.append(np.arange(base, base + n))
        base += n
    P, V, M = (np.concatenate(a).astype(dtype) for a in (P, V, M))
    return P, V, M, groups


def _dev_groups(groups):
    from pyhalma_amd import engine
    off = np.concatenate([[0], np.cumsum([len(g) for g in groups])]).astype(np.int64)
    mem = np.concatenate(groups).astype(np.int32) if groups else np.zeros(1, np.int32)
    return engine.Groups(torch.from_numpy(mem).cuda(), torch.from_numpy(off).cuda(), off)


def _close(a, b, rel=1e-13):
    return abs(a - b) <= rel * max(abs(a), abs(b), 1e-300)


def _check(res, per, kept):
    assert len(per) == len(res.n_cell)
    ctrl = res.control_clean.cpu().numpy()
    fin = res.control.cpu().numpy()
    for g, want in enumerate(per):
        a, b = int(res.offsets_host[g]), int(res.offsets_host[g + 1])
        assert (ctrl[a:b] == want["control_clean"]).all(), g
        assert (fin[a:b] == want["control"]).all(), g
        assert res.rounds[g] == want["rounds"], (g, res.rounds[g], want["rounds"])
        assert res.n_cell[g] == want["n_cell"], g
        assert (res.which_of(g) == want["which"]).all(), g
        for q in range(5):
            assert _close(res.clean[g, q], want["clean"][q]), (g, q, res.clean[g], want["clean"])
    assert len(kept) == res.kept.n_groups
    got = res.kept.to_arrays()
    for k, want in enumerate(kept):
        assert got[k].tolist() == want["members"].tolist(), k
        assert res.npart[k] == want["members"].size
        assert res.kept_n_cell[k] == want["n_cell"]
        assert _close(res.mass[k], want["mass"]) and _close(res.rmax[k], want["rmax"])
        for q in range(3):
            assert _close(res.centre[k, q], want["centre"][q]), k


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("ll", [0.2, 0.4, 0.8])
def test_contaminated_groups_vs_oracle(gpu, ll, dtype):
    from pyhalma_amd import engine
    P, V, M, groups = _contaminated(1, dtype=dtype)
    args = (P[:, 0], P[:, 1], P[:, 2], V[:, 0], V[:, 1], V[:, 2], M)
    res = engine.phase_space_clean(_dev_groups(groups), *args, ll, 1.0, 3.0)
    per, kept = oracle.phase_space_clean(groups, *args, ll, 1.0, 3.0)
    print(f"ll={ll} dtype={np.dtype(dtype).name} rounds={np.bincount(res.rounds)} "
          f"kept={res.kept.n_groups}", flush=True)
    _check(res, per, kept)


def test_outliers_and_index_quirk_vs_oracle(gpu):
    from pyhalma_amd import engine
    rng = np.random.default_rng(2)
    # group 0: five movers among 95 at rest (one round); group 1: a lone member and a
    # mover that the control-index quirk never takes out (ten rounds, zeros); group 2:
    # every member alone in its cell (all removed in the first round)
    p0 = rng.uniform(-1e-3, 1e-3, size=(100, 3))
    v0 = np.zeros((100, 3))
    v0[[3, 17, 40, 41, 99], 0] = 200.0
    p1 = np.concatenate([[[10.5, 0.0, 0.0]], rng.uniform(-1e-3, 1e-3, size=(9, 3))]) + 50.0
    v1 = np.zeros((10, 3))
    v1[1, 0] = 50.0
    p2 = np.stack([np.arange(60) * 5.0, np.zeros(60), np.zeros(60)], 1) + 200.0
    v2 = rng.normal(size=(60, 3))
    P = np.concatenate([p0, p1, p2])
    V = np.concatenate([v0, v1, v2])
    M = np.ones(len(P))
    groups = [np.arange(100), np.arange(100, 110), np.arange(110, 170)]
    args = (P[:, 0], P[:, 1], P[:, 2], V[:, 0], V[:, 1], V[:, 2], M)
    res = engine.phase_space_clean(_dev_groups(groups), *args, 1.0, 1.0, 3.0, min_part=0,
                                   minp=1)
    per, kept = oracle.phase_space_clean(groups, *args, 1.0, 1.0, 3.0, min_part=0, minp=1)
    assert res.rounds.tolist() == [1, 10, 1]
    assert res.clean[1].tolist() == [0.0] * 5 and res.clean[2].tolist() == [0.0] * 5
    _check(res, per, kept)


@pytest.mark.parametrize("psc,min_part", [(False, 50), (True, 1000)])
def test_driver_switches(gpu, psc, min_part):
    from pyhalma_amd import engine
    P, V, M, groups = _contaminated(3, k=12)
    args = (P[:, 0], P[:, 1], P[:, 2], V[:, 0], V[:, 1], V[:, 2], M)
    res = engine.phase_space_clean(_dev_groups(groups), *args, 0.3, 1.0, 3.0, psc=psc,
                                   min_part=min_part, minp=50)
    per, kept = oracle.phase_space_clean(groups, *args, 0.3, 1.0, 3.0, psc=psc,
                                         min_part=min_part, minp=50)
    _check(res, per, kept)


def test_fof_clumps_vs_oracle(gpu):
    from pyhalma_amd import engine, synth
    w = synth.clumps(seed=43, scale=0.05)
    ll = float(w["ll"])
    g = engine.groups_from_labels(oracle.fof_labels_c(w["x"], w["y"], w["z"], ll)
                                  .astype(np.int32), 50)
    args = (w["x"], w["y"], w["z"], w["vx"], w["vy"], w["vz"], w["mass"])
    res = engine.phase_space_clean(g, *args, ll / 16, 1.0, 3.0)
    per, kept = oracle.phase_space_clean(g.to_arrays(), *args, ll / 16, 1.0, 3.0)
    _check(res, per, kept)


def test_wide_grid_uses_64_bit_cell_keys(gpu):
    # one member 3000 ll from a dense core: about 3000 nodes per axis, 2.7e10 cells, so the
    # cell key (ix n + iy) n + iz needs more than 32 bits
    from pyhalma_amd import engine
    rng = np.random.default_rng(7)
    core = rng.normal(size=(3000, 3)) * 0.3
    P = np.concatenate([core, [[3000.0, 1.0, -1.0]]])
    V = rng.normal(size=(len(P), 3)) * 10
    V[:200] += 400.0
    M = rng.uniform(1, 2, len(P))
    groups = [np.arange(len(P))]
    args = (P[:, 0], P[:, 1], P[:, 2], V[:, 0], V[:, 1], V[:, 2], M)
    res = engine.phase_space_clean(_dev_groups(groups), *args, 1.0, 1.0, 3.0)
    per, kept = oracle.phase_space_clean(groups, *args, 1.0, 1.0, 3.0)
    assert per[0]["n_cell"] ** 3 > 2 ** 32
    _check(res, per, kept)


def test_drop_in_calc_cell_and_clean_cell(gpu):
    from oracle import cleaning
    from pyhalma_amd import halo_properties as hp
    P, V, M, groups = _contaminated(5, k=3)
    args = (P[:, 0], P[:, 1], P[:, 2], V[:, 0], V[:, 1], V[:, 2], M)
    ll = 0.3
    for part_list in groups:
        part_list = part_list[::-1].copy()  # any member order
        cx, cy, cz, mass = oracle.center_of_mass(part_list, P[:, 0], P[:, 1], P[:, 2], M)
        R = oracle.furthest_particle(cx, cy, cz, part_list, P[:, 0], P[:, 1], P[:, 2])
        grid = np.arange(-(R + ll), R + ll, 2 * ll)
        n = len(grid)
        arrs = (np.zeros((n, n, n, 3)), np.zeros((n, n, n)), np.zeros((n, n, n)),
                np.zeros((n, n, n)))
        vcm, mc, qc, sc, wx, wy, wz = hp.calc_cell(cx, cy, cz, grid, part_list, *args, *arrs)
        cells, which = cleaning.calc_cell(cx, cy, cz, grid, part_list, *args)
        assert (np.stack([wx, wy, wz], 1) == which).all() and wx.dtype == np.int32
        assert int((qc > 0).sum()) == len(cells)
        for (ix, iy, iz), (vc, cm, q, sig) in cells.items():
            assert qc[ix, iy, iz] == q
            assert _close(mc[ix, iy, iz], cm) and _close(sc[ix, iy, iz], sig, 1e-12)
            for a in range(3):
                assert _close(vcm[ix, iy, iz, a], vc[a], 1e-12)
        got = hp.clean_cell(cx, cy, cz, mass, R, grid, part_list, *args, vcm, mc, qc, sc, ll,
                            1.0, 3.0, wx, wy, wz)
        want = cleaning.clean_cell(part_list, *args, cells, which, ll, 1.0, 3.0)
        assert (got[5] == want[5]).all() and got[5].dtype == np.int32
        assert (np.stack(got[6:9], 1) == want[6]).all()
        for q in range(5):
            assert _close(got[q], want[q])
    with pytest.raises(ValueError):
        hp.calc_cell(0.0, 0.0, 0.0, np.array([0.0, 1.0, 3.0]), groups[0], *args, *arrs)
