"""Pins the reduction and energy-test oracles (CPU only).

Each check is fixed by something other than the oracle's own formula: closed forms on
configurations with known answers (symmetric sets, two bodies at the escape speed, a
3-4-5 triangle), an independent numpy routine (np.average), the argmax and tie rules
of halo_properties.py:546-553 on constructed histograms, and the fixed-point property
of the iterated test (the final bound set passes the test again when v_cm and S are
recomputed from scratch)."""
import math

import numpy as np
import pytest

import oracle


def test_center_of_mass_closed_forms():
    # symmetric pairs: the CM is exactly zero; M is the exact integer sum
    x = np.array([-2.0, 2.0, -0.5, 0.5])
    y = np.array([1.0, -1.0, 3.0, -3.0])
    z = np.zeros(4)
    m = np.array([3.0, 3.0, 5.0, 5.0])
    assert oracle.center_of_mass(np.arange(4), x, y, z, m) == (0.0, 0.0, 0.0, 16.0)
    # two bodies: (m1 x1 + m2 x2) / (m1 + m2)
    cx, cy, cz, M = oracle.center_of_mass([0, 1], np.array([0.0, 4.0]), np.zeros(2),
                                          np.zeros(2), np.array([1.0, 3.0]))
    assert (cx, M) == (3.0, 4.0)
    # no mass: zeros (halo_properties.py:38-43)
    assert oracle.center_of_mass([0, 1], x, y, z, np.zeros(4)) == (0.0, 0.0, 0.0, 0.0)
    # an independent routine on random data, and invariance under reordering part_list
    rng = np.random.default_rng(0)
    X, Y, Z = rng.normal(size=(3, 500))
    W = rng.uniform(1e4, 1e6, 500)
    p = rng.choice(500, 300, replace=False)
    got = oracle.center_of_mass(p, X, Y, Z, W)
    want = [np.average(a[p], weights=W[p]) for a in (X, Y, Z)]
    np.testing.assert_allclose(got[:3], want, rtol=1e-12, atol=1e-15)
    assert got == oracle.center_of_mass(p[::-1], X, Y, Z, W)
    assert oracle.total_mass(np.arange(4), m) == 16.0


def test_cm_velocity_and_furthest_particle():
    v = oracle.cm_velocity(4.0, [0, 1], np.array([1.0, -1.0 / 3.0]), np.zeros(2),
                           np.array([2.0, 2.0]), np.array([1.0, 3.0]))
    assert v == (0.0, 0.0, 2.0)
    assert oracle.cm_velocity(0.0, [0, 1], np.ones(2), np.ones(2), np.ones(2),
                              np.ones(2)) == (0.0, 0.0, 0.0)
    # 3-4-5 triangle around the centre
    x = np.array([0.0, 3.0, 1.0])
    y = np.array([0.0, 4.0, 1.0])
    z = np.zeros(3)
    assert oracle.furthest_particle(0.0, 0.0, 0.0, [0, 1, 2], x, y, z) == 5.0
    assert oracle.furthest_particle(0.0, 0.0, 0.0, [], x, y, z) == 0.0


def _grid(lo, hi, ll):
    return np.linspace(lo - ll, hi + ll, 50)


def test_density_peak_semantics():
    rng = np.random.default_rng(1)
    ll = 0.01
    # a heavy tight clump and a light diffuse one: the peak is the node nearest the
    # heavy clump on every axis
    A = np.array([0.2, 0.7, 0.4])
    B = np.array([0.8, 0.1, 0.9])
    pa = A + rng.normal(scale=1e-4, size=(200, 3))
    pb = B + rng.normal(scale=5e-2, size=(200, 3))
    P = np.concatenate([pa, pb])
    m = np.concatenate([np.full(200, 10.0), np.full(200, 1.0)])
    got = oracle.density_peak(np.arange(400), P[:, 0], P[:, 1], P[:, 2], m, ll)
    for d in range(3):
        g = _grid(P[:, d].min(), P[:, d].max(), ll)
        assert got[d] == g[np.argmin(np.abs(g - A[d]))]
    # two equal clumps: the first node in (ix, iy, iz) order wins, i.e. the clump with
    # the smaller x
    P2 = np.array([[0.1, 0.9, 0.9], [0.1, 0.9, 0.9], [0.9, 0.1, 0.1], [0.9, 0.1, 0.1]])
    got = oracle.density_peak(np.arange(4), P2[:, 0], P2[:, 1], P2[:, 2], np.ones(4), ll)
    assert got[0] == _grid(0.1, 0.9, ll)[np.argmin(np.abs(_grid(0.1, 0.9, ll) - 0.1))]
    assert got[1] == _grid(0.1, 0.9, ll)[np.argmin(np.abs(_grid(0.1, 0.9, ll) - 0.9))]
    # no positive mass: the centre node
    g = _grid(0.1, 0.9, ll)
    assert oracle.density_peak(np.arange(4), P2[:, 0], P2[:, 1], P2[:, 2], np.zeros(4),
                               ll) == (g[25], g[25], g[25])
    with pytest.raises(ValueError):
        oracle.density_peak([], P2[:, 0], P2[:, 1], P2[:, 2], np.ones(4), ll)


def test_two_body_escape_speed():
    """Two equal masses m at separation r moving apart at +-v (v_cm = 0): each has
    S = m/r and E = v^2/2 - factor*G*m/r, so the pair is bound iff v^2 <= 2*factor*G*m/r.
    The offset (2, 3, 6)/1024 differs in every coordinate (the reference's skip rule)
    and has the exact length r = 7/1024 in FP32 and FP64."""
    m, r = 1e10, 7.0 / 1024.0
    x = np.array([0.0, 2.0 / 1024.0])
    y = np.array([0.0, 3.0 / 1024.0])
    z = np.array([0.0, 6.0 / 1024.0])
    zero = np.zeros(2)
    for factor in (1.0, 9.0):
        v_esc = math.sqrt(2.0 * factor * oracle.G_CONST * m / r)
        for scale, want in ((0.999, True), (1.001, False)):
            v = scale * v_esc
            b, it, S = oracle.unbind_iterative([0, 2], [0, 1], x, y, z,
                                               np.array([v, -v]), zero, zero,
                                               np.full(2, m), factor)
            assert S.tolist() == [m / r, m / r]
            assert b.tolist() == [want, want]
            # unbound pair: pass 1 removes both, the empty set ends the iteration
            assert it.tolist() == [1]


def test_iteration_fixed_point_and_single_pass():
    w = __import__("pyhalma_amd.synth", fromlist=["x"]).plummer_with_background(
        n_sphere=600, n_bg=60, seed=3)
    n = len(w["x"])
    members = np.arange(n)
    off = [0, n]
    args = (w["x"], w["y"], w["z"], w["vx"], w["vy"], w["vz"], w["mass"])
    b, it, _ = oracle.unbind_iterative(off, members, *args, 1.0)
    assert it[0] >= 2 and 0 < b.sum() < n
    # the final set is a fixed point: recompute v_cm and S over it from scratch
    sel = members[b]
    m = np.asarray(w["mass"], np.float64)[sel]
    vcm = [math.fsum((m * np.asarray(v)[sel]).tolist()) / math.fsum(m.tolist())
           for v in args[3:6]]
    pos = [np.asarray(a)[sel] for a in args[:3]]
    S = oracle.binding_energy(np.asarray(w["mass"])[sel], *pos, *pos)
    _, ok = oracle.energy_test(S, *(np.asarray(a)[sel] for a in args[3:6]), vcm, 1.0)
    assert ok.all()
    # one pass is the reference's single test with v_cm over all members
    b1, it1, _ = oracle.unbind_iterative(off, members, *args, 1.0, max_iter=1)
    m_all = np.asarray(w["mass"], np.float64)
    vcm_all = [math.fsum((m_all * np.asarray(v)).tolist()) / math.fsum(m_all.tolist())
               for v in args[3:6]]
    S_all = oracle.binding_energy(w["mass"], *args[:3], *args[:3])
    _, ok_all = oracle.energy_test(S_all, *args[3:6], vcm_all, 1.0)
    assert it1.tolist() == [1] and np.array_equal(b1, ok_all)
    # removal only shrinks the set
    assert not np.any(b & ~b1)


def test_rps_and_most_bound_closed_forms():
    """Two gas cells and one star, no subsampling (sizes under the limit): the sums are
    m_j / r by hand; cold/hot split at T = 5e4 K (halo_gas.py:478-480)."""
    G = oracle.G_CONST
    gx = np.array([0.0, 0.5], np.float32)
    gy = np.array([0.0, 0.25], np.float32)
    gz = np.array([0.0, 0.125], np.float32)
    gm = np.array([1e12, 1e6], np.float32)
    gv = np.zeros(2)
    st = (np.array([0.375], np.float32), np.array([0.0625], np.float32),
          np.array([0.5], np.float32), np.array([1e3], np.float32))
    dm = (np.zeros(0, np.float32),) * 4
    # cell 0 (cold) sits at rest in the potential of the others: bound; cell 1 (hot)
    # moves at 3x its escape speed: unbound
    S1 = 1e12 / math.sqrt(0.5 ** 2 + 0.25 ** 2 + 0.125 ** 2) + \
        1e3 / math.sqrt(0.125 ** 2 + 0.1875 ** 2 + 0.375 ** 2)
    vx = np.array([0.0, 3.0 * math.sqrt(2.0 * 2.0 * G * S1)])
    temp = np.array([1e3, 1e6])
    tot, frac, ucold, uhot = oracle.rps(gx, gy, gz, vx, gv, gv, gm, temp, *dm, *st,
                                        0.0, 0.0, 0.0, 100, 1.0, 1)
    assert tot == pytest.approx(1e12 + 1e6, rel=1e-7)
    assert frac == pytest.approx(1e12 / (1e12 + 1e6), rel=1e-7)
    assert ucold == 0.0 and uhot == pytest.approx(1e6, rel=1e-7)
    # the star nearest the heavy cell is the most bound of three
    sx = np.array([0.4, 0.05, 0.3], np.float32)
    sy = np.array([0.03, 0.01, 0.2], np.float32)
    sz = np.array([0.02, 0.04, 0.06], np.float32)
    sm = np.ones(3, np.float32)
    res = oracle.most_bound_particle(gx, gy, gz, gm, *dm, sx, sy, sz, sm,
                                     np.array([7, 8, 9]), 100, 1.0)
    assert res[3] == 8
