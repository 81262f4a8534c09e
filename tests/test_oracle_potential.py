"""Pins the pair-potential oracle (CPU only).

Checks, each chosen so a plausible slip (a dropped coordinate, the wrong skip rule, a
sign, a transposed source/test) fails one of them:
  * the reference's own Fortran outputs (golden fixture written from oracle/_ref);
  * the reference routine itself, when oracle/_ref is built here;
  * closed forms: two and three bodies, linearity in the masses, 1/lambda scaling,
    permutation invariance of the sources;
  * the three-coordinate skip rule on hand-built pairs;
  * the analytic Plummer binding energy as a statistical check.
"""
import math

import numpy as np
import pytest

import oracle

GOLD_CASES = ("self", "cross", "lattice", "dups")


def _case(g, name):
    return tuple(g[f"{name}_{k}"] for k in ("m", "x", "y", "z", "tx", "ty", "tz"))


@pytest.fixture(scope="module")
def gold(golden_dir):
    return np.load(f"{golden_dir}/pot_ref.npz")


@pytest.mark.parametrize("name", GOLD_CASES)
def test_f32_restatement_matches_reference_outputs(gold, name):
    args = _case(gold, name)
    ref = gold[f"{name}_ref"]
    got = oracle.pot_sum_f32(*args, nthreads=1)
    np.testing.assert_allclose(got, ref, rtol=2e-5, atol=0)


@pytest.mark.parametrize("name", GOLD_CASES)
def test_f64_oracle_matches_reference_outputs(gold, name):
    args = _case(gold, name)
    ref = gold[f"{name}_ref"].astype(np.float64)
    got = oracle.pot_sum_f64(*args)
    # the reference sums in FP32: its own error is ~ n * 2^-24 relative at worst
    n = len(args[0])
    np.testing.assert_allclose(got, ref, rtol=max(1e-5, n * 6e-8), atol=0)


@pytest.mark.skipif(not oracle.have_ref(), reason="oracle/_ref not built")
def test_restatement_matches_reference_routine_fresh_inputs():
    rng = np.random.default_rng(7)
    p = rng.normal(size=(400, 3)).astype(np.float32)
    m = rng.uniform(1, 3, 400).astype(np.float32)
    t = rng.normal(size=(123, 3)).astype(np.float32)
    ref = oracle.ref_pot_sum_f32(m, p[:, 0], p[:, 1], p[:, 2], t[:, 0], t[:, 1], t[:, 2], 2)
    got = oracle.pot_sum_f32(m, p[:, 0], p[:, 1], p[:, 2], t[:, 0], t[:, 1], t[:, 2], 2)
    np.testing.assert_allclose(got, ref, rtol=2e-5)


def test_two_and_three_bodies_closed_form():
    # masses and positions exact in FP32; r = 5 and 13 exactly (3-4-5, 5-12-13)
    m = np.array([2.0, 7.0, 3.0], np.float32)
    x = np.array([0.0, 3.0, 5.0], np.float32)
    y = np.array([0.0, 4.0, 12.0], np.float32)
    z = np.array([1.0, 2.0, 3.0], np.float32)
    # z differences are nonzero so no pair is skipped; distances in 3-D
    d01 = math.sqrt(3 ** 2 + 4 ** 2 + 1)
    d02 = math.sqrt(5 ** 2 + 12 ** 2 + 2 ** 2)
    d12 = math.sqrt(2 ** 2 + 8 ** 2 + 1)
    want = np.array([7 / d01 + 3 / d02, 2 / d01 + 3 / d12, 2 / d02 + 7 / d12])
    got = oracle.pot_sum_f64(m, x, y, z, x, y, z)
    np.testing.assert_allclose(got, want, rtol=1e-15)


def test_skip_rule_any_equal_coordinate():
    # the source shares x (then y, then z, then nothing) with the test at the origin
    m = np.ones(4, np.float32)
    x = np.array([0.0, 1.0, 1.0, 1.0], np.float32)
    y = np.array([1.0, 0.0, 1.0, 1.0], np.float32)
    z = np.array([1.0, 1.0, 0.0, 2.0], np.float32)
    t = np.zeros(1, np.float32)
    got = oracle.pot_sum_f64(m, x, y, z, t, t, t)
    assert got[0] == pytest.approx(1.0 / math.sqrt(6.0), rel=1e-15)
    got32 = oracle.pot_sum_f32(m, x, y, z, t, t, t)
    assert got32[0] == pytest.approx(1.0 / math.sqrt(6.0), rel=1e-6)


def test_linearity_scaling_and_permutation():
    rng = np.random.default_rng(3)
    p = rng.normal(size=(300, 3)).astype(np.float32)
    m1 = rng.uniform(1, 2, 300).astype(np.float32)
    m2 = rng.uniform(1, 2, 300).astype(np.float32)
    args = (p[:, 0], p[:, 1], p[:, 2], p[:, 0], p[:, 1], p[:, 2])
    s1 = oracle.pot_sum_f64(m1, *args)
    s2 = oracle.pot_sum_f64(m2, *args)
    s12 = oracle.pot_sum_f64((m1.astype(np.float64) + m2).astype(np.float32), *args)
    np.testing.assert_allclose(s12, s1 + s2, rtol=1e-7)  # masses re-rounded to FP32
    lam = np.float32(4.0)  # power of two: positions scale exactly
    q = p * lam
    sl = oracle.pot_sum_f64(m1, q[:, 0], q[:, 1], q[:, 2], q[:, 0], q[:, 1], q[:, 2])
    np.testing.assert_allclose(sl * 4.0, s1, rtol=1e-13)
    perm = rng.permutation(300)
    sp = oracle.pot_sum_f64(m1[perm], p[perm, 0], p[perm, 1], p[perm, 2], *args[3:])
    np.testing.assert_allclose(sp, s1, rtol=1e-13)


def test_f64_oracle_against_fsum_of_terms():
    rng = np.random.default_rng(11)
    p = rng.normal(size=(257, 3)).astype(np.float32)
    m = rng.uniform(1, 2, 257).astype(np.float32)
    got = oracle.pot_sum_f64(m, p[:, 0], p[:, 1], p[:, 2], p[:, 0], p[:, 1], p[:, 2])
    P = p.astype(np.float64)
    for i in (0, 100, 256):
        d = P - P[i]
        keep = (d != 0).all(axis=1)
        terms = m[keep].astype(np.float64) / np.sqrt((d[keep] ** 2).sum(axis=1))
        assert got[i] == pytest.approx(math.fsum(terms), rel=1e-14)


def test_plummer_binding_energy_statistical():
    from pyhalma_amd import synth

    rng = np.random.default_rng(5)
    n, a, mp = 4000, 0.005, 1e5
    pos = synth.plummer_positions(rng, n, a).astype(np.float32)
    m = np.full(n, mp, np.float32)
    S = oracle.pot_sum_f64(m, pos[:, 0], pos[:, 1], pos[:, 2], pos[:, 0], pos[:, 1],
                           pos[:, 2], nthreads=0)
    G = oracle.G_CONST
    W = -0.5 * G * mp * S.sum()
    M = n * mp
    W_analytic = -3 * math.pi * G * M * M / (32 * a)
    assert W == pytest.approx(W_analytic, rel=0.05)


def test_g_constant_bits():
    g = 4.3 * 1e-3
    g *= 1e-6
    assert oracle.G_CONST == g
    assert oracle.G_CONST == pytest.approx(4.3e-9, rel=1e-15)
