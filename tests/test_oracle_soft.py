"""Pins the softened-potential oracles (CPU only): a single source in closed form, the
eps -> 0 limit against the (pinned) unsoftened FP64 oracle, and the gradient oracle
against central finite differences of the potential oracle."""
import math

import numpy as np

import oracle


def test_single_source_closed_form():
    m = np.array([3.0e9], np.float32)
    src = [np.array([v], np.float32) for v in (0.25, -0.5, 0.125)]
    t = [np.array([1.0]), np.array([0.5]), np.array([-0.375])]
    eps = 0.0625
    d = [1.0 - 0.25, 0.5 + 0.5, -0.375 - 0.125]
    r2 = sum(v * v for v in d)
    S = oracle.pot_soft_f64(m, *src, *t, eps)
    assert S[0] == float(m[0]) / math.sqrt(r2 + eps * eps)
    a = oracle.acc_soft_f64(m, *src, *t, eps)
    w = float(m[0]) * (r2 + eps * eps) ** -1.5
    np.testing.assert_allclose(a[0], [w * (0.25 - 1.0), w * (-0.5 - 0.5), w * (0.125 + 0.375)],
                               rtol=1e-15)


def test_eps_zero_matches_unsoftened_oracle():
    rng = np.random.default_rng(4)
    x, y, z = (rng.normal(size=400).astype(np.float32) for _ in range(3))
    m = rng.uniform(1e4, 1e6, 400).astype(np.float32)
    S0 = oracle.pot_soft_f64(m, x, y, z, x.astype(np.float64), y.astype(np.float64),
                             z.astype(np.float64), 0.0)
    S = oracle.pot_sum_f64(m, x, y, z, x, y, z)
    np.testing.assert_allclose(S0, S, rtol=1e-13)


def test_gradient_is_the_derivative_of_the_potential():
    rng = np.random.default_rng(7)
    x, y, z = (rng.normal(scale=0.01, size=300).astype(np.float32) for _ in range(3))
    m = rng.uniform(1e4, 1e6, 300).astype(np.float32)
    eps = 0.002
    t = rng.normal(scale=0.01, size=(5, 3))
    a = oracle.acc_soft_f64(m, x, y, z, t[:, 0], t[:, 1], t[:, 2], eps)
    h = 1e-6
    for k in range(3):
        tp, tm = t.copy(), t.copy()
        tp[:, k] += h
        tm[:, k] -= h
        Sp = oracle.pot_soft_f64(m, x, y, z, tp[:, 0], tp[:, 1], tp[:, 2], eps)
        Sm = oracle.pot_soft_f64(m, x, y, z, tm[:, 0], tm[:, 1], tm[:, 2], eps)
        fd = (Sp - Sm) / (2 * h)
        np.testing.assert_allclose(a[:, k], fd, rtol=1e-5, atol=1e-6 * np.abs(a).max())
