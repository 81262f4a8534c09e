"""Pair-potential parity on the GPU (libhalma through its C ABI, via pyhalma_amd).

Bars:
  * FP32 compensated sums: |S_gpu - S_64| <= B * S_gpu with B = halma_pot_error_bound()
    (rigorous), and the observed error <= 1e-6 relative;
  * FP64 ordered sums: bit-identical to the C oracle (same fixed summation order);
  * plain FP64 comparator: within 1e-13 of the oracle (different order);
  * the reference's own FP32 outputs (golden fixture): within FP32 noise.
"""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

GOLD_CASES = ("self", "cross", "lattice", "dups")


def _case(g, name):
    return tuple(g[f"{name}_{k}"] for k in ("m", "x", "y", "z", "tx", "ty", "tz"))


@pytest.fixture(scope="module")
def gold(golden_dir):
    return np.load(f"{golden_dir}/pot_ref.npz")


def _rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return np.abs(a - b) / np.maximum(np.abs(b), 1e-300)


@pytest.mark.parametrize("name", GOLD_CASES)
def test_pot_f32_golden_and_oracle(gpu, gold, name):
    from pyhalma_amd import engine

    args = _case(gold, name)
    print(f"case={name} n_src={len(args[0])} n_test={len(args[4])}", flush=True)
    out32, out64 = engine.pot_sum(*args, with_f64=True)
    out32 = out32.cpu().numpy()
    out64 = out64.cpu().numpy()
    ref64 = oracle.pot_sum_f64(*args)
    B = engine.pot_error_bound()
    rel = _rel(out64, ref64)
    assert rel.max() <= B, (rel.max(), B)
    assert rel.max() <= 1e-6
    np.testing.assert_allclose(out32, gold[f"{name}_ref"], rtol=max(1e-5, len(args[0]) * 6e-8))
    assert out32.dtype == np.float32


@pytest.mark.parametrize("name", GOLD_CASES)
def test_pot_f64_ordered_bitwise(gpu, gold, name):
    from pyhalma_amd import engine

    args = _case(gold, name)
    got = engine.pot_sum_f64_ordered(*args).cpu().numpy()
    want = oracle.pot_sum_f64(*args)
    assert np.array_equal(got.view(np.uint64), want.view(np.uint64)), _rel(got, want).max()


def test_pot_f64_plain(gpu, gold):
    from pyhalma_amd import engine

    args = _case(gold, "self")
    got = engine.pot_sum_f64_plain(*args).cpu().numpy()
    want = oracle.pot_sum_f64(*args)
    assert _rel(got, want).max() < 1e-13


@pytest.mark.parametrize("ns,nt", [(1, 1), (2, 2), (63, 65), (129, 1), (1000, 777),
                                   (4099, 3001)])
def test_pot_sizes(gpu, ns, nt):
    from pyhalma_amd import engine

    rng = np.random.default_rng(ns * 7 + nt)
    s = rng.normal(size=(ns, 4)).astype(np.float32)
    s[:, 3] = rng.uniform(1e4, 1e6, ns)
    t = rng.normal(size=(nt, 3)).astype(np.float32)
    args = (s[:, 3], s[:, 0], s[:, 1], s[:, 2], t[:, 0], t[:, 1], t[:, 2])
    print(f"ns={ns} nt={nt}", flush=True)
    got = engine.pot_sum(*args, out_dtype=np.float64).cpu().numpy()
    want = oracle.pot_sum_f64(*args)
    assert _rel(got, want).max() <= engine.pot_error_bound()


def test_pot_tuning_variants(gpu):
    from pyhalma_amd import engine, _lib

    rng = np.random.default_rng(99)
    p = rng.normal(size=(2500, 3)).astype(np.float32)
    m = rng.uniform(1, 2, 2500).astype(np.float32)
    args = (m, p[:, 0], p[:, 1], p[:, 2], p[:, 0], p[:, 1], p[:, 2])
    want = oracle.pot_sum_f64(*args)
    keep = {k: engine.tune_get(k) for k in (_lib.TUNE_TESTS_PER_LANE, _lib.TUNE_SUBTILE,
                                            _lib.TUNE_MIN_CHUNK)}
    try:
        for R in (1, 2, 4):
            for sub in (4, 8):
                for chunk in (64, 512):
                    engine.tune_set(_lib.TUNE_TESTS_PER_LANE, R)
                    engine.tune_set(_lib.TUNE_SUBTILE, sub)
                    engine.tune_set(_lib.TUNE_MIN_CHUNK, chunk)
                    print(f"R={R} sub={sub} chunk={chunk}", flush=True)
                    got = engine.pot_sum(*args, out_dtype=np.float64).cpu().numpy()
                    assert _rel(got, want).max() <= engine.pot_error_bound()
    finally:
        for k, v in keep.items():
            engine.tune_set(k, v)


def test_pot_batched_matches_per_halo(gpu):
    from pyhalma_amd import engine

    rng = np.random.default_rng(5)
    sizes = [1, 70, 300, 0, 1500, 64, 129]
    tsizes = [3, 70, 11, 5, 1500, 0, 200]
    so = np.concatenate([[0], np.cumsum(sizes)])
    to = np.concatenate([[0], np.cumsum(tsizes)])
    s = rng.normal(size=(so[-1], 4)).astype(np.float32)
    s[:, 3] = rng.uniform(1, 5, so[-1])
    t = rng.normal(size=(to[-1], 3)).astype(np.float32)
    got = engine.pot_sum_batched(so, s[:, 3], s[:, 0], s[:, 1], s[:, 2], to, t[:, 0],
                                 t[:, 1], t[:, 2]).cpu().numpy()
    B = engine.pot_error_bound()
    for h in range(len(sizes)):
        a, b = so[h], so[h + 1]
        c, d = to[h], to[h + 1]
        if d == c:
            continue
        want = oracle.pot_sum_f64(s[a:b, 3], s[a:b, 0], s[a:b, 1], s[a:b, 2], t[c:d, 0],
                                  t[c:d, 1], t[c:d, 2])
        if b == a:
            assert np.all(got[c:d] == 0)
        else:
            assert _rel(got[c:d], want).max() <= B, h


def test_particle_dropin_semantics(gpu):
    from pyhalma_amd import particle as particle_mod
    from pyhalma_amd import halo_gas

    P = particle_mod.particle
    rng = np.random.default_rng(1)
    x = rng.normal(size=50)
    m = rng.uniform(1, 2, 50)
    out = P.brute_force_binding_energy(8, 50, m, x, x * 2, x * 3, 50, x, x * 2, x * 3)
    assert out.dtype == np.float32 and out.shape == (50,)
    with pytest.raises(ValueError):
        P.brute_force_binding_energy(8, 49, m, x, x, x, 50, x, x, x)
    empty = halo_gas.brute_force_binding_energy_fortran(m, x, x, x, [], [], [])
    assert isinstance(empty, np.ndarray) and empty.size == 0 and empty.dtype == np.float64
    s = P.serial_brute_force_binding_energy(50, m, x, x * 2, x * 3, 50, x, x * 2, x * 3)
    np.testing.assert_array_equal(s, out)
