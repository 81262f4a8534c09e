"""Friends-of-friends parity on the GPU: labels bit-identical to the oracle (minimum
member index per component) on the golden fixture, random/clustered/tie/duplicate
inputs in FP32 and FP64, the clumped and bridged workloads at test sizes, and repeat
runs (the lock-free union-find must give the same partition every time)."""
import zlib

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def _gpu_labels(x, y, z, ll):
    from pyhalma_amd import engine

    return engine.fof_labels(x, y, z, ll).cpu().numpy().astype(np.int64)


def test_fof_golden(gpu, golden_dir):
    g = np.load(f"{golden_dir}/fof_ref.npz")
    for name in ("clumps_f64", "clumps_f32", "ties"):
        x = g[f"{name}_x"]
        ll = float(g[f"{name}_ll"])
        print(f"case={name} n={len(x)}", flush=True)
        got = _gpu_labels(x[:, 0], x[:, 1], x[:, 2], ll)
        assert np.array_equal(got, g[f"{name}_labels"]), name


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("kind", ["uniform", "percolating", "clustered", "dups", "lattice",
                                  "chain"])
def test_fof_vs_oracle(gpu, dtype, kind):
    rng = np.random.default_rng(zlib.crc32(kind.encode()))
    if kind == "uniform":
        d, ll = rng.uniform(0, 1, (5000, 3)), 0.02
    elif kind == "percolating":
        d = rng.uniform(0, 1, (20000, 3))
        ll = 0.8 * (1.0 / 20000) ** (1 / 3)
    elif kind == "clustered":
        d = np.concatenate([rng.normal(rng.uniform(-1, 1, 3), 0.01, (n, 3))
                            for n in (3000, 1000, 500, 50, 5)] + [rng.uniform(-1, 1, (3000, 3))])
        ll = 0.004
    elif kind == "dups":
        b = rng.uniform(0, 1, (3000, 3))
        d = np.concatenate([b, b[:1000], b[:10], np.repeat(b[:1], 500, 0)])
        ll = 0.01
    elif kind == "lattice":
        g = np.arange(30) * 0.125
        X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
        d = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1)
        d = np.concatenate([d, d + [0.0625, 0.125, 0.0]])
        ll = 0.125
    else:  # exact-ll chain along a diagonal plus near misses
        t = np.arange(2000)[:, None] * np.array([[0.375, 0.5, 0.0]])  # step 0.625
        d = np.concatenate([t, t + [0, 0, 0.62500001]])
        ll = 0.625
    d = d.astype(dtype)
    d = d[rng.permutation(len(d))]
    want = oracle.fof_labels_c(d[:, 0], d[:, 1], d[:, 2], ll)
    print(f"kind={kind} dtype={np.dtype(dtype).name} n={len(d)} groups="
          f"{len(np.unique(want))}", flush=True)
    got = _gpu_labels(d[:, 0], d[:, 1], d[:, 2], ll)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_fof_long_z_columns(gpu, dtype):
    """Lines of particles along z, so that one column of sub-cells holds hundreds of cells
    and the link lanes cut it into segments: links must cross every segment boundary.
    Spacings below, at and just above ll (inclusive test; the margin of 1e-3 ll outlives
    FP32 rounding of the z values), dense runs (many particles per cell)
    split by one gap wider than ll, and pairs of lines half a linking length apart in x
    (links between neighbouring columns).  The group counts are fixed by construction;
    the labels must equal the oracle's."""
    ll = 0.5
    lines, expect = [], 0
    for j, (step, n, linked) in enumerate([(0.95 * ll, 300, True), (ll, 300, True),
                                          (ll * 1.001, 300, False),
                                          (0.3 * ll, 400, True)]):
        z = np.arange(n) * step
        lines.append(np.stack([np.full(n, 100.0 * j), np.zeros(n), z], 1))
        expect += 1 if linked else n
    # dense runs: 40 particles in each of 30 consecutive sub-cells, then a gap of 1.2 ll
    rng = np.random.default_rng(7)
    side = ll / np.sqrt(3.0)
    zr = np.concatenate([c * side + rng.uniform(0.05, 0.95, 40) * side for c in range(30)])
    run = np.stack([np.full(zr.size, 500.0), np.full(zr.size, 0.01), zr], 1)
    lines += [run, run + [0.0, 0.0, zr.max() + 1.2 * ll - zr.min()]]
    expect += 2
    # two lines 0.5 ll apart in x, each with gaps of 1.5 ll that only the other bridges
    za = np.arange(0, 200) * 1.5 * ll
    lines += [np.stack([np.full(za.size, 900.0), np.zeros(za.size), za], 1),
              np.stack([np.full(za.size, 900.0 + 0.5 * ll), np.zeros(za.size),
                        za + 0.75 * ll], 1)]
    expect += 1
    d = np.concatenate(lines).astype(dtype)
    d = d[rng.permutation(len(d))]
    want = oracle.fof_labels_c(d[:, 0], d[:, 1], d[:, 2], ll)
    print(f"dtype={np.dtype(dtype).name} n={len(d)} groups={len(np.unique(want))}", flush=True)
    assert len(np.unique(want)) == expect
    got = _gpu_labels(d[:, 0], d[:, 1], d[:, 2], ll)
    assert np.array_equal(got, want)


def test_fof_clumps_workload_and_repeat(gpu):
    from pyhalma_amd import synth

    w = synth.clumps(seed=3, scale=0.03)
    x, y, z, ll = w["x"], w["y"], w["z"], w["ll"]
    print(f"clumps n={len(x)} ll={ll:.4g}", flush=True)
    want = oracle.fof_labels_c(x, y, z, ll)
    for _ in range(3):
        assert np.array_equal(_gpu_labels(x, y, z, ll), want)


def test_fof_bridged_pairs_merge(gpu):
    from pyhalma_amd import synth

    w = synth.bridged_pairs(seed=5, scale=0.01)
    x, y, z, ll = w["x"], w["y"], w["z"], w["ll"]
    print(f"bridged n={len(x)} ll={ll:.4g}", flush=True)
    want = oracle.fof_labels_c(x, y, z, ll)
    got = _gpu_labels(x, y, z, ll)
    assert np.array_equal(got, want)
    # every pair's two spheres and bridge end up in one group holding most of the pair
    pid = w["pair_id"]
    for p in np.unique(pid)[:20]:
        lab, cnt = np.unique(got[pid == p], return_counts=True)
        assert cnt.max() > 0.5 * cnt.sum()


def test_fof_errors(gpu):
    from pyhalma_amd import _lib, engine

    x = np.array([0.0, 1e9, 2.0])
    with pytest.raises(_lib.HalmaError, match="2\\^21"):
        engine.fof_labels(x, x, x, 1e-3)
    with pytest.raises(_lib.HalmaError, match="not finite"):
        engine.fof_labels(np.array([0.0, np.nan]), np.zeros(2), np.zeros(2), 0.1)
    with pytest.raises(_lib.HalmaError, match="linking length"):
        engine.fof_labels(x, x, x, 0.0)
    assert engine.fof_labels(np.zeros(0), np.zeros(0), np.zeros(0), 1.0).numel() == 0


def test_fof_path_graph_repeats(gpu):
    """Regression: a path graph (every link unique, exact-ll ties) must come out whole on
    every run.  A root test on a stale parent value once split such chains."""
    from pyhalma_amd import engine

    t = np.arange(2000)[:, None] * np.array([[0.375, 0.5, 0.0]])
    d = np.concatenate([t, t + [0, 0, 0.62500001]])
    d = d[np.random.default_rng(1).permutation(len(d))]
    want = oracle.fof_labels_c(d[:, 0], d[:, 1], d[:, 2], 0.625)
    for rep in range(10):
        got = engine.fof_labels(d[:, 0], d[:, 1], d[:, 2], 0.625).cpu().numpy()
        assert np.array_equal(got, want), rep
