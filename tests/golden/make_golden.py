"""Writes the golden fixtures under tests/golden/ (run in the build container, where the
reference tree is mounted; the fixtures, not the reference, travel to the GPU host).

  pot_ref.npz  inputs and outputs of the reference's own brute_force_binding_energy
               (particle_subroutines.f90:466-514), compiled from its Fortran source into
               oracle/_ref by `make -C oracle ref` and called through ctypes.
  fof_ref.npz  inputs, linking lengths and FoF labels from scipy's cKDTree (the
               reference's own neighbour-search dependency, fof.py:8,20) through
               connected components; labels are the minimum member index.

    make -C oracle all ref && python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle  # noqa: E402


def _plummer(rng, n, a):
    u = rng.uniform(1e-9, 1.0, n)
    r = a / np.sqrt(u ** (-2.0 / 3.0) - 1.0)
    v = rng.normal(size=(n, 3))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    return r[:, None] * v


def pot_cases(rng):
    cases = {}
    # 1. self potential of a Plummer clump (tests == sources; self pairs skipped)
    p = _plummer(rng, 700, 0.005).astype(np.float32)
    m = np.exp(rng.uniform(np.log(1e4), np.log(1e6), 700)).astype(np.float32)
    cases["self"] = (m, p[:, 0], p[:, 1], p[:, 2], p[:, 0], p[:, 1], p[:, 2])
    # 2. distinct test and source sets (gas tests against star sources)
    s = _plummer(rng, 513, 0.01).astype(np.float32)
    t = (_plummer(rng, 257, 0.02) + 0.003).astype(np.float32)
    ms = np.full(513, 1e5, np.float32)
    cases["cross"] = (ms, s[:, 0], s[:, 1], s[:, 2], t[:, 0], t[:, 1], t[:, 2])
    # 3. lattice: many pairs share one or two coordinates (skip rule of :499-501)
    g = np.arange(6, dtype=np.float32) * np.float32(0.125)
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    lx, ly, lz = X.ravel(), Y.ravel(), Z.ravel()
    ml = rng.uniform(1.0, 2.0, lx.size).astype(np.float32)
    cases["lattice"] = (ml, lx, ly, lz, lx, ly, lz)
    # 4. duplicated positions and a test set that shares coordinates with sources
    d = _plummer(rng, 100, 0.01).astype(np.float32)
    d = np.concatenate([d, d[:30]])
    md = rng.uniform(1e4, 1e5, len(d)).astype(np.float32)
    tt = d[::3].copy()
    tt[::2, 0] = np.float32(0.001)  # x equal to nothing; y,z equal to a source's
    cases["dups"] = (md, d[:, 0], d[:, 1], d[:, 2], tt[:, 0], tt[:, 1], tt[:, 2])
    return cases


def fof_cases(rng):
    cases = {}
    # clumps + field
    c = [(_plummer(rng, n, 0.002) + rng.uniform(-0.2, 0.2, 3)) for n in (300, 150, 60, 7)]
    f = rng.uniform(-0.25, 0.25, size=(400, 3))
    x = np.concatenate(c + [f])
    x = x[rng.permutation(len(x))]
    cases["clumps_f64"] = (x.astype(np.float64), 0.0035)
    cases["clumps_f32"] = (x.astype(np.float32), 0.0035)
    # exact ties: points spaced exactly ll apart along axes and 3-4-5 diagonals
    ll = 0.5
    pts = [(0.0, 0.0, 0.0), (0.5, 0.0, 0.0), (1.0, 0.0, 0.0), (1.3, 0.4, 0.0),
           (3.0, 3.0, 3.0), (3.0, 3.5, 3.0), (3.0, 3.5, 3.5000001), (5.0, 5.0, 5.0),
           (5.0, 5.0, 5.0), (7.0, 0.0, 0.0), (7.3, 0.4, 0.0)]
    cases["ties"] = (np.array(pts, np.float64), ll)
    return cases


def main():
    rng = np.random.default_rng(20261202)
    if not oracle.have_ref():
        raise SystemExit("oracle/_ref not built: make -C oracle ref")
    out = {}
    for name, (m, x, y, z, tx, ty, tz) in pot_cases(rng).items():
        ref = oracle.ref_pot_sum_f32(m, x, y, z, tx, ty, tz, nthreads=1)
        for k, v in zip(("m", "x", "y", "z", "tx", "ty", "tz"), (m, x, y, z, tx, ty, tz)):
            out[f"{name}_{k}"] = v
        out[f"{name}_ref"] = ref
    np.savez_compressed(os.path.join(HERE, "pot_ref.npz"), **out)

    out = {}
    for name, (x, ll) in fof_cases(rng).items():
        lab = oracle.fof_labels_scipy(x[:, 0], x[:, 1], x[:, 2], ll)
        out[f"{name}_x"] = x
        out[f"{name}_ll"] = np.float64(ll)
        out[f"{name}_labels"] = lab
    np.savez_compressed(os.path.join(HERE, "fof_ref.npz"), **out)
    print("wrote", os.listdir(HERE))


if __name__ == "__main__":
    main()
