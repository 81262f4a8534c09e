"""Full-size checks of the measured catalogues on the GPU.

* FoF labels of the 2.1M-particle clump catalogue and of the 8M-particle bridged pairs
  equal the oracle's (its sub-cell C FoF, pinned against scipy's KD-tree, the plain
  cell-list oracle and brute force), particle for particle;
* on the clump catalogue, every S the batched unbinding keeps -- the first pass's (the
  symmetric kernel's) for every member, the last pass's for the members bound at the
  end -- is within 1e-6 of the FP64 direct sum over the same set (the library's FP64
  fixed-order sum, bit-identical to the oracle's C sum), and the bound masks and pass
  counts equal those of the plain FP64 comparator.
"""
import numpy as np
import pytest

import oracle

pytestmark = [pytest.mark.gpu, pytest.mark.slow]


@pytest.fixture(scope="module")
def clumps():
    from pyhalma_amd import synth

    return synth.clumps()


def test_fof_labels_clumps_full_size(gpu, clumps):
    from pyhalma_amd import engine

    w = clumps
    lab = engine.fof_labels(w["x"], w["y"], w["z"], w["ll"]).cpu().numpy()
    want = oracle.fof_labels_cells(w["x"], w["y"], w["z"], w["ll"])
    print(f"n={len(lab)} components={len(np.unique(want))}", flush=True)
    assert np.array_equal(lab, want)


def test_fof_labels_bridged_full_size(gpu):
    from pyhalma_amd import engine, synth

    w = synth.bridged_pairs()
    lab = engine.fof_labels(w["x"], w["y"], w["z"], w["ll"]).cpu().numpy()
    want = oracle.fof_labels_cells(w["x"], w["y"], w["z"], w["ll"])
    print(f"n={len(lab)} components={len(np.unique(want))}", flush=True)
    assert np.array_equal(lab, want)


def _max_rel(groups, md, pot, bound):
    from pyhalma_amd import engine

    off = groups.offsets_host
    worst = 0.0
    for h in range(groups.n_groups):
        a, b = int(off[h]), int(off[h + 1])
        sel = bound[a:b].bool()
        if int(sel.sum()) == 0:
            continue
        mm, xx, yy, zz = (t[a:b][sel] for t in (md.m, md.x, md.y, md.z))
        s64 = engine.pot_sum_f64_ordered(mm, xx, yy, zz, xx, yy, zz)
        worst = max(worst, float(((pot[a:b][sel] - s64).abs() / s64).max()))
    return worst


def test_unbinding_potential_error_full_size(gpu, clumps):
    import torch

    from pyhalma_amd import engine

    w = clumps
    dev = torch.device("cuda", 0)
    D = {k: torch.from_numpy(w[k]).to(dev) for k in ("x", "y", "z", "vx", "vy", "vz", "mass")}
    groups = engine.fof_groups(D["x"], D["y"], D["z"], w["ll"], 50)[1]
    md = engine.gather_members(groups, D["x"], D["y"], D["z"], D["mass"], D["vx"], D["vy"],
                               D["vz"])
    g = md.groups
    args = (g, md.x, md.y, md.z, md.vx, md.vy, md.vz, md.m)
    first = engine.unbind(*args, factor=1.0, max_iter=1, with_pot=True)
    e1 = _max_rel(g, md, first.pot, torch.ones_like(first.bound))
    r = engine.unbind(*args, factor=1.0, with_pot=True)
    e2 = _max_rel(g, md, r.pot, r.bound)
    pb, piters = engine.unbind_plain_f64(*args, factor=1.0)
    print(f"haloes={g.n_groups} first-pass max {e1:.2e} last-pass max {e2:.2e} "
          f"passes={r.stats['passes']} resummed={r.stats['resummed']}", flush=True)
    assert e1 <= 1e-6 and e2 <= 1e-6
    assert torch.equal(pb, r.bound)
    assert np.array_equal(piters, r.iters.cpu().numpy())
