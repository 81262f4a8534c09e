"""The pair-kernel autotuner: it never returns a setting slower than the one it started
from, it leaves that setting applied, and sums under it stay within the rigorous bound
of the FP64 oracle."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def test_autotune_picks_no_slower_setting_and_stays_exact(gpu, tmp_path):
    from pyhalma_amd import autotune, engine, synth

    w = synth.clumps(seed=31, scale=0.02)
    lab = oracle.fof_labels_c(w["x"], w["y"], w["z"], w["ll"])
    g = engine.groups_from_labels(lab.astype(np.int32), 50)
    md = engine.gather_members(g, w["x"], w["y"], w["z"], w["mass"])
    before = autotune.current()
    try:
        cache = str(tmp_path / "tune.json")
        setting, ms, log = autotune.autotune(md.groups.offsets_host, md.m, md.x, md.y, md.z,
                                             reps=2, rounds=1, cache=cache)
        print(f"default {log[0][1]:.3f} ms -> {setting} {ms:.3f} ms ({len(log)} trials)",
              flush=True)
        assert ms <= log[0][1]
        assert autotune.current() == setting
        # the cache returns the same setting without timing anything
        s2, ms2, log2 = autotune.autotune(md.groups.offsets_host, md.m, md.x, md.y, md.z,
                                          cache=cache)
        assert s2 == setting and log2 == []
        off = md.groups.offsets_host
        x, y, z, m = (t.cpu().numpy() for t in (md.x, md.y, md.z, md.m))
        S = engine.pot_sum_batched(off, m, x, y, z, off, x, y, z).cpu().numpy()
        for h in (0, len(off) // 2, len(off) - 2):
            a, b = off[h], off[h + 1]
            want = oracle.pot_sum_f64(m[a:b], x[a:b], y[a:b], z[a:b], x[a:b], y[a:b], z[a:b])
            assert np.all(np.abs(S[a:b] - want) <= engine.pot_error_bound() * want)
    finally:
        autotune.apply(before)


def test_autotune_unbind_keeps_masks(gpu):
    """The unbinding tuner returns a setting no slower than its start, leaves it applied,
    and the unbinding under it still matches the oracle's masks."""
    from pyhalma_amd import autotune, engine, synth

    w = synth.clumps(seed=32, scale=0.01)
    lab = oracle.fof_labels_c(w["x"], w["y"], w["z"], w["ll"])
    g = engine.groups_from_labels(lab.astype(np.int32), 50)
    md = engine.gather_members(g, w["x"], w["y"], w["z"], w["mass"], w["vx"], w["vy"], w["vz"])
    before = autotune.current()
    try:
        setting, ms, log = autotune.autotune_unbind(md.groups, md.x, md.y, md.z, md.vx, md.vy,
                                                    md.vz, md.m, reps=2, rounds=1)
        print(f"default {log[0][1]:.3f} ms -> {setting} {ms:.3f} ms ({len(log)} trials)",
              flush=True)
        assert ms <= log[0][1]
        assert autotune.current() == setting
        r = engine.unbind(g, w["x"], w["y"], w["z"], w["vx"], w["vy"], w["vz"], w["mass"], 1.0)
        bound, iters, _ = oracle.unbind_iterative(
            g.offsets_host, g.members[: g.n_members].cpu().numpy(), w["x"], w["y"], w["z"],
            w["vx"], w["vy"], w["vz"], w["mass"], 1.0, 100)
        assert np.array_equal(r.bound.cpu().numpy().astype(bool), bound)
        assert np.array_equal(r.iters.cpu().numpy(), iters)
    finally:
        autotune.apply(before)
