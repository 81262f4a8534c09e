"""Autotuner for the chunked FP32 pair kernel's tile shape and work decomposition.

The knobs (``halma_tune_set``): tests per lane R (1, 2, 4), FP32 sub-tile length SUB
(4, 8), persistent waves per CU, queue items per wave and the smallest source chunk.
``autotune`` times the batched pair sum of the caller's haloes (every member a test and
a source) for candidate settings; ``autotune_unbind`` times the whole iterated
unbinding of the caller's groups, whose incremental passes are where the chunked kernel
runs (full passes take the symmetric kernel, which has no knobs).  Both search one knob
at a time from the current setting (coordinate descent, two rounds), timing each
candidate with HIP events on the current stream (median of ``reps`` calls after one
warm-up), and keep the fastest.  Results can be cached in a JSON file keyed by the
device name and a signature of the halo sizes.
"""
import json
import os

import numpy as np
import torch

from . import _lib, engine

KNOBS = (
    ("tests_per_lane", _lib.TUNE_TESTS_PER_LANE, (1, 2, 4)),
    ("subtile", _lib.TUNE_SUBTILE, (4, 8)),
    ("waves_per_cu", _lib.TUNE_WAVES_PER_CU, (8, 12, 16, 24, 32)),
    ("items_per_wave", _lib.TUNE_ITEMS_PER_WAVE, (2, 4, 8)),
    ("min_chunk", _lib.TUNE_MIN_CHUNK, (256, 512, 1024, 2048)),
)


def current():
    return {name: engine.tune_get(key) for name, key, _ in KNOBS}


def apply(setting):
    for name, key, _ in KNOBS:
        if name in setting:
            engine.tune_set(key, int(setting[name]))


def _signature(offsets):
    sizes = np.diff(np.asarray(offsets, np.int64))
    return (f"H{len(sizes)}_n{int(sizes.sum())}_p{float((sizes.astype(np.float64) ** 2).sum()):.3e}"
            f"_max{int(sizes.max()) if len(sizes) else 0}")


def _time_setting(setting, run, reps):
    apply(setting)
    run()  # warm-up
    times = []
    for _ in range(reps):
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        run()
        e1.record()
        e1.synchronize()
        times.append(e0.elapsed_time(e1))
    return float(np.median(times))


def autotune(offsets, m, x, y, z, reps=3, rounds=2, cache=None, verbose=False):
    """Tune the batched pair sum on the haloes given by CSR host ``offsets`` over
    member-ordered FP32 device arrays (x, y, z, m).  Leaves the best setting applied and
    returns (setting, ms, log).  ``cache``: path of a JSON file to read and update."""
    offsets = np.asarray(offsets, np.int64)
    m, x, y, z = (engine.to_dev(a, np.float32) for a in (m, x, y, z))

    def run():
        engine.pot_sum_batched(offsets, m, x, y, z, offsets, x, y, z)

    return _search(run, f"pot|{_signature(offsets)}", reps, rounds, cache, verbose)


def autotune_unbind(groups, x, y, z, vx, vy, vz, m, factor=1.0, reps=3, rounds=2, cache=None,
                    verbose=False):
    """Tune on the iterated unbinding of ``groups`` (an ``engine.Groups`` over the given
    device arrays, as ``engine.unbind`` takes them).  Leaves the best setting applied and
    returns (setting, ms, log)."""

    def run():
        engine.unbind(groups, x, y, z, vx, vy, vz, m, factor=factor)

    return _search(run, f"unbind|{_signature(groups.offsets_host)}", reps, rounds, cache,
                   verbose)


def _search(run, signature, reps, rounds, cache, verbose):
    key = None
    if cache is not None:
        key = f"{torch.cuda.get_device_name(0)}|{signature}"
        try:
            with open(cache) as f:
                known = json.load(f)
        except (OSError, ValueError):
            known = {}
        if key in known:
            apply(known[key]["setting"])
            return known[key]["setting"], known[key]["ms"], []
    best = current()
    best_ms = _time_setting(best, run, reps)
    log = [(dict(best), best_ms)]
    for _ in range(rounds):
        improved = False
        for name, _, values in KNOBS:
            for v in values:
                if v == best[name]:
                    continue
                cand = dict(best)
                cand[name] = v
                ms = _time_setting(cand, run, reps)
                log.append((cand, ms))
                if verbose:
                    print(f"autotune {cand} {ms:.3f} ms", flush=True)
                if ms < best_ms * 0.99:
                    best, best_ms, improved = cand, ms, True
        if not improved:
            break
    apply(best)
    if cache is not None:
        try:
            with open(cache) as f:
                known = json.load(f)
        except (OSError, ValueError):
            known = {}
        known[key] = {"setting": best, "ms": best_ms}
        os.makedirs(os.path.dirname(os.path.abspath(cache)), exist_ok=True)
        with open(cache, "w") as f:
            json.dump(known, f, indent=1)
    return best, best_ms, log
