"""Benchmark: FoF + per-group reductions + batched iterative unbinding on one GPU.

Workload: 1000 Plummer clumps in a 40 Mpc box, clump sizes p(N) ~ N^-1.385 on
[100, 20000] (~1.9M clump stars) plus 10% uniform field stars (~2.1M particles in all),
FP32 positions, velocities and log-uniform masses 1e4-1e6 Msun, linking length 0.2x the
mean interparticle spacing, groups of >= 50 members, unbinding in physical mode
(E = K + U, factor 1) iterated to convergence.  Synthetic data generated from a seed;
inputs are resident in HBM before the timed region.

One step = friends-of-friends labels -> groups -> one gather of the members' data into
group order -> group reductions (mass, centre of mass, CM velocity, furthest particle)
-> density peaks -> batched unbinding.  The plain FP64
comparator (same method: one launch per halo, FP64 sums, no batching, no graphs) runs
once after the timed steps; the CPU baseline runs the oracle's FoF port and the
reference's own FP32 pair routine (compiled from its Fortran source, OpenMP on all host
cores) on a bounded sub-box of the same workload.

    python bench.py [--gpus 1] [--steps 5] [--warmup 2]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = "particles/s FoF+unbinding at 2M stars, batched FP32 vs plain FP64 on same GPU"
PEAK_FP32_TFLOPS = 157.3   # MI355X FP32 vector (AMD published, packed FMA)
FLOP_PER_PAIR = 11         # particle_subroutines.f90:502-505 as written
# HBM bytes per k_pot_sym launch on this workload, averaged over the 9 launches of a step
# (the first pass carries almost all of it): FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0 from two
# rocprofv3 --pmc passes, each scaled by what the counter reads for a 256 MiB device copy
# (profiles/bench_step_2M_hbm_traffic.md)
SYM_TRAFFIC_BYTES = 80.9e6
MINP = 50


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--seed", type=int, default=20261204)
    ap.add_argument("--no-plain", action="store_true", help="skip the plain FP64 comparator")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the built-in pair-kernel tiling instead of autotuning it on "
                         "the workload's haloes before the timed steps")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="run one 256 MiB device copy first (scales profiler byte counters)")
    ap.add_argument("--cpu-frac", type=float, default=0.5,
                    help="fraction of the box (along x) used for the CPU baseline sample")
    ap.add_argument("--workload", default="clumps",
                    choices=["clumps", "gas", "bridged", "centring", "substructure",
                             "cleaning", "progenitors", "sources"],
                    help="clumps (default): the headline 2.1M-particle step; the others are "
                         "the gas batch, the bridged-pair FoF stress case, the "
                         "potential-minimum centring, the substructure splitting of the "
                         "bridged pairs, and the phase-space cleaning, progenitor "
                         "matching and source gathering of the clumps at full size")
    return ap.parse_args()


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def main():
    args = parse()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(the halo-finding path would shard FoF by spatial slabs and haloes by size "
              "across GPUs)",
              file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch

    from pyhalma_amd import engine, synth

    if args.workload != "clumps":
        torch.cuda.init()
        out = {"gas": bench_gas, "bridged": bench_bridged, "centring": bench_centring,
               "substructure": bench_substructure,
               "cleaning": bench_cleaning,
               "progenitors": bench_progenitors,
               "sources": bench_sources}[args.workload](args)
        print(json.dumps(out))
        return

    torch.cuda.init()
    dev = torch.device("cuda", 0)
    t0 = time.time()
    w = synth.clumps(seed=args.seed)
    n = len(w["x"])
    ll = w["ll"]
    log(f"workload: {n} particles, {len(w['sizes'])} clumps, ll={ll:.5f} Mpc "
        f"(generated in {time.time() - t0:.1f}s)")
    torch.cuda.synchronize()
    th0 = time.perf_counter()
    D = {k: torch.from_numpy(w[k]).to(dev) for k in ("x", "y", "z", "vx", "vy", "vz", "mass")}
    torch.cuda.synchronize()
    h2d_s = time.perf_counter() - th0  # first upload: also pays allocator set-up
    for k in ("x", "y", "z", "vx", "vy", "vz", "mass"):  # steady-state upload time
        D[k].copy_(torch.from_numpy(w[k]))
    torch.cuda.synchronize()
    th0 = time.perf_counter()
    for k in ("x", "y", "z", "vx", "vy", "vz", "mass"):
        D[k].copy_(torch.from_numpy(w[k]))
    torch.cuda.synchronize()
    h2d_s = min(h2d_s, time.perf_counter() - th0)
    h2d_bytes = sum(int(D[k].numel()) * D[k].element_size() for k in D)
    if args.calibrate_copy:
        src = torch.ones(64 << 20, dtype=torch.float32, device=dev)  # 256 MiB
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    def step(time_pot=False):
        lab = engine.fof_labels(D["x"], D["y"], D["z"], ll)
        groups = engine.groups_from_labels(lab, MINP)
        md = engine.gather_members(groups, D["x"], D["y"], D["z"], D["mass"], D["vx"], D["vy"],
                                   D["vz"])
        g = md.groups
        red = engine.group_reduce(g, md.x, md.y, md.z, md.m, md.vx, md.vy, md.vz)
        peak = engine.density_peak(g, md.x, md.y, md.z, md.m, ll, reduced=red)
        ub = engine.unbind(g, md.x, md.y, md.z, md.vx, md.vy, md.vz, md.m, factor=1.0,
                           time_pot=time_pot)
        return lab, groups, red, peak, ub

    tuning = None
    if not args.no_autotune:
        from pyhalma_amd import autotune
        _, groups, _, _, _ = step()
        md = engine.gather_members(groups, D["x"], D["y"], D["z"], D["mass"], D["vx"], D["vy"],
                                   D["vz"])
        t_at = time.time()
        # the chunked kernel's knobs, timed on the step's own unbinding
        setting, ms, trials = autotune.autotune_unbind(md.groups, md.x, md.y, md.z, md.vx,
                                                       md.vy, md.vz, md.m, factor=1.0, reps=5)
        tuning = {"setting": setting, "unbind_ms": ms, "trials": len(trials),
                  "default_ms": trials[0][1] if trials else ms}
        log(f"autotune: {setting} {ms:.3f} ms (default {tuning['default_ms']:.3f} ms, "
            f"{len(trials)} trials, {time.time() - t_at:.1f}s)")
        del md
    for i in range(args.warmup):
        step()
        log(f"warmup {i + 1}/{args.warmup} done")
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    acc = dict.fromkeys(("pot_ns", "pot_launches", "pairs", "sym_ns", "sym_launches",
                         "full_pairs", "f32_launches"), 0)
    for i in range(args.steps):
        lab, groups, red, peak, ub = step(time_pot=True)
        for k in acc:
            acc[k] += ub.stats[k]
    torch.cuda.synchronize()
    t_end = time.perf_counter()
    step_s = (t_end - t_start) / args.steps
    log(f"timed {args.steps} steps: {step_s * 1e3:.2f} ms/step; groups={groups.n_groups} "
        f"members={groups.n_members} passes={ub.stats['passes']} "
        f"retested={ub.stats['retested']}")

    # roofline of the dominant kernel (the symmetric pair kernel of the full passes),
    # from HIP events around its own launches.  Algorithmic work: the reference's loop
    # over the same pairs, 11 FLOP per ordered (test, source) pair; the kernel evaluates
    # each unordered pair once and accumulates both terms, so the FP32 instructions it
    # executes are about half of that count (executed_* below).
    sym_s = acc["sym_ns"] * 1e-9 / max(acc["sym_launches"], 1)
    pairs_per_launch = acc["full_pairs"] / max(acc["sym_launches"], 1)
    achieved = FLOP_PER_PAIR * pairs_per_launch / sym_s / 1e12
    f32_pairs = acc["pairs"] - acc["full_pairs"]
    f32_s = (acc["pot_ns"] - acc["sym_ns"]) * 1e-9
    roofline = {"bound": "mfma", "achieved": round(achieved, 3),
                "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                "frac": round(achieved / PEAK_FP32_TFLOPS, 4), "traffic": SYM_TRAFFIC_BYTES,
                "roof": "compute, FP32 (157.3 TFLOP/s, the dense FP32 figure); the kernel "
                        "issues FP32 vector instructions, no matrix instructions",
                "kernel": "k_pot_sym (symmetric FP32 pair kernel, full passes)",
                "pairs_per_launch": int(pairs_per_launch),
                "launch_ms": round(sym_s * 1e3, 4),
                "launches_per_step": acc["sym_launches"] / args.steps,
                "executed_pair_evals_per_launch": int(pairs_per_launch // 2),
                "chunked_kernel": {"kernel": "k_pot_f32 (incremental passes)",
                                   "launches_per_step": acc["f32_launches"] / args.steps,
                                   "ms_per_step": f32_s * 1e3 / args.steps,
                                   "tflops": (FLOP_PER_PAIR * f32_pairs / f32_s / 1e12
                                              if f32_s > 0 else None)}}

    out = {
        "metric": METRIC, "value": n / step_s, "unit": "particles/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": step_s * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded Plummer clumps + uniform field)",
        "config": {"workload": "1000 Plummer clumps in a 40 Mpc box, p(N)~N^-1.385 on "
                               "[100, 20000], +10% field, FP32, ll=0.2 mean spacing, "
                               "groups>=50, physical-mode iterative unbinding",
                   "particles": n, "groups": groups.n_groups,
                   "members": groups.n_members, "passes": ub.stats["passes"]},
        "roofline": roofline,
        "tuning": tuning,
        # inputs handed over from (pageable) host memory instead of resident in HBM
        "pcie_inclusive": {"h2d_ms": h2d_s * 1e3, "h2d_bytes": h2d_bytes,
                           "value": n / (step_s + h2d_s)},
    }

    if not args.no_plain:
        # the plain FP64 comparator on the same groups (FoF and CSR are shared: FoF has
        # one implementation; the comparator replaces reductions + unbinding)
        torch.cuda.synchronize()
        tp0 = time.perf_counter()
        lab2 = engine.fof_labels(D["x"], D["y"], D["z"], ll)
        g2 = engine.groups_from_labels(lab2, MINP)
        red2 = engine.group_reduce(g2, D["x"], D["y"], D["z"], D["mass"], D["vx"], D["vy"],
                                   D["vz"])
        pb, piters = engine.unbind_plain_f64(g2, D["x"], D["y"], D["z"], D["vx"], D["vy"],
                                             D["vz"], D["mass"], factor=1.0)
        torch.cuda.synchronize()
        plain_s = time.perf_counter() - tp0
        same_groups = torch.equal(g2.members[: g2.n_members], groups.members[: groups.n_members])
        same_masks = torch.equal(pb, ub.bound)
        same_iters = bool(np.array_equal(piters, ub.iters.cpu().numpy()))
        del red2
        out["plain_f64"] = {"ms": plain_s * 1e3, "value": n / plain_s,
                            "speedup": plain_s / step_s, "groups_identical": bool(same_groups),
                            "masks_identical": bool(same_masks),
                            "passes_identical": same_iters}
        log(f"plain FP64: {plain_s * 1e3:.1f} ms -> speedup {plain_s / step_s:.1f}x; "
            f"masks identical: {same_masks}")

    # per-halo potential error against the FP64 ordered sum (first pass, sample of haloes)
    sizes = groups.sizes()
    order = np.argsort(sizes)
    sample = np.unique(np.concatenate([order[-3:], order[:: max(1, len(order) // 12)]]))
    mem = groups.members[: groups.n_members]
    errs = []
    for h in sample:
        a, b = int(groups.offsets_host[h]), int(groups.offsets_host[h + 1])
        idx = mem[a:b].long()
        args4 = [D[k][idx] for k in ("mass", "x", "y", "z")]
        s32 = engine.pot_sum(*args4, args4[1], args4[2], args4[3], out_dtype=np.float64)
        s64 = engine.pot_sum_f64_ordered(*args4, args4[1], args4[2], args4[3])
        errs.append(float(((s32 - s64).abs() / s64).max()))
    out["pot_rel_err_max"] = max(errs)
    out["pot_error_bound"] = engine.pot_error_bound()

    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(w, args.cpu_frac)
    print(json.dumps(out))


def _timed(fn, steps, warmup):
    import torch

    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        r = fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / steps, r


def _line(metric, value, unit, step_s, args, config, **extra):
    d = {"metric": metric, "value": value, "unit": unit, "n_gpus": 1, "steps": args.steps,
         "warmup": args.warmup, "ms_per_step": step_s * 1e3, "higher_is_better": True,
         "scaling": "weak", "vs_baseline": None, "dtype": "f32",
         "data": "synthetic (seeded)", "config": config}
    d.update(extra)
    return d


def bench_gas(args):
    """Gas batch: 500 NFW gas haloes of 1e3-2e5 cells, one exact energy test per cell
    (factor 2, sources = the halo's cells, halo_gas.py:285-492 without subsampling),
    every halo in one call.  Full-size parity: a random sample of cells is re-decided
    with the FP64 fixed-order sum on the GPU."""
    import numpy as np
    import torch

    from pyhalma_amd import engine, halo_gas, synth

    t0 = time.time()
    w = synth.gas_haloes()
    off = w["offsets"]
    H = len(off) - 1
    n = int(off[-1])
    pairs = float((np.diff(off).astype(np.float64) ** 2).sum())
    log(f"gas workload: {H} haloes, {n} cells, {pairs:.3e} pairs ({time.time() - t0:.1f}s)")
    dev = torch.device("cuda", 0)
    g = [torch.from_numpy(w[k]).to(dev) for k in ("x", "y", "z", "vx", "vy", "vz", "mass",
                                                   "temp")]
    bulk = torch.from_numpy(w["bulk"]).to(dev)
    step_s, (out, bound) = _timed(lambda: halo_gas.RPS_batched(g, off, bulk), args.steps,
                                  args.warmup)
    log(f"gas: {step_s * 1e3:.1f} ms/step")
    # full-size parity on a sample: FP64 fixed-order sums of sampled cells
    rng = np.random.default_rng(1)
    sample = np.sort(rng.choice(n, 256, replace=False))
    halo = np.searchsorted(off, sample, side="right") - 1
    mism = 0
    for i, h in zip(sample, halo):
        a, b = off[h], off[h + 1]
        S = engine.pot_sum_f64_ordered(g[6][a:b], g[0][a:b], g[1][a:b], g[2][a:b],
                                       g[0][i:i + 1], g[1][i:i + 1], g[2][i:i + 1])
        dv = np.array([w[k][i] for k in ("vx", "vy", "vz")], np.float64) - w["bulk"][h]
        kin = 0.5 * ((dv[0] * dv[0] + dv[1] * dv[1]) + dv[2] * dv[2])
        E = kin - 2.0 * engine_g() * float(S[0])
        mism += int((E <= 0.0) != bool(bound[i].item()))
    return _line("gas cells/s, exact batched energy test (500 NFW haloes, 1e3-2e5 cells)",
                 n / step_s, "cells/s", step_s, args,
                 {"workload": "gas batch: 500 NFW haloes, cells log-uniform 1e3-2e5, FP32, "
                              "factor 2, sources = each halo's cells", "cells": n,
                  "haloes": H, "pairs": pairs},
                 pairs_per_s=pairs / step_s,
                 achieved_tflops=pairs * FLOP_PER_PAIR / step_s / 1e12,
                 sample_mismatches=mism, sample_size=int(len(sample)),
                 bound_fraction=float(bound.float().mean().item()))


def engine_g():
    return 4.3 * 1e-3 * 1e-6


def bench_bridged(args):
    """Bridged stress case: 8M stars in 4000 Plummer pairs joined by a 5% bridge; FoF
    labels + CSR groups + gather + reductions + iterated unbinding of the groups.  FoF
    must merge each pair: every group of >= 50 members holds exactly one pair's
    particles, checked against the generator's pair ids."""
    import numpy as np
    import torch

    from pyhalma_amd import engine, synth

    t0 = time.time()
    w = synth.bridged_pairs()
    n = len(w["x"])
    ll = w["ll"]
    log(f"bridged workload: {n} particles, ll={ll:.5f} ({time.time() - t0:.1f}s)")
    dev = torch.device("cuda", 0)
    D = {k: torch.from_numpy(w[k]).to(dev) for k in ("x", "y", "z", "vx", "vy", "vz", "mass")}

    def step():
        lab = engine.fof_labels(D["x"], D["y"], D["z"], ll)
        groups = engine.groups_from_labels(lab, MINP)
        md = engine.gather_members(groups, D["x"], D["y"], D["z"], D["mass"], D["vx"], D["vy"],
                                   D["vz"])
        red = engine.group_reduce(md.groups, md.x, md.y, md.z, md.m, md.vx, md.vy, md.vz)
        ub = engine.unbind(md.groups, md.x, md.y, md.z, md.vx, md.vy, md.vz, md.m, factor=1.0)
        return groups, red, ub

    step_s, (groups, red, ub) = _timed(step, args.steps, args.warmup)
    mem = groups.members[: groups.n_members].cpu().numpy()
    offs = groups.offsets_host
    pid = w["pair_id"][mem]
    seg_min = np.minimum.reduceat(pid, offs[:-1]) if groups.n_groups else np.array([])
    seg_max = np.maximum.reduceat(pid, offs[:-1]) if groups.n_groups else np.array([])
    pure = int((seg_min == seg_max).sum())
    pairs_hit = len(np.unique(seg_min[seg_min == seg_max]))
    log(f"bridged: {step_s * 1e3:.1f} ms/step; groups={groups.n_groups} pure={pure} "
        f"pairs found={pairs_hit}")
    return _line("particles/s FoF+unbinding, 8M-particle bridged pairs", n / step_s,
                 "particles/s", step_s, args,
                 {"workload": "4000 Plummer pairs at 1-3 a joined by 5% bridges, 8M FP32, "
                              "groups>=50, physical-mode unbinding", "particles": n,
                  "groups": groups.n_groups, "passes": ub.stats["passes"]},
                 groups_single_pair=pure, pairs_merged=pairs_hit,
                 members=groups.n_members, pairs_evaluated=ub.stats["pairs"],
                 bound_fraction=float(ub.bound.float().mean().item()))


def bench_substructure(args):
    """Substructure splitting (pyHALMA.py:703-766) of the bridged-pair workload: FoF
    labels + CSR groups at ll, then two sub-linking levels (ll/2, ll/4) over all groups
    at once, minp_sub 25, distance factor 0.75, minp 50.  CPU baseline: the oracle's
    restatement of the same block (C cell-list FoF per group and level) on the first
    groups of the same catalogue."""
    import numpy as np
    import torch

    from pyhalma_amd import engine, synth

    t0 = time.time()
    w = synth.bridged_pairs()
    n = len(w["x"])
    ll = float(w["ll"])
    min_ll = ll / 4.0 * 0.99
    log(f"substructure workload: {n} particles, ll={ll:.5f} ({time.time() - t0:.1f}s)")
    dev = torch.device("cuda", 0)
    D = {k: torch.from_numpy(w[k]).to(dev) for k in ("x", "y", "z", "mass")}

    def step():
        lab = engine.fof_labels(D["x"], D["y"], D["z"], ll)
        groups = engine.groups_from_labels(lab, MINP)
        sub = engine.split_substructure(groups, D["x"], D["y"], D["z"], D["mass"], ll,
                                        minp_sub=25, dec_factor=2.0, min_ll=min_ll,
                                        dist_factor=0.75, minp=MINP)
        return groups, sub

    step_s, (groups, sub) = _timed(step, args.steps, args.warmup)
    log(f"substructure: {step_s * 1e3:.1f} ms/step; groups in={groups.n_groups} "
        f"out={sub.groups.n_groups} found={sub.n_found}")
    extra = {}
    if not args.no_cpu:
        import oracle

        k = 40
        gl = groups.to_arrays()[:k]
        ns = int(sum(len(g) for g in gl))
        tc = time.perf_counter()
        _, cnt = oracle.split_substructure(gl, w["x"], w["y"], w["z"], w["mass"], ll, 25, 2.0,
                                           min_ll, 0.75, MINP)
        cs = time.perf_counter() - tc
        extra["cpu_baseline"] = {"value": ns / cs, "unit": "members/s", "cores": 1,
                                 "kind": "port", "seconds": cs,
                                 "sample": f"the first {k} groups ({ns} members, {cnt} "
                                           "substructures): the oracle's restatement, C "
                                           "cell-list FoF per group and level"}
    return _line("particles/s FoF + substructure splitting, 8M-particle bridged pairs",
                 n / step_s, "particles/s", step_s, args,
                 {"workload": "4000 Plummer pairs at 1-3 a joined by 5% bridges, 8M FP32, "
                              "FoF at ll = 0.35 a, sub-FoF at ll/2 and ll/4",
                  "particles": n, "groups_in": groups.n_groups,
                  "groups_out": sub.groups.n_groups}, substructures=sub.n_found,
                 members_in=groups.n_members, **extra)


def bench_cleaning(args):
    """Phase-space cleaning (pyHALMA.py:789-843) of the headline clump catalogue: FoF
    labels + CSR groups at ll, then calc_cell / clean_cell of every group at once with
    the reference's parameter file values (cells of 2 ll, sig_fil 3, q_fil 1, groups
    of at least 50 members cleaned, minp 50).  CPU baseline: the oracle's restatement
    of the same block on the first groups of the same catalogue."""
    import numpy as np
    import torch

    from pyhalma_amd import engine, synth

    t0 = time.time()
    w = synth.clumps(seed=args.seed)
    n = len(w["x"])
    ll = float(w["ll"])
    log(f"cleaning workload: {n} particles, ll={ll:.5f} ({time.time() - t0:.1f}s)")
    dev = torch.device("cuda", 0)
    keys = ("x", "y", "z", "vx", "vy", "vz", "mass")
    D = {k: torch.from_numpy(w[k]).to(dev) for k in keys}

    def step():
        lab = engine.fof_labels(D["x"], D["y"], D["z"], ll)
        groups = engine.groups_from_labels(lab, MINP)
        res = engine.phase_space_clean(groups, *(D[k] for k in keys), ll, 1.0, 3.0,
                                       psc=True, min_part=50, minp=MINP)
        return groups, res

    step_s, (groups, res) = _timed(step, args.steps, args.warmup)
    removed = int(groups.n_members - int(res.control_clean.sum().item()))
    log(f"cleaning: {step_s * 1e3:.1f} ms/step; groups {groups.n_groups} kept "
        f"{res.kept.n_groups}; members removed {removed}; rounds max {res.rounds.max()}")
    extra = {}
    if not args.no_cpu:
        import oracle

        k = 40
        gl = groups.to_arrays()[:k]
        ns = int(sum(len(g) for g in gl))
        tc = time.perf_counter()
        oracle.phase_space_clean(gl, *(w[q] for q in keys), ll, 1.0, 3.0)
        cs = time.perf_counter() - tc
        extra["cpu_baseline"] = {"value": ns / cs, "unit": "members/s", "cores": 1,
                                 "kind": "port", "seconds": cs,
                                 "sample": f"the first {k} groups ({ns} members): the "
                                           "oracle's restatement (sparse cells, one group "
                                           "at a time)"}
    return _line("particles/s FoF + phase-space cleaning, 2.1M-particle clumps",
                 n / step_s, "particles/s", step_s, args,
                 {"workload": "1000 Plummer clumps in a 40 Mpc box, 2.1M FP32, FoF at ll, "
                              "cells of 2 ll, sig_fil 3, q_fil 1",
                  "particles": n, "groups": groups.n_groups, "kept": res.kept.n_groups},
                 members_in=groups.n_members, members_removed=removed,
                 max_rounds=int(res.rounds.max()) if len(res.rounds) else 0, **extra)


def bench_progenitors(args):
    """Progenitor matching (pyHALMA.py:1211-1266) on the headline clump catalogue: the
    new haloes are the FoF groups at ll (1001 haloes, 1.91M members).  The old catalogue
    splits each new halo's members into 1-4 progenitors (Dirichlet fractions), leaves
    10 % of them out of every old halo, and joins the first progenitor of 5 % of the
    haloes to their neighbour's (old haloes that break apart); particle IDs are a random
    permutation offset by 10^9, masses the summed particle masses.  One step =
    halma_progenitors over all new haloes (inputs resident).  CPU baseline: the oracle's
    restatement (np.isin per (new, old) halo pair, as the reference) on the first new
    haloes."""
    import numpy as np
    import torch

    from pyhalma_amd import engine, synth

    t0 = time.time()
    w = synth.clumps(seed=args.seed)
    n = len(w["x"])
    ll = float(w["ll"])
    dev = torch.device("cuda", 0)
    D = {k: torch.from_numpy(w[k]).to(dev) for k in ("x", "y", "z", "mass")}
    rng = np.random.default_rng(args.seed)
    oripa = rng.permutation(n).astype(np.int64) + 10 ** 9
    _, new = engine.fof_groups(D["x"], D["y"], D["z"], ll, MINP)
    ids_d = torch.from_numpy(oripa).to(dev)
    new_mass = engine.group_reduce(new, D["x"], D["y"], D["z"], D["mass"])[:, 0].cpu().numpy()
    H = new.n_groups
    mem = new.members[: new.n_members].cpu().numpy().astype(np.int64)
    halo_of = np.repeat(np.arange(H), new.sizes())
    k = rng.integers(1, 5, H)
    frac = rng.dirichlet(np.full(4, 0.7), H) * (np.arange(4)[None, :] < k[:, None])
    cum = np.cumsum(frac / frac.sum(1, keepdims=True), 1)
    part = (rng.random(mem.size)[:, None] > cum[halo_of]).sum(1).clip(0, 3)
    label = halo_of * 4 + part
    joined = np.nonzero(rng.random(H - 1) < 0.05)[0] + 1  # first progenitor -> neighbour's
    label[np.isin(halo_of, joined) & (part == 0)] -= 4
    label[rng.random(mem.size) < 0.1] = -1                 # not in any old halo
    keep = label >= 0
    order = np.argsort(label[keep], kind="stable")
    old_members = mem[keep][order]
    _, old_npart = np.unique(label[keep], return_counts=True)
    omm = np.bincount(np.unique(label[keep], return_inverse=True)[1],
                      weights=w["mass"][mem[keep]].astype(np.float64))
    old_ids = ids_d[torch.from_numpy(old_members).to(dev)]
    old_lists = np.split(oripa[old_members], np.cumsum(old_npart)[:-1])
    log(f"progenitor workload: {H} new haloes ({new.n_members} members), "
        f"{old_npart.size} old haloes ({old_members.size} members) ({time.time() - t0:.1f}s)")

    def step():
        return engine.progenitors(new, ids_d, new_mass, old_npart, old_ids, omm)

    step_s, (pro1, pro2, nmerg, mtype) = _timed(step, args.steps, args.warmup)
    log(f"progenitors: {step_s * 1e3:.2f} ms/step; with a progenitor {int((pro1 > 0).sum())}; "
        f"n_mergers max {int(nmerg.max())}; types {np.unique(mtype).tolist()}")
    extra = {}
    if not args.no_cpu:
        import oracle

        k = 8
        groups = new.to_arrays()[:k]
        tc = time.perf_counter()
        want = oracle.progenitors(groups, oripa, new_mass[:k], old_lists, omm)
        cs = time.perf_counter() - tc
        same = all((a[:k] == b).all() for a, b in zip((pro1, pro2, nmerg, mtype), want))
        ns = int(sum(len(g) for g in groups))
        extra["cpu_baseline"] = {"value": ns / cs, "unit": "new members/s", "cores": 1,
                                 "kind": "port", "seconds": cs,
                                 "sample": f"the first {k} new haloes ({ns} members) against "
                                           f"all {old_npart.size} old haloes: the oracle's "
                                           "restatement, np.isin per halo pair"}
        extra["sample_matches_oracle"] = bool(same)
    return _line("new members/s progenitor matching, 2.1M-particle clumps",
                 new.n_members / step_s, "new members/s", step_s, args,
                 {"workload": "FoF groups at ll of 1000 Plummer clumps (2.1M FP32) against "
                              "an old catalogue of 1-4 progenitors each, 10 % unassigned, 5 % "
                              "joined to a neighbour; random int64 IDs",
                  "new_haloes": H, "old_haloes": int(old_npart.size),
                  "new_members": new.n_members, "old_members": int(old_members.size)},
                 with_progenitor=int((pro1 > 0).sum()), max_mergers=int(nmerg.max()),
                 merger_types={str(t): int((mtype == t).sum()) for t in np.unique(mtype)},
                 **extra)


def bench_sources(args):
    """Source gathering of the gas path (st_gas_dm_particles_inside, halo_gas.py:223-277)
    for every halo of the headline clump catalogue at once: the 1001 FoF groups' centres,
    radius 3 half-mass radii; stars = the 2.1M clump particles, dark matter = 4000 Plummer
    particles around each group plus a 2M uniform field (6M), gas = a synthetic AMR
    snapshot of two 16^3 patches (levels 6 and 7) around each group (8.2M cells).  One step
    = gas particles from the AMR cells in every sphere + the dark matter and stars in
    every sphere (ascending indices, masses in Msun).  CPU baseline: the oracle's
    restatement on the first haloes (scipy KD-trees built once, as the reference does;
    the AMR loops in pure Python, which the reference compiles with numba)."""
    import numpy as np
    import torch

    from pyhalma_amd import engine, gathering, synth

    t0 = time.time()
    w = synth.clumps(seed=args.seed)
    ll = float(w["ll"])
    dev = torch.device("cuda", 0)
    D = {k: torch.from_numpy(w[k]).to(dev) for k in ("x", "y", "z", "mass")}
    _, groups = engine.fof_groups(D["x"], D["y"], D["z"], ll, MINP)
    red = engine.group_reduce(groups, D["x"], D["y"], D["z"], D["mass"])
    r12 = engine.half_mass_radius(groups, D["x"], D["y"], D["z"], D["mass"], red)
    centres = red[:, 1:4].cpu().numpy()
    radii = 3.0 * r12.cpu().numpy()
    H = len(radii)
    dm = synth.dark_matter_around(centres)
    amr = synth.amr_around(centres)
    grid = gathering.AmrGrid(amr["L"], amr["ncoarse"], amr["levels"], amr["nx"], amr["ny"],
                             amr["nz"], amr["rx"], amr["ry"], amr["rz"], amr["cells"])
    rng = np.random.default_rng(args.seed)
    oripa = rng.permutation(len(w["x"])).astype(np.int32)
    DM = tuple(torch.from_numpy(dm[k]).to(dev) for k in ("x", "y", "z", "mass"))
    ST = (D["x"], D["y"], D["z"], D["mass"] * 1e-17, torch.from_numpy(oripa).to(dev))
    rete, rho_B = 0.5, 2.7e10
    log(f"source workload: {H} haloes, {len(dm['x'])} DM, {len(w['x'])} stars, "
        f"{grid.n_cells} AMR cells in {grid.n_patches} patches ({time.time() - t0:.1f}s)")

    def step():
        return gathering.gather_sources(grid, rho_B, rete, DM, ST, centres, radii)

    step_s, out = _timed(step, args.steps, args.warmup)
    n_gas = int(out["gas_offsets"][-1])
    n_dm = int(out["dm_offsets"][-1])
    n_st = int(out["st_offsets"][-1])
    log(f"sources: {step_s * 1e3:.2f} ms/step; gas {n_gas}, dm {n_dm}, stars {n_st}")
    extra = {}
    if not args.no_cpu:
        from scipy.spatial import KDTree

        import oracle.gathering as og

        k = 10
        dm_tree = KDTree(np.array((dm["x"], dm["y"], dm["z"])).T.astype(np.float64))
        st_tree = KDTree(np.array((w["x"], w["y"], w["z"])).T.astype(np.float64))
        tc = time.perf_counter()
        same = True
        for h in range(k):
            c, R = centres[h], radii[h]
            g = og.amr_to_particles(amr["L"], amr["ncoarse"], amr["levels"], amr["rx"],
                                    amr["ry"], amr["rz"], amr["cells"], R, *c, rho_B)
            keep = og.inside(g[0], g[1], g[2], R, *c)
            di = sorted(dm_tree.query_ball_point(c, R))
            si = sorted(st_tree.query_ball_point(c, R))
            a, b = out["gas_offsets"][h], out["gas_offsets"][h + 1]
            same &= bool(np.array_equal(out["gas"]["x"][a:b].cpu().numpy(), g[0][keep]))
            a, b = out["dm_offsets"][h], out["dm_offsets"][h + 1]
            same &= out["dm_index"][a:b].cpu().tolist() == di
            a, b = out["st_offsets"][h], out["st_offsets"][h + 1]
            same &= out["st_index"][a:b].cpu().tolist() == si
        cs = time.perf_counter() - tc
        extra["cpu_baseline"] = {"value": k / cs, "unit": "haloes/s", "cores": 1,
                                 "kind": "port", "seconds": cs,
                                 "sample": f"the first {k} haloes: AMR cells by the oracle's "
                                           "pure-Python loops, dark matter and stars by "
                                           "scipy KD-trees built once (untimed)"}
        extra["sample_matches_oracle"] = same
    return _line("haloes/s source gathering (gas from AMR cells, dark matter, stars)",
                 H / step_s, "haloes/s", step_s, args,
                 {"workload": "1001 clump haloes, R = 3 half-mass radii; 2.1M stars, 6M dark "
                              "matter, 8.2M AMR cells (2 patches of 16^3 per halo)",
                  "haloes": H, "dm_particles": len(dm["x"]), "stars": len(w["x"]),
                  "amr_cells": grid.n_cells},
                 gas_found=n_gas, dm_found=n_dm, stars_found=n_st, **extra)


def bench_centring(args):
    """Differentiable centring: 200 Hernquist haloes of 5e4-5e5 particles, seeds offset
    by 0.1-0.5 a; autograd-driven descent on the softened potential (eps = 0.5 kpc) to
    the potential-minimum centre of every halo at once."""
    import numpy as np
    import torch

    from pyhalma_amd import centring, synth

    t0 = time.time()
    w = synth.hernquist_haloes()
    off = w["offsets"]
    n = int(off[-1])
    log(f"centring workload: {len(off) - 1} haloes, {n} particles ({time.time() - t0:.1f}s)")
    dev = torch.device("cuda", 0)
    D = [torch.from_numpy(w[k]).to(dev) for k in ("x", "y", "z", "mass")]
    eps = 5e-4
    step_s, res = _timed(lambda: centring.refine_centres(*D, off, w["seeds"], eps,
                                                         max_iter=100, tol=1e-5),
                         args.steps, args.warmup)
    c = res.centres.cpu().numpy()
    err = np.linalg.norm(c - w["centres"], axis=1) / w["a"]
    it = res.iters.cpu().numpy()
    log(f"centring: {step_s * 1e3:.1f} ms/step; iterations max {it.max()} "
        f"median {np.median(it)}; |c - c_true|/a median {np.median(err):.2e}")
    return _line("particles/s potential-minimum centring (200 Hernquist haloes)", n / step_s,
                 "particles/s", step_s, args,
                 {"workload": "200 Hernquist haloes of 5e4-5e5, seeds 0.1-0.5 a off, FP32, "
                              "eps 0.5 kpc", "particles": n, "haloes": len(off) - 1},
                 iterations_max=int(it.max()), iterations_median=float(np.median(it)),
                 offset_over_a_median=float(np.median(err)),
                 offset_over_a_max=float(err.max()))


def cpu_baseline(w, frac):
    """CPU reference path on a sub-box (x < -L/2 + frac*L) of the same workload: the
    oracle's cell-list FoF port, groups >= 50, numpy reductions, and the reference's own
    FP32 pair routine (particle_subroutines.f90:466-514 compiled from its Fortran source
    into oracle/_ref, OpenMP on every host core; the oracle's C port of the same loop
    when that build is absent) inside the same iterated energy test."""
    import numpy as np

    import oracle

    ncores = os.cpu_count() or 1
    L = w["box"]
    sel = w["x"] < (-L / 2 + frac * L)
    x, y, z = w["x"][sel], w["y"][sel], w["z"][sel]
    vx, vy, vz, m = w["vx"][sel], w["vy"][sel], w["vz"][sel], w["mass"][sel]
    nsub = int(sel.sum())
    use_ref = oracle.have_ref()
    pot = (lambda *a: oracle.ref_pot_sum_f32(*a, nthreads=ncores)) if use_ref else \
        (lambda *a: oracle.pot_sum_f32(*a, nthreads=ncores))
    t0 = time.perf_counter()
    lab = oracle.fof_labels_c(x, y, z, w["ll"])
    groups = [np.asarray(g) for g in oracle.labels_to_groups(lab, 50)]
    cm = [oracle.center_of_mass(g, x, y, z, m) for g in groups]
    G = oracle.G_CONST
    for g, (cx, cy, cz, M) in zip(groups, cm):
        b = np.ones(len(g), bool)
        for _ in range(100):
            s = g[b]
            if s.size == 0:
                break
            mm = m[s].astype(np.float64)
            vc = [float((mm * v[s]).sum() / mm.sum()) for v in (vx, vy, vz)]
            S = pot(m[s], x[s], y[s], z[s], x[s], y[s], z[s]).astype(np.float64)
            E = 0.5 * ((vx[s] - vc[0]) ** 2 + (vy[s] - vc[1]) ** 2 + (vz[s] - vc[2]) ** 2) - G * S
            if (E <= 0).all():
                break
            pos = np.nonzero(b)[0]
            b[pos[E > 0]] = False
    dt = time.perf_counter() - t0
    return {"value": nsub / dt, "unit": "particles/s", "cores": ncores,
            "kind": "reference" if use_ref else "port",
            "sample": f"sub-box x < -L/2 + {frac}L of the same workload: {nsub} particles, "
                      f"{len(groups)} groups >= 50; FoF = C cell-list port, pair sums = "
                      f"{'reference Fortran routine' if use_ref else 'C port of it'} "
                      f"(FP32, OpenMP {ncores} threads), iterated energy test; {dt:.1f} s",
            "seconds": dt}


if __name__ == "__main__":
    main()
