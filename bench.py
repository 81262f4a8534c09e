"""Benchmark: FoF + per-group reductions + batched iterative unbinding on one GPU.

Default workload (clumps): 1000 Plummer clumps in a 40 Mpc box, clump sizes
p(N) ~ N^-1.385 on [100, 20000] (~1.9M clump stars) plus 10% uniform field stars (~2.1M
particles in all), FP32 positions, velocities and log-uniform masses 1e4-1e6 Msun, linking
length 0.2x the mean interparticle spacing, groups of >= 50 members, unbinding in physical
mode (E = K + U, factor 1) iterated to convergence.  Synthetic data generated from a seed;
inputs are resident in HBM before the timed region.

One step = friends-of-friends labels -> groups -> one gather of the members' data into
group order -> group reductions (mass, centre of mass, CM velocity, furthest particle)
-> density peaks -> batched unbinding.  After the timed steps: the plain FP64 comparator
(same method: one launch per halo, FP64 sums, no batching, no graphs), the per-halo
potential error of every halo against the library's FP64 fixed-order sum (first pass and
last pass), and the CPU baseline leg: the reference's own FP32 pair routine (compiled from
its Fortran source, OpenMP on the host cores) with the oracle's FoF on a bounded
sub-box, median of several runs, plus the oracle's labels of the whole catalogue as the
checker of the GPU's.

Other workloads (--workload): clumps_alpha2 (the same with the literal slope 2),
plummer (20k Plummer + 4k background, FP64), gas (500 NFW gas haloes), bridged (8M
stars in 4000 bridged pairs), centring (200 Hernquist haloes: potential energy forward
and backward, then the potential-minimum centres), substructure, cleaning, progenitors,
sources.

    python bench.py [--gpus 1] [--steps 5] [--warmup 2] [--workload clumps]
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
PEAK_HBM_GBS = 8000.0      # MI355X HBM3E (AMD published)
FLOP_PER_PAIR = 11         # particle_subroutines.f90:502-505 as written
FLOP_PER_SOFT_PAIR = 12    # the same term with the softening length added to r^2
FLOP_PER_ACC_PAIR = 20     # acceleration: 3 sub, r^2 + eps^2 (6), 1/r (2), r^-3 (2),
#                            m r^-3 (1), three products and sums (6)
FOF_BYTES_PER_PARTICLE = 16  # 12 B FP32 xyz read + 4 B label written
# HBM bytes per k_pot_sym launch on the default workload, averaged over the launches of a
# step (the first pass carries almost all of it): FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0 from
# two rocprofv3 --pmc passes, each scaled by what the counter reads for a 256 MiB device
# copy (profiles/clumps_2M_hbm_traffic.md)
SYM_TRAFFIC_BYTES = 80.0e6
MINP = 50
VALU_ROOF = ("FP32 vector ALU (157.3 TFLOP/s, MI355X's dense FP32 figure, which counts "
             "packed FMA); the kernel issues FP32 vector instructions, no matrix "
             "instructions and no packed ones")


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--seed", type=int, default=20261204)
    ap.add_argument("--no-plain", action="store_true", help="skip the plain FP64 comparator")
    ap.add_argument("--no-cpu", action="store_true",
                    help="skip the CPU baseline leg (and the oracle checks it holds)")
    ap.add_argument("--no-check", action="store_true",
                    help="skip the GPU FP64 potential-error pass after the timed steps")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the built-in pair-kernel tiling instead of autotuning it on "
                         "the workload's haloes before the timed steps")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="run one 256 MiB device copy first (scales profiler byte counters)")
    ap.add_argument("--cpu-frac", type=float, default=0.15,
                    help="fraction of the box (along x) used for the CPU baseline sample")
    ap.add_argument("--cpu-runs", type=int, default=5,
                    help="timed CPU baseline runs (the median is reported) after one warm-up")
    ap.add_argument("--workload", default="clumps",
                    choices=["clumps", "clumps_alpha2", "plummer", "gas", "bridged",
                             "centring", "substructure", "cleaning", "progenitors",
                             "sources"],
                    help="clumps (default): the headline 2.1M-particle step; clumps_alpha2 "
                         "the same with the literal size slope 2; plummer the 20k + 4k "
                         "FP64 set; gas, bridged, centring the other measured sets; "
                         "substructure, cleaning, progenitors, sources the rows around "
                         "the path at full size")
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
    import torch

    torch.cuda.init()
    fn = {"clumps": lambda a: bench_clumps(a, 1.385),
          "clumps_alpha2": lambda a: bench_clumps(a, 2.0),
          "plummer": bench_plummer, "gas": bench_gas, "bridged": bench_bridged,
          "centring": bench_centring, "substructure": bench_substructure,
          "cleaning": bench_cleaning, "progenitors": bench_progenitors,
          "sources": bench_sources}[args.workload]
    print(json.dumps(fn(args)))


# --------------------------------------------------------------------- helpers ----
def _threads():
    """Host threads the CPU baselines use: OMP_NUM_THREADS when set (the GPU box sets
    it to its CPU share), else every CPU."""
    try:
        return max(1, int(os.environ.get("OMP_NUM_THREADS", "0")) or (os.cpu_count() or 1))
    except ValueError:
        return os.cpu_count() or 1


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _median_runs(fn, runs, warmup=1):
    """Median wall time of ``runs`` calls after ``warmup`` untimed ones; the times and
    the last result."""
    import numpy as np

    r = None
    for _ in range(warmup):
        r = fn()
    ts = []
    for _ in range(max(1, runs)):
        t0 = time.perf_counter()
        r = fn()
        ts.append(time.perf_counter() - t0)
    return float(np.median(ts)), [round(t, 4) for t in ts], r


def _cpu_entry(value, unit, cores, kind, med, ts, sample):
    return {"value": value, "unit": unit, "cores": cores, "kind": kind,
            "cpu_model": _cpu_model(), "seconds_median": med, "runs_seconds": ts,
            "sample": sample}


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


def _sym_roofline(acc, steps, traffic=None):
    """Roofline of the symmetric pair kernel of the unbinding's full passes from the
    accumulated unbinding statistics (its launches bracketed by HIP events on the
    library's stream).  Algorithmic work: the reference's loop over the same pairs, 11 FLOP
    per ordered (test, source) pair; the kernel evaluates each unordered pair once for
    both terms, so the instructions it executes are about half of that count."""
    sym_s = acc["sym_ns"] * 1e-9 / max(acc["sym_launches"], 1)
    pairs_per_launch = acc["full_pairs"] / max(acc["sym_launches"], 1)
    achieved = FLOP_PER_PAIR * pairs_per_launch / sym_s / 1e12 if sym_s > 0 else 0.0
    f32_pairs = acc["pairs"] - acc["full_pairs"]
    f32_s = (acc["pot_ns"] - acc["sym_ns"]) * 1e-9
    return {"bound": "valu", "achieved": round(achieved, 3), "peak": PEAK_FP32_TFLOPS,
            "unit": "TFLOP/s", "frac": round(achieved / PEAK_FP32_TFLOPS, 4),
            "traffic": traffic, "roof": VALU_ROOF,
            "kernel": "k_pot_sym (symmetric FP32 pair kernel, full passes)",
            "flop_per_pair": FLOP_PER_PAIR, "pairs_per_launch": int(pairs_per_launch),
            "launch_ms": round(sym_s * 1e3, 4),
            "launches_per_step": acc["sym_launches"] / steps,
            "executed_pair_evals_per_launch": int(pairs_per_launch // 2),
            "chunked_kernel": {"kernel": "k_pot_f32 (incremental passes)",
                               "launches_per_step": acc["f32_launches"] / steps,
                               "ms_per_step": f32_s * 1e3 / steps,
                               "tflops": (FLOP_PER_PAIR * f32_pairs / f32_s / 1e12
                                          if f32_s > 0 else None)}}


_UB_KEYS = ("pot_ns", "pot_launches", "pairs", "sym_ns", "sym_launches", "full_pairs",
            "f32_launches", "resummed", "retested")


def _unbind_step(D, ll, factor=1.0, fof_events=None):
    """The step of the clump-like workloads on device arrays D: FoF labels -> groups ->
    member gather -> reductions -> density peaks -> iterated unbinding.  With
    ``fof_events`` (a list), a pair of HIP events recorded on the current stream around
    the FoF call is appended per step."""
    import torch

    from pyhalma_amd import engine

    def step(time_pot=False):
        if fof_events is not None:
            s = torch.cuda.current_stream()
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record(s)
        lab = engine.fof_labels(D["x"], D["y"], D["z"], ll)
        if fof_events is not None:
            e1.record(s)
            fof_events.append((e0, e1))
        groups = engine.groups_from_labels(lab, MINP)
        md = engine.gather_members(groups, D["x"], D["y"], D["z"], D["mass"], D["vx"], D["vy"],
                                   D["vz"])
        g = md.groups
        red = engine.group_reduce(g, md.x, md.y, md.z, md.m, md.vx, md.vy, md.vz)
        peak = engine.density_peak(g, md.x, md.y, md.z, md.m, ll, reduced=red)
        ub = engine.unbind(g, md.x, md.y, md.z, md.vx, md.vy, md.vz, md.m, factor=factor,
                           time_pot=time_pot)
        return lab, groups, md, red, peak, ub

    return step


def _timed_unbind_steps(step, steps, warmup):
    import torch

    for i in range(warmup):
        step()
        log(f"warmup {i + 1}/{warmup} done")
    acc = dict.fromkeys(_UB_KEYS, 0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        r = step(time_pot=True)
        for k in acc:
            acc[k] += r[-1].stats[k]
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / steps, acc, r


def _pot_errors(groups, md, factor, bound=None, pot=None):
    """Relative error of the unbinding's own S against the library's FP64 fixed-order sum
    (halma_pot_sum_f64_ordered) for every halo: with ``pot`` None, the first pass (one
    pass over all members, the symmetric kernel's S); else ``pot`` of the members still
    bound at the end against the FP64 sum over the final bound set.  Returns the per-halo
    maxima and a summary."""
    import numpy as np
    import torch

    from pyhalma_amd import engine

    if pot is None:
        r1 = engine.unbind(groups, md.x, md.y, md.z, md.vx, md.vy, md.vz, md.m, factor=factor,
                           max_iter=1, with_pot=True)
        pot = r1.pot
        bound = torch.ones_like(r1.bound)
    off = groups.offsets_host
    per_halo = np.zeros(groups.n_groups)
    allrel = []
    for h in range(groups.n_groups):
        a, b = int(off[h]), int(off[h + 1])
        sel = bound[a:b].bool()
        if int(sel.sum()) == 0:
            continue
        mm, xx, yy, zz = (t[a:b][sel] for t in (md.m, md.x, md.y, md.z))
        s64 = engine.pot_sum_f64_ordered(mm, xx, yy, zz, xx, yy, zz)
        rel = ((pot[a:b][sel] - s64).abs() / s64).cpu().numpy()
        rel = rel[np.isfinite(rel)]
        if rel.size:
            per_halo[h] = rel.max()
            allrel.append(rel)
    allrel = np.concatenate(allrel) if allrel else np.zeros(1)
    return {"haloes": int(groups.n_groups), "members": int(allrel.size),
            "max": float(allrel.max()), "p99": float(np.quantile(allrel, 0.99)),
            "median": float(np.median(allrel)),
            "members_above_1e-6": int((allrel > 1e-6).sum()),
            "haloes_above_1e-6": int((per_halo > 1e-6).sum()),
            "per_halo_max_p50": float(np.median(per_halo))}


def _cpu_unbind_path(x, y, z, vx, vy, vz, m, ll, factor, runs, fof="cells", label=""):
    """CPU reference path on the given particles: FoF (the oracle's C sub-cell FoF, or
    its restatement of fof.py:18-79 over scipy's KD-tree), groups >= 50, and the iterated
    energy test with the reference's own FP32 pair routine
    (particle_subroutines.f90:466-514 compiled from its Fortran source into oracle/_ref,
    OpenMP on the host threads; the oracle's C port of the same loop when that build is
    absent).  Median of ``runs`` runs after one warm-up."""
    import numpy as np

    import oracle

    nthreads = _threads()
    use_ref = oracle.have_ref()
    pot = (lambda *a: oracle.ref_pot_sum_f32(*a, nthreads=nthreads)) if use_ref else \
        (lambda *a: oracle.pot_sum_f32(*a, nthreads=nthreads))
    G = oracle.G_CONST

    def run():
        if fof == "bfs":
            groups = [np.asarray(g) for g in
                      oracle.fof_bfs(np.stack([x, y, z], 1).astype(np.float64), ll)
                      if len(g) >= MINP]
        else:
            lab = oracle.fof_labels_cells(x, y, z, ll)
            groups = [np.asarray(g) for g in oracle.labels_to_groups(lab, MINP)]
        kept = 0
        for g in groups:
            oracle.center_of_mass(g, x, y, z, m)
            b = np.ones(len(g), bool)
            for _ in range(100):
                s = g[b]
                if s.size == 0:
                    break
                mm = m[s].astype(np.float64)
                vc = [float((mm * v[s]).sum() / mm.sum()) for v in (vx, vy, vz)]
                S = pot(m[s], x[s], y[s], z[s], x[s], y[s], z[s]).astype(np.float64)
                E = 0.5 * ((vx[s] - vc[0]) ** 2 + (vy[s] - vc[1]) ** 2
                           + (vz[s] - vc[2]) ** 2) - factor * G * S
                if (E <= 0).all():
                    break
                pos = np.nonzero(b)[0]
                b[pos[E > 0]] = False
            kept += int(b.sum())
        return len(groups), kept

    med, ts, (ng, kept) = _median_runs(run, runs)
    n = len(x)
    fof_txt = ("FoF = the oracle's restatement of fof.py:18-79 over scipy's KD-tree"
               if fof == "bfs" else "FoF = the oracle's C sub-cell FoF")
    return _cpu_entry(n / med, "particles/s", nthreads, "reference" if use_ref else "port",
                      med, ts,
                      f"{label}{n} particles, {ng} groups >= {MINP}; {fof_txt}; pair sums = "
                      f"{'the reference Fortran routine' if use_ref else 'the C port of it'} "
                      f"(FP32, OpenMP {nthreads} threads) in the iterated energy test; "
                      f"median of {len(ts)} runs after a warm-up")


# ---------------------------------------------------------------- workloads ----
def bench_clumps(args, alpha):
    """The headline step (alpha 1.385, the survey's resolution of the 2M / 1000-halo
    target) or the literal slope 2 (fewer particles: the same 1000 clumps)."""
    import numpy as np
    import torch

    from pyhalma_amd import engine, synth

    dev = torch.device("cuda", 0)
    t0 = time.time()
    w = synth.clumps(seed=args.seed, alpha=alpha)
    n = len(w["x"])
    ll = w["ll"]
    log(f"workload: {n} particles, {len(w['sizes'])} clumps, alpha={alpha}, ll={ll:.5f} Mpc "
        f"(generated in {time.time() - t0:.1f}s)")
    keys = ("x", "y", "z", "vx", "vy", "vz", "mass")
    torch.cuda.synchronize()
    th0 = time.perf_counter()
    D = {k: torch.from_numpy(w[k]).to(dev) for k in keys}
    torch.cuda.synchronize()
    h2d_s = time.perf_counter() - th0  # first upload: also pays allocator set-up
    for _ in range(2):  # steady-state upload time
        torch.cuda.synchronize()
        th0 = time.perf_counter()
        for k in keys:
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

    step = _unbind_step(D, ll)
    tuning = None
    if not args.no_autotune:
        from pyhalma_amd import autotune
        _, groups, md, _, _, _ = step()
        t_at = time.time()
        # the chunked kernel's knobs, timed on the step's own unbinding
        setting, ms, trials = autotune.autotune_unbind(md.groups, md.x, md.y, md.z, md.vx,
                                                       md.vy, md.vz, md.m, factor=1.0, reps=5)
        tuning = {"setting": setting, "unbind_ms": ms, "trials": len(trials),
                  "default_ms": trials[0][1] if trials else ms}
        log(f"autotune: {setting} {ms:.3f} ms (default {tuning['default_ms']:.3f} ms, "
            f"{len(trials)} trials, {time.time() - t_at:.1f}s)")
        del md
    step_s, acc, (lab, groups, md, red, peak, ub) = _timed_unbind_steps(step, args.steps,
                                                                       args.warmup)
    log(f"timed {args.steps} steps: {step_s * 1e3:.2f} ms/step; groups={groups.n_groups} "
        f"members={groups.n_members} passes={ub.stats['passes']} "
        f"retested={ub.stats['retested']} resummed={ub.stats['resummed']}")
    slope = "1.385" if alpha != 2.0 else "2 (the literal slope)"
    out = {
        "metric": METRIC, "value": n / step_s, "unit": "particles/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": step_s * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded Plummer clumps + uniform field)",
        "config": {"workload": f"1000 Plummer clumps in a 40 Mpc box, p(N)~N^-{slope} on "
                               "[100, 20000], +10% field, FP32, ll=0.2 mean spacing, "
                               "groups>=50, physical-mode iterative unbinding",
                   "alpha": alpha, "particles": n, "groups": groups.n_groups,
                   "members": groups.n_members, "passes": ub.stats["passes"],
                   "retested_fp64": ub.stats["retested"],
                   "resummed_for_1e-6": ub.stats["resummed"]},
        "roofline": _sym_roofline(acc, args.steps,
                                  SYM_TRAFFIC_BYTES if alpha != 2.0 else None),
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
        same_masks = torch.equal(pb, ub.bound)
        same_iters = bool(np.array_equal(piters, ub.iters.cpu().numpy()))
        del red2
        out["plain_f64"] = {"ms": plain_s * 1e3, "value": n / plain_s,
                            "speedup": plain_s / step_s, "masks_identical": bool(same_masks),
                            "passes_identical": same_iters}
        log(f"plain FP64: {plain_s * 1e3:.1f} ms -> speedup {plain_s / step_s:.1f}x; "
            f"masks identical: {same_masks}")

    if not args.no_check:
        # the unbinding's own S against the FP64 fixed-order sum, every halo: the first
        # pass (symmetric kernel) and the last pass (after the incremental updates)
        tc = time.perf_counter()
        first = _pot_errors(md.groups, md, 1.0)
        r = engine.unbind(md.groups, md.x, md.y, md.z, md.vx, md.vy, md.vz, md.m, factor=1.0,
                          with_pot=True)
        last = _pot_errors(md.groups, md, 1.0, bound=r.bound, pot=r.pot)
        out["pot_rel_err"] = {"first_pass": first, "last_pass_bound_members": last,
                              "fp64_check": "halma_pot_sum_f64_ordered on the GPU",
                              "seconds": round(time.perf_counter() - tc, 2)}
        out["pot_error_bound"] = engine.pot_error_bound()
        log(f"potential error: first pass max {first['max']:.2e}, last pass max "
            f"{last['max']:.2e} (haloes above 1e-6: {first['haloes_above_1e-6']}, "
            f"{last['haloes_above_1e-6']})")

    if not args.no_cpu:
        import oracle

        # the oracle's labels of the whole catalogue, as the checker of the GPU's
        tc = time.perf_counter()
        want = oracle.fof_labels_cells(w["x"], w["y"], w["z"], ll)
        same = bool(np.array_equal(lab.cpu().numpy(), want))
        out["fof_labels_identical_full_size"] = same
        log(f"FoF labels vs the oracle (all {n} particles): identical={same} "
            f"({time.perf_counter() - tc:.1f}s)")
        L = w["box"]
        sel = w["x"] < (-L / 2 + args.cpu_frac * L)
        sub = [w[k][sel] for k in keys]
        out["cpu_baseline"] = _cpu_unbind_path(*sub[:6], sub[6], ll, 1.0, args.cpu_runs,
                                               label=f"sub-box x < -L/2 + {args.cpu_frac} L "
                                                     "of the same workload: ")
        log(f"cpu baseline: {out['cpu_baseline']['value']:.3e} particles/s")
    return out


def bench_plummer(args):
    """The CPU reference path's set at full size: one Plummer sphere of 20,000 stars
    (a = 5 kpc, 1e5 Msun each, FP64, virial Gaussian velocities) plus 4,000 uniform
    background stars in a 200 kpc box at 3 v_esc; FoF at 0.2 x the mean spacing, groups
    >= 50, physical-mode iterated unbinding.  CPU baseline: the same path on the host
    (FoF = the oracle's restatement of fof.py:18-79 over scipy's KD-tree, pair sums = the
    reference's Fortran routine)."""
    import numpy as np
    import torch

    from pyhalma_amd import engine, synth

    w = synth.plummer_with_background()
    n = len(w["x"])
    ll = float(w["ll"])
    dev = torch.device("cuda", 0)
    keys = ("x", "y", "z", "vx", "vy", "vz", "mass")
    D = {k: torch.from_numpy(w[k]).to(dev) for k in keys}
    step = _unbind_step(D, ll)
    step_s, acc, (lab, groups, md, red, peak, ub) = _timed_unbind_steps(step, args.steps,
                                                                       args.warmup)
    log(f"plummer: {step_s * 1e3:.3f} ms/step; groups={groups.n_groups} "
        f"bound={int(ub.bound.sum().item())} passes={ub.stats['passes']}")
    out = _line("particles/s FoF+unbinding, Plummer 20k + 4k background (FP64 inputs)",
                n / step_s, "particles/s", step_s, args,
                {"workload": "Plummer sphere of 20,000 (a = 5 kpc, 1e5 Msun, FP64) + 4,000 "
                             "uniform background at 3 v_esc in a 200 kpc box, ll = 0.2 mean "
                             "spacing, groups>=50, physical-mode iterative unbinding",
                 "particles": n, "groups": groups.n_groups, "members": groups.n_members,
                 "passes": ub.stats["passes"]},
                roofline=_sym_roofline(acc, args.steps),
                bound_members=int(ub.bound.sum().item()))
    out["data"] = "synthetic (seeded Plummer sphere + uniform background)"
    out["dtype"] = "f32 pair sums on FP64 inputs"
    if not args.no_check:
        first = _pot_errors(md.groups, md, 1.0)
        r = engine.unbind(md.groups, md.x, md.y, md.z, md.vx, md.vy, md.vz, md.m, factor=1.0,
                          with_pot=True)
        out["pot_rel_err"] = {"first_pass": first,
                              "last_pass_bound_members": _pot_errors(md.groups, md, 1.0,
                                                                     bound=r.bound,
                                                                     pot=r.pot)}
    if not args.no_cpu:
        import oracle

        want = oracle.fof_labels_cells(w["x"], w["y"], w["z"], ll)
        out["fof_labels_identical"] = bool(np.array_equal(lab.cpu().numpy(), want))
        mem = groups.members[: groups.n_members].cpu().numpy()
        bound, iters, _ = oracle.unbind_iterative(groups.offsets_host, mem, w["x"], w["y"],
                                                  w["z"], w["vx"], w["vy"], w["vz"],
                                                  w["mass"], 1.0)
        out["masks_identical"] = bool(np.array_equal(ub.bound.cpu().numpy().astype(bool),
                                                     bound))
        out["passes_identical"] = bool(np.array_equal(ub.iters.cpu().numpy(), iters))
        out["cpu_baseline"] = _cpu_unbind_path(*(w[k] for k in keys), ll, 1.0, args.cpu_runs,
                                               fof="bfs", label="the whole set: ")
        log(f"plummer cpu baseline: {out['cpu_baseline']['value']:.3e} particles/s; "
            f"labels {out['fof_labels_identical']} masks {out['masks_identical']}")
    return out


def bench_gas(args):
    """Gas batch: 500 NFW gas haloes of 1e3-2e5 cells, one exact energy test per cell
    (factor 2, sources = the halo's cells, halo_gas.py:285-492 without subsampling),
    every halo in one call.  Full-size check: a random sample of cells re-decided with the
    library's FP64 fixed-order sum.  CPU baseline: the reference's Fortran pair routine
    (OpenMP) in the same energy test on a sample of the haloes."""
    import numpy as np
    import torch

    from pyhalma_amd import engine, halo_gas, synth

    t0 = time.time()
    w = synth.gas_haloes()
    off = w["offsets"]
    H = len(off) - 1
    n = int(off[-1])
    sizes = np.diff(off)
    pairs = float((sizes.astype(np.float64) ** 2).sum())
    log(f"gas workload: {H} haloes, {n} cells, {pairs:.3e} pairs ({time.time() - t0:.1f}s)")
    dev = torch.device("cuda", 0)
    g = [torch.from_numpy(w[k]).to(dev) for k in ("x", "y", "z", "vx", "vy", "vz", "mass",
                                                   "temp")]
    bulk = torch.from_numpy(w["bulk"]).to(dev)
    kstats = []

    def step():
        st = {}
        r = halo_gas.RPS_batched(g, off, bulk, stats=st)
        kstats.append(st)
        return r

    step_s, (out, bound) = _timed(step, args.steps, args.warmup)
    timed = kstats[args.warmup:]
    pair_s = float(np.mean([s["pair_ns"] for s in timed])) * 1e-9
    achieved = FLOP_PER_PAIR * timed[-1]["pairs"] / pair_s / 1e12
    log(f"gas: {step_s * 1e3:.1f} ms/step; pair kernel {pair_s * 1e3:.1f} ms "
        f"({achieved:.1f} TFLOP/s)")
    # the self-term runs on the symmetric kernel in bands of 64 blocks of 512 members: one
    # launch per tile (P <= Q) of the halo with the most bands
    blocks = -(-(-(-sizes // 128)) // 4)
    bands = -(-blocks // 64)
    launches = int((bands * (bands + 1) // 2).max())
    roofline = {"bound": "valu", "achieved": round(achieved, 3), "peak": PEAK_FP32_TFLOPS,
                "unit": "TFLOP/s", "frac": round(achieved / PEAK_FP32_TFLOPS, 4),
                "traffic": None, "roof": VALU_ROOF,
                "kernel": "k_pot_sym (symmetric FP32 pair kernel, one launch per tile of 64 x 64 "
                          "blocks of 512 cells; the HIP-event window also holds the tile merges)",
                "flop_per_pair": FLOP_PER_PAIR,
                "pairs_per_launch": int(timed[-1]["pairs"]) // launches,
                "launch_ms": round(pair_s * 1e3 / launches, 3), "launches_per_step": launches,
                "window_ms": round(pair_s * 1e3, 3)}
    # full-size check on a sample: FP64 fixed-order sums of sampled cells
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
    extra = {}
    if not args.no_cpu:
        import oracle

        nthreads = _threads()
        use_ref = oracle.have_ref()
        pot = (lambda *a: oracle.ref_pot_sum_f32(*a, nthreads=nthreads)) if use_ref else \
            (lambda *a: oracle.pot_sum_f32(*a, nthreads=nthreads))
        # every 25th halo by size (20 haloes spanning the distribution)
        pick = np.argsort(sizes)[::25]
        cells = int(sizes[pick].sum())
        G = engine_g()

        def run():
            nb = 0
            for h in pick:
                a, b = off[h], off[h + 1]
                S = pot(w["mass"][a:b], w["x"][a:b], w["y"][a:b], w["z"][a:b], w["x"][a:b],
                        w["y"][a:b], w["z"][a:b]).astype(np.float64)
                dv = [w[k][a:b].astype(np.float64) - w["bulk"][h][q]
                      for q, k in enumerate(("vx", "vy", "vz"))]
                E = 0.5 * ((dv[0] * dv[0] + dv[1] * dv[1]) + dv[2] * dv[2]) - 2.0 * G * S
                nb += int((E <= 0).sum())
            return nb

        med, ts, _ = _median_runs(run, args.cpu_runs)
        spairs = float((sizes[pick].astype(np.float64) ** 2).sum())
        extra["cpu_baseline"] = _cpu_entry(
            cells / med, "cells/s", nthreads, "reference" if use_ref else "port", med, ts,
            f"every 25th halo by size ({len(pick)} haloes, {cells} cells, {spairs:.3e} "
            f"pairs): the energy test with the "
            f"{'reference Fortran pair routine' if use_ref else 'C port of it'} (FP32, "
            f"OpenMP {nthreads} threads); median of {len(ts)} runs after a warm-up")
        log(f"gas cpu baseline: {cells / med:.3e} cells/s")
    return _line("gas cells/s, exact batched energy test (500 NFW haloes, 1e3-2e5 cells)",
                 n / step_s, "cells/s", step_s, args,
                 {"workload": "gas batch: 500 NFW haloes, cells log-uniform 1e3-2e5, FP32, "
                              "factor 2, sources = each halo's cells", "cells": n,
                  "haloes": H, "pairs": pairs},
                 roofline=roofline, pairs_per_s=pairs / step_s,
                 sample_mismatches=mism, sample_size=int(len(sample)),
                 bound_fraction=float(bound.float().mean().item()), **extra)


def engine_g():
    return 4.3 * 1e-3 * 1e-6


def bench_bridged(args):
    """Bridged stress case: 8M stars in 4000 Plummer pairs joined by a 5% bridge; FoF
    labels + CSR groups + gather + reductions + density peaks + iterated unbinding.  FoF
    must merge each pair: every group of >= 50 members holds exactly one pair's
    particles (checked against the generator's pair ids), and the oracle's labels of all
    8M particles equal the GPU's.  Roofline: the FoF labels (HBM, 16 B per particle),
    timed with HIP events on the stream of its kernels."""
    import numpy as np
    import torch

    from pyhalma_amd import engine, synth

    t0 = time.time()
    w = synth.bridged_pairs()
    n = len(w["x"])
    ll = w["ll"]
    log(f"bridged workload: {n} particles, ll={ll:.5f} ({time.time() - t0:.1f}s)")
    dev = torch.device("cuda", 0)
    keys = ("x", "y", "z", "vx", "vy", "vz", "mass")
    D = {k: torch.from_numpy(w[k]).to(dev) for k in keys}
    fof_ms = []
    step = _unbind_step(D, ll, fof_events=fof_ms)
    step_s, acc, (lab, groups, md, red, peak, ub) = _timed_unbind_steps(step, args.steps,
                                                                       args.warmup)
    torch.cuda.synchronize()
    fof_s = float(np.mean([a.elapsed_time(b) for a, b in fof_ms[args.warmup:]])) * 1e-3
    mem = groups.members[: groups.n_members].cpu().numpy()
    offs = groups.offsets_host
    pid = w["pair_id"][mem]
    seg_min = np.minimum.reduceat(pid, offs[:-1]) if groups.n_groups else np.array([])
    seg_max = np.maximum.reduceat(pid, offs[:-1]) if groups.n_groups else np.array([])
    pure = int((seg_min == seg_max).sum())
    pairs_hit = len(np.unique(seg_min[seg_min == seg_max]))
    gbs = FOF_BYTES_PER_PARTICLE * n / fof_s / 1e9
    log(f"bridged: {step_s * 1e3:.1f} ms/step (FoF {fof_s * 1e3:.2f} ms, {gbs:.0f} GB/s); "
        f"groups={groups.n_groups} pure={pure} pairs found={pairs_hit}")
    roofline = {"bound": "hbm", "achieved": round(gbs, 2), "peak": PEAK_HBM_GBS,
                "unit": "GB/s", "frac": round(gbs / PEAK_HBM_GBS, 5), "traffic": None,
                "kernel": "friends-of-friends labels (halma_fof_labels: every kernel of one "
                          "call, 16 B per particle algorithmic)",
                "launch_ms": round(fof_s * 1e3, 3),
                "unbinding_k_pot_sym": _sym_roofline(acc, args.steps)}
    out = _line("particles/s FoF+unbinding, 8M-particle bridged pairs", n / step_s,
                "particles/s", step_s, args,
                {"workload": "4000 Plummer pairs at 1-3 a joined by 5% bridges, 8M FP32, "
                             "groups>=50, physical-mode unbinding", "particles": n,
                 "groups": groups.n_groups, "passes": ub.stats["passes"]},
                roofline=roofline, groups_single_pair=pure, pairs_merged=pairs_hit,
                members=groups.n_members, pairs_evaluated=ub.stats["pairs"],
                bound_fraction=float(ub.bound.float().mean().item()))
    if not args.no_cpu:
        import oracle

        tc = time.perf_counter()
        want = oracle.fof_labels_cells(w["x"], w["y"], w["z"], ll)
        out["fof_labels_identical_full_size"] = bool(np.array_equal(lab.cpu().numpy(), want))
        log(f"FoF labels vs the oracle (all {n}): {out['fof_labels_identical_full_size']} "
            f"({time.perf_counter() - tc:.1f}s)")
        # the unbinding of the first 60 groups against the oracle (masks and passes)
        k = min(60, groups.n_groups)
        sub_off = offs[: k + 1]
        bnd, it, _ = oracle.unbind_iterative(sub_off, mem[: sub_off[-1]], w["x"], w["y"],
                                             w["z"], w["vx"], w["vy"], w["vz"], w["mass"], 1.0)
        out["unbinding_sample_identical"] = {
            "groups": int(k),
            "masks": bool(np.array_equal(ub.bound[: sub_off[-1]].cpu().numpy().astype(bool),
                                         bnd)),
            "passes": bool(np.array_equal(ub.iters[:k].cpu().numpy(), it))}
        sel = w["pair_id"] < 40
        out["cpu_baseline"] = _cpu_unbind_path(*(w[k][sel] for k in keys), ll, 1.0,
                                               args.cpu_runs, fof="bfs",
                                               label="the first 40 pairs: ")
        log(f"bridged cpu baseline: {out['cpu_baseline']['value']:.3e} particles/s")
    return out


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
        cs, ts, (_, cnt) = _median_runs(
            lambda: oracle.split_substructure(gl, w["x"], w["y"], w["z"], w["mass"], ll, 25,
                                              2.0, min_ll, 0.75, MINP), args.cpu_runs, warmup=0)
        extra["cpu_baseline"] = _cpu_entry(
            ns / cs, "members/s", 1, "port", cs, ts,
            f"the first {k} groups ({ns} members, {cnt} substructures): the oracle's "
            f"restatement, C cell-list FoF per group and level, one core; median of {len(ts)} "
            "runs")
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
        cs, ts, _ = _median_runs(
            lambda: oracle.phase_space_clean(gl, *(w[q] for q in keys), ll, 1.0, 3.0),
            args.cpu_runs, warmup=0)
        extra["cpu_baseline"] = _cpu_entry(
            ns / cs, "members/s", 1, "port", cs, ts,
            f"the first {k} groups ({ns} members): the oracle's restatement (sparse cells, "
            f"one group at a time), one core; median of {len(ts)} runs")
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
        cs, ts, want = _median_runs(
            lambda: oracle.progenitors(groups, oripa, new_mass[:k], old_lists, omm),
            args.cpu_runs, warmup=0)
        same = all((a[:k] == b).all() for a, b in zip((pro1, pro2, nmerg, mtype), want))
        ns = int(sum(len(g) for g in groups))
        extra["cpu_baseline"] = _cpu_entry(
            ns / cs, "new members/s", 1, "port", cs, ts,
            f"the first {k} new haloes ({ns} members) against all {old_npart.size} old "
            f"haloes: the oracle's restatement, np.isin per halo pair, one core; median of "
            f"{len(ts)} runs")
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
        def run():
            res = []
            for h in range(k):
                c, R = centres[h], radii[h]
                g = og.amr_to_particles(amr["L"], amr["ncoarse"], amr["levels"], amr["rx"],
                                        amr["ry"], amr["rz"], amr["cells"], R, *c, rho_B)
                keep = og.inside(g[0], g[1], g[2], R, *c)
                res.append((g[0][keep], sorted(dm_tree.query_ball_point(c, R)),
                            sorted(st_tree.query_ball_point(c, R))))
            return res

        cs, ts, res = _median_runs(run, args.cpu_runs, warmup=0)
        same = True
        for h, (gx, di, si) in enumerate(res):
            a, b = out["gas_offsets"][h], out["gas_offsets"][h + 1]
            same &= bool(np.array_equal(out["gas"]["x"][a:b].cpu().numpy(), gx))
            a, b = out["dm_offsets"][h], out["dm_offsets"][h + 1]
            same &= out["dm_index"][a:b].cpu().tolist() == di
            a, b = out["st_offsets"][h], out["st_offsets"][h + 1]
            same &= out["st_index"][a:b].cpu().tolist() == si
        extra["cpu_baseline"] = _cpu_entry(
            k / cs, "haloes/s", 1, "port", cs, ts,
            f"the first {k} haloes: AMR cells by the oracle's pure-Python loops, dark matter "
            f"and stars by scipy KD-trees built once (untimed), one core; median of "
            f"{len(ts)} runs")
        extra["sample_matches_oracle"] = same
    return _line("haloes/s source gathering (gas from AMR cells, dark matter, stars)",
                 H / step_s, "haloes/s", step_s, args,
                 {"workload": "1001 clump haloes, R = 3 half-mass radii; 2.1M stars, 6M dark "
                              "matter, 8.2M AMR cells (2 patches of 16^3 per halo)",
                  "haloes": H, "dm_particles": len(dm["x"]), "stars": len(w["x"]),
                  "amr_cells": grid.n_cells},
                 gas_found=n_gas, dm_found=n_dm, stars_found=n_st, **extra)


def bench_centring(args):
    """Differentiable centring: 200 Hernquist haloes of 5e4-5e5 particles (eps = 0.5 kpc).
    One step = the halo potential energy W_h = -G/2 sum m_i S_i of every halo with its
    autograd backward dW/dx (the softened sum and the hand-written acceleration kernel
    over the sum of n_h^2 pairs: with positions that require a gradient the forward launch
    evaluates both in one pass, and the backward launches nothing) + the potential-minimum
    centre of every halo from seeds 0.1-0.5 a off (autograd-driven descent).  Roofline: the
    fused pass (FP32 VALU; algorithmic work 12 FLOP per pair for S plus 20 for the
    acceleration), its time from HIP events on the stream around the forward call; the two
    kernels unfused (S alone, the acceleration alone) are timed once after the steps.
    Checks (CPU leg, the oracle's FP64 sums on the three smallest haloes): W, the
    gradient, and each centre against the FP64 brute-force minimum over the halo's
    particles."""
    import numpy as np
    import torch

    from pyhalma_amd import centring, synth

    t0 = time.time()
    w = synth.hernquist_haloes()
    off = w["offsets"]
    H = len(off) - 1
    n = int(off[-1])
    sizes = np.diff(off).astype(np.float64)
    pairs = float((sizes * (sizes - 1)).sum())
    log(f"centring workload: {H} haloes, {n} particles, {pairs:.3e} pairs "
        f"({time.time() - t0:.1f}s)")
    dev = torch.device("cuda", 0)
    X = [torch.from_numpy(w[k]).to(dev) for k in ("x", "y", "z")]
    m = torch.from_numpy(w["mass"]).to(dev)
    pos = torch.stack(X, 1).contiguous()
    eps = 5e-4
    evs = []

    def step():
        s = torch.cuda.current_stream()
        e = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
        p = pos.detach().requires_grad_(True)
        e[0].record(s)
        W = centring.potential_energy(p, m, off, eps)
        e[1].record(s)
        W.sum().backward()
        e[2].record(s)
        res = centring.refine_centres(X[0], X[1], X[2], m, off, w["seeds"], eps, max_iter=100,
                                      tol=1e-5)
        e[3].record(s)
        evs.append(e)
        return W.detach(), p.grad, res

    step_s, (W, grad, res) = _timed(step, args.steps, args.warmup)
    torch.cuda.synchronize()
    tm = np.array([[e[k].elapsed_time(e[k + 1]) for k in range(3)] for e in evs[args.warmup:]])
    fwd_s, bwd_s, cen_s = tm.mean(0) * 1e-3
    c = res.centres.cpu().numpy()
    err = np.linalg.norm(c - w["centres"], axis=1) / w["a"]
    it = res.iters.cpu().numpy()
    fused_flop = FLOP_PER_SOFT_PAIR + FLOP_PER_ACC_PAIR
    fwd_tf = fused_flop * pairs / fwd_s / 1e12
    # the two kernels unfused, once each (S alone; the acceleration alone)
    from pyhalma_amd import engine
    packed = torch.cat([pos, m[:, None]], 1).contiguous()
    s = torch.cuda.current_stream()
    e = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
    e[0].record(s)
    engine.pot_soft_self(off, packed, eps, pot=True)
    e[1].record(s)
    engine.pot_soft_self(off, packed, eps, pot=False, acc=True)
    e[2].record(s)
    torch.cuda.synchronize()
    pot_ms, acc_ms = e[0].elapsed_time(e[1]), e[1].elapsed_time(e[2])
    pot_tf = FLOP_PER_SOFT_PAIR * pairs / pot_ms / 1e9
    acc_tf = FLOP_PER_ACC_PAIR * pairs / acc_ms / 1e9
    log(f"centring: {step_s * 1e3:.1f} ms/step (forward with the acceleration "
        f"{fwd_s * 1e3:.1f} ms {fwd_tf:.1f} TFLOP/s, backward {bwd_s * 1e3:.2f} ms, centres "
        f"{cen_s * 1e3:.2f} ms; unfused: S {pot_ms:.1f} ms {pot_tf:.1f} TFLOP/s, acceleration "
        f"{acc_ms:.1f} ms {acc_tf:.1f} TFLOP/s); iterations max {it.max()} median "
        f"{np.median(it)}; |c - c_true|/a median {np.median(err):.2e}")
    roofline = {"bound": "valu", "achieved": round(fwd_tf, 3), "peak": PEAK_FP32_TFLOPS,
                "unit": "TFLOP/s", "frac": round(fwd_tf / PEAK_FP32_TFLOPS, 4),
                "traffic": None, "roof": VALU_ROOF,
                "kernel": "k_soft_tile<S and acceleration, tests = the halo's members> (the "
                          "potential-energy forward with the gradient fused)",
                "flop_per_pair": fused_flop, "pairs_per_launch": pairs,
                "launch_ms": round(fwd_s * 1e3, 3), "launches_per_step": 1,
                "unfused": {"S": {"ms": pot_ms, "flop_per_pair": FLOP_PER_SOFT_PAIR,
                                  "tflops": pot_tf, "frac": pot_tf / PEAK_FP32_TFLOPS},
                            "acceleration": {"ms": acc_ms, "flop_per_pair": FLOP_PER_ACC_PAIR,
                                             "tflops": acc_tf,
                                             "frac": acc_tf / PEAK_FP32_TFLOPS}}}
    out = _line("particles/s halo potential energy forward + backward + potential-minimum "
                "centres (200 Hernquist haloes)", n / step_s, "particles/s", step_s, args,
                {"workload": "200 Hernquist haloes of 5e4-5e5 (log-uniform), FP32, eps 0.5 "
                             "kpc; W forward + autograd backward over sum n^2 pairs, then "
                             "centres from seeds 0.1-0.5 a off",
                 "particles": n, "haloes": H, "pairs": pairs},
                roofline=roofline,
                phases_ms={"forward": fwd_s * 1e3, "backward": bwd_s * 1e3,
                           "centres": cen_s * 1e3},
                iterations_max=int(it.max()), iterations_median=float(np.median(it)),
                offset_over_a_median=float(np.median(err)),
                offset_over_a_max=float(err.max()))
    if not args.no_cpu:
        import oracle

        nthreads = _threads()
        G = centring.G_CONST
        checks = []
        Wn = W.cpu().numpy()
        gn = grad.cpu().numpy()
        for h in np.argsort(sizes)[:3]:
            a, b = int(off[h]), int(off[h + 1])
            args4 = (w["mass"][a:b], w["x"][a:b], w["y"][a:b], w["z"][a:b])
            t64 = [w[k][a:b].astype(np.float64) for k in ("x", "y", "z")]
            S = oracle.pot_soft_f64(*args4, *t64, eps, nthreads=nthreads)
            W64 = -0.5 * G * float(np.sum(w["mass"][a:b].astype(np.float64) * S))
            A = oracle.acc_soft_f64(*args4, *t64, eps, nthreads=nthreads)
            g64 = -G * w["mass"][a:b, None].astype(np.float64) * A
            gerr = float(np.abs(gn[a:b] - g64).max() / np.abs(g64).max())
            Sc = float(oracle.pot_soft_f64(*args4, *(c[h:h + 1, q] for q in range(3)), eps)[0])
            k = int(np.argmax(S))
            checks.append({"halo": int(h), "members": b - a,
                           "W_rel_err": abs(Wn[h] - W64) / abs(W64),
                           "grad_err_over_max": gerr,
                           "centre_deeper_than_every_particle": bool(Sc >= S[k]),
                           "centre_to_deepest_particle_over_eps": float(
                               np.linalg.norm(c[h] - np.array([t[k] for t in t64])) / eps)})
        out["fp64_checks"] = checks
        # the reference's own FP32 pair routine (unsoftened: the softened sum has no
        # reference counterpart) over the smallest halo, twice per step (forward and
        # backward pass over the same pairs)
        use_ref = oracle.have_ref()
        pot = (lambda *a: oracle.ref_pot_sum_f32(*a, nthreads=nthreads)) if use_ref else \
            (lambda *a: oracle.pot_sum_f32(*a, nthreads=nthreads))
        h = int(np.argsort(sizes)[0])
        a, b = int(off[h]), int(off[h + 1])
        q = [w[k][a:b] for k in ("mass", "x", "y", "z")]
        med, ts, _ = _median_runs(lambda: pot(*q, q[1], q[2], q[3]), args.cpu_runs)
        out["cpu_baseline"] = _cpu_entry(
            (b - a) / (2 * med), "particles/s", nthreads, "reference" if use_ref else "port",
            med, ts,
            f"the smallest halo ({b - a} particles, {float(b - a) ** 2:.3e} pairs): the "
            f"{'reference Fortran pair routine' if use_ref else 'C port of it'} (FP32, "
            f"unsoftened, OpenMP {nthreads} threads) over its members, counted twice per "
            f"step (forward and backward); median of {len(ts)} runs after a warm-up")
        log(f"centring checks: {checks}")
    return out


if __name__ == "__main__":
    main()
