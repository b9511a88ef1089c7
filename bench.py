"""Throughput of the LMFF force path (Tersoff + ILP with normals and Tap) on one MI355X.

Default workload (the line the metric is quoted on): the 10-layer alternating graphene/hBN
stack, 260x260 rectangular cells per layer (2,704,000 atoms, layer spacing 3.35 A), seeded
0.02 A jitter, 300 K Maxwell-Boltzmann velocities, NVE dt = 1 fs, neighbor rebuild every 10
steps with a 2 A skin, ILP 16 A with Tap.  A "step" is one Velocity-Verlet step of the whole
system (force evaluation + integration + the amortised neighbor rebuild), with all state
resident in HBM.  The FP64 path gives `value`; the mixed-precision path (FP32 ILP pair math,
FP64 displacements, Tersoff and accumulation) is timed on the same input and reported beside
it with its force error against FP64.  After the timed steps each precision runs a separate
1,000-step NVE trajectory (--drift-steps) whose total energy, sampled every 10 steps, is fitted
by a straight line: the slope is the drift.  The line also carries the 228,800-atom
graphene/hBN bilayer (FP64 and mixed step times) and a short twisted-bilayer job stream
(packed vs sequential throughput, p50/p99 latency).

Other workloads (not the default line):
  --workload multilayer   the same measurement on another stack: --layers L --nx NX --ny NY,
                          e.g. --layers 4 --nx 1118 --ny 1118 --drift-steps 100 for the
                          19,998,784-atom 4-layer stack
  --workload bilayer      the 228,800-atom heterobilayer alone, with its kernel roofline
  --workload serving      the twisted-bilayer Langevin job stream (--jobs, --steps)

Prints one JSON line.  python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "lammps-lmff_amd"))

from lmff import capi, structures as S  # noqa: E402

HBM_PEAK_GBS = 8000.0
FP64_VALU_PEAK = 78.6   # TFLOP/s, MI355X FP64 vector
FP32_VALU_PEAK = 157.3  # TFLOP/s, MI355X FP32 vector
ILP_PAIR_FLOP = 120     # algorithmic FLOP per ILP ordered pair (exp, sqrt, divide count 1 each)


def log(msg):
    """Progress on stderr (the result line is the only stdout output)."""
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


def build_ctx(s, precision):
    ctx = capi.Context(0)
    capi.setup_for(ctx, s, precision=precision, skin=2.0, every=10, delay=0, check=0)
    sid = ctx.add_system(s)
    ctx.integrator(sid, capi.LMFF_NVE, dt=0.001)
    return ctx, sid


def one_shot(s, precision, repeat=1):
    """Pair::compute through the host-buffer entry point (x in, f/energy/virial out); with
    repeat > 1 also the median wall times (PCIe transfers included) of later calls with a fresh
    list (new_list=1, LAMMPS' reneighbouring steps) and of as many calls reusing it."""
    ctx = capi.Context(0)
    try:
        capi.setup_for(ctx, s, precision=precision)
        r = ctx.pair_compute(s, 1, 2)
        walls = {1: [], 0: [], "f": []}
        for nl in (1, 0):
            for _ in range(repeat - 1):
                t0 = time.perf_counter()
                ctx.pair_compute(s, 1, 2, new_list=nl)
                walls[nl].append(time.perf_counter() - t0)
        for _ in range(repeat - 1):
            t0 = time.perf_counter()
            ctx.pair_compute(s, 0, 0, new_list=0)
            walls["f"].append(time.perf_counter() - t0)
        if walls[1]:
            r["wall_s"] = float(np.median(walls[1]))
            r["wall_reuse_s"] = float(np.median(walls[0]))
            r["wall_reuse_force_s"] = float(np.median(walls["f"]))
        return r
    finally:
        ctx.close()


def timed_run(ctx, steps, warmup, profile):
    """The timed run is the production path: the steps between rebuilds replayed from a
    captured hipGraph, HIP events on the engine's stream only around the whole run (lmff_run
    synchronises before it returns).  With `profile`, a second run of the same length
    afterwards records events around every kernel family on that stream for the per-launch
    kernel times (its own total, which the markers inflate, is not the step time)."""
    ctx.evaluate(1, 2)
    if warmup:
        ctx.run(warmup, 0, use_graph=True)
    t0 = time.perf_counter()
    ctx.run(steps, steps, use_graph=True)
    wall = time.perf_counter() - t0
    ms_total, _, _ = ctx.profile_read()
    out = {"wall_s": wall, "gpu_ms": ms_total, "fam_ms": None, "fam_n": None}
    if profile:
        ctx.profile(True)
        ctx.run(steps, steps, use_graph=False)
        ms_prof, fam_ms, fam_n = ctx.profile_read()
        ctx.profile(False)
        out.update({"fam_ms": fam_ms.tolist(), "fam_n": fam_n.tolist(), "gpu_ms_profiled": ms_prof})
    return out


def drift_run(ctx, sid, natoms, nsteps, dt_ps=0.001, every=10):
    """NVE energy drift: total energy (pair + kinetic) every `every` steps over an nsteps run,
    fitted by least squares against time; slope in eV/atom/ps with its standard error, and the
    RMS of the residual fluctuation.  Run in segments of at most 1,000 steps (the trajectory
    and its step count carry on from one to the next), each reported on stderr."""
    rec, ms_total, done = [], 0.0, 0
    while done < nsteps:
        seg = min(1000, nsteps - done)
        rec += ctx.run(seg, every, use_graph=True)[0]
        ms_total += ctx.profile_read()[0]
        done += seg
        if nsteps > 1000:
            log(f"  drift run: {done} of {nsteps} steps")
    t = np.array([r["step"] for r in rec], dtype=np.float64) * dt_ps
    e = np.array([r["pe"] + r["ke"] for r in rec]) / natoms
    A = np.stack([t - t.mean(), np.ones_like(t)], 1)
    coef, *_ = np.linalg.lstsq(A, e, rcond=None)
    res = e - A @ coef
    dof = max(len(t) - 2, 1)
    se = float(np.sqrt((res ** 2).sum() / dof / ((t - t.mean()) ** 2).sum())) if len(t) > 2 else float("nan")
    return {"steps": nsteps, "samples": len(t), "slope_eV_per_atom_per_ps": float(coef[0]), "slope_stderr": se,
            "fluctuation_rms_eV_per_atom": float(np.sqrt((res ** 2).mean())),
            "e_first_eV_per_atom": float(e[0]), "e_last_eV_per_atom": float(e[-1]),
            "temp_last": float(rec[-1]["temp"]), "gpu_ms_per_step": ms_total / nsteps}


def cpu_baseline(s, seconds_budget=20.0):
    """The oracle (FP64 C restatement of the reference path) on this host's cores."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle  # noqa: E402

    oracle.build()
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    pot = capi.POTENTIALS
    o = oracle.Oracle(3, os.path.join(pot, "BNC.tersoff"), s.elements, os.path.join(pot, "BNCH.ILP"), s.elements,
                      threads=cores)
    o.set_system(s)
    t0 = time.perf_counter()
    o.build_list()
    t_list = time.perf_counter() - t0
    f_ref = None
    steps = 0
    t_force = 0.0
    o.compute_lmff(1)  # untimed: maps the per-thread force arrays
    while t_force < seconds_budget or steps == 0:
        t1 = time.perf_counter()
        o.compute_lmff(1)
        t_force += time.perf_counter() - t1
        steps += 1
        if steps >= 3:
            break
    t2 = time.perf_counter()
    f_ref = o.compute(1, 2)
    t_stock = time.perf_counter() - t2
    per_step = t_force / steps + t_list / 10.0  # list rebuild amortised over 10 steps
    return {"value": s.n / per_step, "unit": "atom-steps/s", "cores": cores, "kind": "port",
            "sample": f"reference-LMFF CPU restatement (the fused path of LMFF/pair_lmff.cpp:1121-1613, "
                      f"oracle/lmff_oracle.c orc_compute_lmff, OpenMP threads = host cores): {steps} force "
                      f"evaluations + 1 list build of the same {s.n}-atom system, integration excluded",
            "stock_oracle_value": s.n / (t_stock + t_list / 10.0),
            "stock_oracle_note": "the stock tersoff + ilp/graphene/hbn push form (orc_compute), the parity checker",
            "cpu_model": _cpu_model()}, f_ref


def patch_baseline(s, f_gpu, seed, frac=0.01, halo=24.0):
    """cpu_baseline leg of the multilayer workloads: the oracle on a seeded rectangular patch
    covering `frac` of the periodic x-y box, cut out with a `halo`-wide rim (open boundaries).
    The force on an atom depends only on atoms within rc + two normal-neighbour hops (16 + 2 +
    2 A < 24 A), so every atom of the patch interior sees the environment it has in the full
    system: the oracle's forces there are the reference values for the GPU's FP64 forces on
    those atoms (the spot check), and its time is the CPU baseline for this sample."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle  # noqa: E402

    oracle.build()
    rng = np.random.default_rng(seed)
    L = s.boxhi - s.boxlo
    c = s.boxlo[:2] + rng.random(2) * L[:2]
    w = 0.5 * L[:2] * np.sqrt(frac)
    d = s.x[:, :2] - c
    d -= L[:2] * np.round(d / L[:2])
    keep = np.nonzero((np.abs(d) < w + halo).all(1))[0]
    inner = (np.abs(d[keep]) < w).all(1)
    x = s.x[keep].copy()
    x[:, :2] = c + d[keep]
    lo = np.array([c[0] - w[0] - halo - 1.0, c[1] - w[1] - halo - 1.0, s.boxlo[2]])
    hi = np.array([c[0] + w[0] + halo + 1.0, c[1] + w[1] + halo + 1.0, s.boxhi[2]])
    sub = S.System(x=x, type=s.type[keep].copy(), tag=s.tag[keep].copy(), mol=s.mol[keep].copy(), boxlo=lo,
                   boxhi=hi, periodic=(0, 0, s.periodic[2]), elements=list(s.elements))
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    pot = capi.POTENTIALS
    o = oracle.Oracle(3, os.path.join(pot, "BNC.tersoff"), s.elements, os.path.join(pot, "BNCH.ILP"), s.elements,
                      threads=cores)
    o.set_system(sub)
    t0 = time.perf_counter()
    o.build_list()
    t_list = time.perf_counter() - t0
    # the timed path: the reference's fused LMFF CPU path restated (orc_compute_lmff), timed on
    # its second call (the first maps the per-thread force arrays); the stock four-pass oracle is
    # the checker and is timed beside it
    o.compute_lmff(1)
    t1 = time.perf_counter()
    o.compute_lmff(1)
    t_force = time.perf_counter() - t1
    t2 = time.perf_counter()
    r = o.compute(1, 0)
    t_stock = time.perf_counter() - t2
    fr = r["f"][inner]
    fg = f_gpu[keep[inner]]
    rms = float(np.sqrt((fr ** 2).sum(1).mean()))
    err = {"atoms_checked": int(inner.sum()), "fraction_of_system": float(inner.sum()) / s.n,
           "max_abs": float(np.abs(fg - fr).max()), "rel_max_over_rms": float(np.abs(fg - fr).max()) / rms,
           "patch_seed": seed}
    per_step = t_force + t_list / 10.0
    base = {"value": sub.n / per_step, "unit": "atom-steps/s", "cores": cores, "kind": "port",
            "sample": f"reference-LMFF CPU restatement (the fused path of LMFF/pair_lmff.cpp:1121-1613, "
                      f"oracle/lmff_oracle.c orc_compute_lmff, OpenMP threads = host cores): 1 force "
                      f"evaluation + 1 list build (amortised over 10 steps) of a {sub.n}-atom patch (seeded, "
                      f"{frac:.0%} of the box plus a {halo:g} A rim) of the same stack, integration excluded",
            "stock_oracle_value": sub.n / (t_stock + t_list / 10.0),
            "stock_oracle_note": "the stock tersoff + ilp/graphene/hbn push form (orc_compute), the parity checker",
            "cpu_model": _cpu_model()}
    return base, err


def profiled_traffic(kernel, tag):
    """HBM bytes per launch of the first kernel whose name starts with `kernel` in the committed
    rocprofv3 FETCH_SIZE/WRITE_SIZE summary profiles/hbm_traffic_<tag>.md (calibrated against a
    plain copy), or None when no summary covers this workload.  Read from the file, not
    measured in this run (counters need their own profiler pass)."""
    path = os.path.join(REPO, "profiles", f"hbm_traffic_{tag}.md")
    try:
        lines = open(path).read().splitlines()
    except OSError:
        return None
    for line in lines:
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) == 6 and cells[0].startswith(f"`{kernel}"):
            return (float(cells[4]) + float(cells[5])) * 1e6
    return None


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


FAMILIES = ["list_build", "normals", "ilp_pair", "tersoff_bonds", "gather", "integrate", "reduce"]


def precision_runs(s, steps, warmup, drift_steps, profile=True, host_repeat=1):
    """FP64 and mixed contexts on the same input: single-shot forces (host-buffer
    Pair::compute), then per precision the timed resident run, the profiled run of the same
    length and the drift trajectory."""
    log(f"{s.n} atoms: single-shot Pair::compute, FP64 and mixed")
    r64 = one_shot(s, capi.LMFF_FP64, repeat=host_repeat)
    rmx = one_shot(s, capi.LMFF_MIXED)
    df = rmx["f"] - r64["f"]
    mixed_err = {"max_abs": float(np.abs(df).max()), "rms": float(np.sqrt((df ** 2).sum(1).mean()))}
    res = {}
    stats = None
    for name, prec in (("fp64", capi.LMFF_FP64), ("mixed", capi.LMFF_MIXED)):
        log(f"{name} resident run: {warmup} warmup + {steps} timed steps"
            + (f", {steps} profiled, {drift_steps}-step drift run" if profile else ""))
        ctx, sid = build_ctx(s, prec)
        if stats is None:
            stats = ctx.pair_stats()
        res[name] = timed_run(ctx, steps, warmup, profile)
        if drift_steps:
            res[name]["drift"] = drift_run(ctx, sid, s.n, drift_steps)
        ctx.close()
    log("resident runs done")
    return r64, rmx, stats, res, mixed_err


def roofline_ilp(stats, ilp_ms, natoms, tag):
    """Roofline of the dominant kernel, the FP64 ILP pair evaluation: the half-pair kernel
    k_ilp_half<double> (each pair once, partner terms in fixed point) with k_ilp_combine (the
    partner sums into each atom), timed together by the HIP events around the ILP step on the
    engine's stream.  Algorithmic FLOP per launch: 120 per ordered pair within the cutoff (exp,
    sqrt and divide count 1 each): the repulsion and taper of each direction (about 102) and half
    of the pair's van der Waals term, which is what the half-pair kernel evaluates once per
    unordered pair (both repulsion directions, one van der Waals term).  Algorithmic bytes per
    launch: 4 B per (ordered) list entry plus 128 B per atom (x, normal, force and S in and out)."""
    flop = stats["ilp_pairs"] * ILP_PAIR_FLOP
    achieved = flop / (ilp_ms * 1e-3) / 1e12
    t_half = profiled_traffic("k_ilp_half<double, false, false", tag)
    t_comb = profiled_traffic("k_ilp_combine<double, false, false", tag)
    traffic = t_half + t_comb if t_half is not None and t_comb is not None else None
    return {"bound": "valu", "kernel": "k_ilp_half<double> + k_ilp_combine<double> (FP64 ILP pairs, force-only variants)",
            "achieved": achieved, "peak": FP64_VALU_PEAK, "unit": "TFLOP/s", "frac": achieved / FP64_VALU_PEAK,
            "traffic": traffic,
            "traffic_source": (f"profiles/hbm_traffic_{tag}.md (rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE of "
                               f"this workload, copy-calibrated; read from the file, not this run)")
            if traffic is not None else None,
            "kernel_ms": ilp_ms, "flop_per_launch": flop,
            "algorithmic_bytes_per_launch": 4.0 * stats["ilp_entries"] + 128.0 * natoms,
            "ilp_pairs": stats["ilp_pairs"], "ilp_list_entries": stats["ilp_entries"]}


def stack_workload(layers, nx, ny, seed):
    kinds = ["G" if k % 2 == 0 else "BN" for k in range(layers)]
    s = S.layered_stack(nx, ny, kinds, seed=seed)
    desc = (f"{layers}-layer alternating graphene/hBN stack, {nx}x{ny} cells/layer, {s.n} atoms, spacing 3.35 A, "
            f"NVE dt=1 fs, rebuild every 10 steps, skin 2 A, ILP 16 A with Tap")
    return s, desc


def stack_line(args, layers, nx, ny, seed, extras=True):
    """The metric's measurement on a layered stack: FP64 `value`, mixed beside it with its force
    error, fitted drift of both, the ILP kernel roofline, the CPU baseline on a 1 % patch with
    the GPU's FP64 spot check there; with extras, the bilayer and a short job stream."""
    s, desc = stack_workload(layers, nx, ny, seed)
    n = s.n
    r64, rmx, stats, res, mixed_err = precision_runs(s, args.steps, args.warmup, args.drift_steps)
    ms64 = res["fp64"]["gpu_ms"] / args.steps
    msmx = res["mixed"]["gpu_ms"] / args.steps
    d64 = res["fp64"].get("drift")
    dmx = res["mixed"].get("drift")
    drift = None
    if d64 and dmx:
        drift = {"fp64": d64, "mixed": dmx,
                 "mixed_over_fp64_abs_slope": abs(dmx["slope_eV_per_atom_per_ps"]) / abs(d64["slope_eV_per_atom_per_ps"])
                 if d64["slope_eV_per_atom_per_ps"] != 0 else None,
                 "note": "least-squares slope of (pe + ke)/N against time over the run, sampled every 10 steps"}
    tag = f"multilayer_{n}"
    out = {
        "metric": "atom-steps/s FP64 vs mixed at RMS force err<=1e-5 eV/A; job p50/p99 latency",
        "value": n / (ms64 * 1e-3),
        "unit": "atom-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms64,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded generator: jittered lattice, 300 K Maxwell-Boltzmann velocities)",
        "config": {"workload": desc, "natoms": n, "precision": "fp64"},
        "mixed": {"value": n / (msmx * 1e-3), "ms_per_step": msmx, "speedup_vs_fp64": ms64 / msmx,
                  "force_err_vs_fp64": mixed_err, "ilp_kernel_ms": res["mixed"]["fam_ms"][2],
                  "note": "FP32 ILP pair and normal math from FP64 displacements; Tersoff, list build, "
                          "integration and all accumulation in FP64"},
        "drift": drift,
        "kernel_ms_per_launch": {"fp64": dict(zip(FAMILIES, res["fp64"]["fam_ms"])),
                                 "mixed": dict(zip(FAMILIES, res["mixed"]["fam_ms"]))},
        "roofline": roofline_ilp(stats, res["fp64"]["fam_ms"][2], n, tag),
        "cpu_baseline": None,
        "oracle_force_error": None,
        "wall_s": {"fp64": res["fp64"]["wall_s"], "mixed": res["mixed"]["wall_s"]},
    }
    if not args.no_cpu:
        log("CPU baseline and FP64 spot check: oracle on a seeded 1 % patch")
        out["cpu_baseline"], out["oracle_force_error"] = patch_baseline(s, r64["f"], seed=20240 + layers)
    if extras:
        out["bilayer_228800"] = bilayer_summary(args)
        if not args.no_serving:
            out["serving"] = serving_short(args)
    return out


def bilayer_summary(args):
    """The 228,800-atom graphene/hBN heterobilayer (220x130 cells per layer): FP64 and mixed
    step times over the same number of steps (no drift run, no profiling)."""
    s = S.layered_stack(220, 130, ["G", "BN"], seed=12346)
    res = {}
    for name, prec in (("fp64", capi.LMFF_FP64), ("mixed", capi.LMFF_MIXED)):
        ctx, _ = build_ctx(s, prec)
        res[name] = timed_run(ctx, args.steps, args.warmup, profile=False)["gpu_ms"] / args.steps
        ctx.close()
    log("bilayer done")
    return {"workload": f"graphene/hBN heterobilayer 220x130 cells/layer, {s.n} atoms, NVE dt=1 fs, rebuild every 10 "
                        f"steps, skin 2 A, ILP 16 A with Tap", "fp64_ms_per_step": res["fp64"],
            "mixed_ms_per_step": res["mixed"], "fp64_value": s.n / (res["fp64"] * 1e-3),
            "mixed_value": s.n / (res["mixed"] * 1e-3), "speedup_mixed_vs_fp64": res["fp64"] / res["mixed"]}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--drift-steps", type=int, default=1000,
                    help="length of the NVE trajectory per precision whose energy slope is the drift (0: none)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-serving", action="store_true", help="default line: skip the short job stream")
    ap.add_argument("--no-extras", action="store_true", help="default line: skip the bilayer and the job stream")
    ap.add_argument("--workload", choices=["stack", "multilayer", "bilayer", "serving"], default="stack")
    ap.add_argument("--nx", type=int, default=None)
    ap.add_argument("--ny", type=int, default=None)
    ap.add_argument("--layers", type=int, default=10, help="multilayer: number of layers")
    ap.add_argument("--jobs", type=int, default=128, help="serving: number of jobs")
    ap.add_argument("--serve-jobs", type=int, default=64, help="default line: jobs in the short stream")
    ap.add_argument("--serve-steps", type=int, default=2000, help="default line: steps per job of the short stream")
    ap.add_argument("--rate", type=float, default=0.0,
                    help="serving: Poisson arrival rate in jobs/s for the latency run (0: derived from capacity)")
    ap.add_argument("--copy-ref", type=int, default=0,
                    help="only run a plain device copy of this many bytes 5 times (counter calibration)")
    args = ap.parse_args()
    if args.copy_ref:
        import torch

        a = torch.empty(args.copy_ref // 4, dtype=torch.float32, device="cuda").fill_(1.0)
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_ref_bytes": args.copy_ref}))
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the multi-GPU layout is a design "
              f"note, not a measured path", file=sys.stderr)
        sys.exit(2)
    if args.workload == "serving":
        print(json.dumps(serving_bench(args)))
        return
    if args.workload == "bilayer":
        print(json.dumps(bilayer_bench(args)))
        return
    if args.workload == "multilayer":
        print(json.dumps(stack_line(args, args.layers, args.nx or 260, args.ny or 260, 12347, extras=False)))
        return
    print(json.dumps(stack_line(args, 10, 260, 260, 12347, extras=not args.no_extras)))


def bilayer_bench(args):
    """The 228,800-atom heterobilayer on its own: FP64 value, mixed beside it, the ILP kernel
    roofline at this size and the oracle on the whole system as CPU baseline."""
    nx = args.nx or 220
    ny = args.ny or 130
    s = S.layered_stack(nx, ny, ["G", "BN"], seed=12346)
    n = s.n
    r64, rmx, stats, res, mixed_err = precision_runs(s, args.steps, args.warmup, args.drift_steps, host_repeat=6)
    ms64 = res["fp64"]["gpu_ms"] / args.steps
    msmx = res["mixed"]["gpu_ms"] / args.steps
    out = {
        "metric": "atom-steps/s FP64 vs mixed (bilayer)", "value": n / (ms64 * 1e-3), "unit": "atom-steps/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms64, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic (seeded generator)",
        "config": {"workload": f"graphene/hBN heterobilayer {nx}x{ny} cells/layer, {n} atoms, NVE dt=1 fs, "
                               f"rebuild every 10 steps, skin 2 A, ILP 16 A with Tap", "natoms": n},
        "mixed": {"value": n / (msmx * 1e-3), "ms_per_step": msmx, "speedup_vs_fp64": ms64 / msmx,
                  "force_err_vs_fp64": mixed_err},
        "drift": {"fp64": res["fp64"].get("drift"), "mixed": res["mixed"].get("drift")},
        "kernel_ms_per_launch": {"fp64": dict(zip(FAMILIES, res["fp64"]["fam_ms"])),
                                 "mixed": dict(zip(FAMILIES, res["mixed"]["fam_ms"]))},
        "roofline": roofline_ilp(stats, res["fp64"]["fam_ms"][2], n, f"bilayer_{n}"),
        "cpu_baseline": None, "oracle_force_error": None,
        "host_buffer_pair_compute": {
            "note": "Pair::compute from host arrays, FP64, energy + virial: with a fresh list (upload "
                    "x/type/tag/mol, list build, forces, download f) and reusing it (upload x, forces, "
                    "download f); per step of a LAMMPS loop that rebuilds every 10 steps; PCIe included",
            "wall_ms": r64["wall_s"] * 1e3, "wall_ms_reuse": r64["wall_reuse_s"] * 1e3,
            "wall_ms_reuse_forces_only": r64["wall_reuse_force_s"] * 1e3,
            "ms_per_step_every10": (r64["wall_s"] + 9 * r64["wall_reuse_s"]) / 10 * 1e3,
            "atom_steps_per_s_every10": n / ((r64["wall_s"] + 9 * r64["wall_reuse_s"]) / 10)},
    }
    if not args.no_cpu:
        log("CPU baseline (oracle)")
        cpu, fref = cpu_baseline(s)
        d = r64["f"] - fref["f"]
        rmsf = np.sqrt((fref["f"] ** 2).sum(1).mean())
        out["cpu_baseline"] = cpu
        out["oracle_force_error"] = {
            "fp64_max_abs": float(np.abs(d).max()), "fp64_rel_max_over_rms": float(np.abs(d).max() / rmsf),
            "fp64_energy_rel": float(abs(r64["energy"] - fref["energy"]) / abs(fref["energy"])),
            "mixed_rms": float(np.sqrt(((rmx["f"] - fref["f"]) ** 2).sum(1).mean()))}
    return out


def _fresh(js, rate=None, seed=77):
    """Copies of jobs with Poisson arrivals at `rate` (all at t=0 when rate is None)."""
    from lmff import serving

    out = []
    t = 0.0
    rng = np.random.default_rng(seed)
    for j in js:
        if rate:
            t += rng.exponential(1.0 / rate)
        out.append(serving.Job(job_id=j.job_id, system=j.system, steps=j.steps, arrival=t, seed=j.seed))
    return out


def _progress(label, every_s=30.0):
    """on_tick callback of Server.run: a progress line at most every every_s seconds."""
    last = [time.perf_counter()]

    def cb(srv, now):
        t = time.perf_counter()
        if t - last[0] >= every_s:
            last[0] = t
            log(f"{label}: {len(srv.resident)} jobs resident, {now:.0f} s into the stream")
    return cb


def sequential_direct(js, precision=capi.LMFF_MIXED, tick=10):
    """The jobs one at a time through the plain single-system path: one context, each job added,
    advanced by one lmff_run call over all its steps (graph-replayed, thermo at the end only)
    and removed.  The strictest sequential baseline: no per-tick host round trips."""
    ctx = capi.Context(0)
    try:
        capi.setup_graphene(ctx, precision=precision, every=tick)
        warm = js[0]
        sid = ctx.add_system(warm.system)
        ctx.integrator(sid, warm.style, dt=0.001, t_target=warm.t_target, damp=warm.damp, seed=warm.seed)
        ctx.run(2 * tick, 0, use_graph=True)
        ctx.remove_system(sid)
        t0 = time.perf_counter()
        last = t0
        atom_steps = 0
        for k, j in enumerate(js):
            sid = ctx.add_system(j.system)
            ctx.integrator(sid, j.style, dt=0.001, t_target=j.t_target, damp=j.damp, seed=j.seed)
            ctx.run(j.steps, j.steps, use_graph=True)
            ctx.remove_system(sid)
            atom_steps += j.natoms * j.steps
            if time.perf_counter() - last >= 30.0:
                last = time.perf_counter()
                log(f"one at a time: {k + 1} of {len(js)} jobs done")
        wall = time.perf_counter() - t0
        return {"makespan_s": wall, "atom_steps_per_s": atom_steps / wall}
    finally:
        ctx.close()


def stream_capacity(js, tick=10):
    """Every job queued at t=0: served packed (one launch per kernel per step for all resident
    jobs), one at a time through the same server code (max one resident job), and one at a
    time through the plain single-system path."""
    from lmff import serving

    srv = serving.Server(precision=capi.LMFF_MIXED, hbm_budget_bytes=200e9, tick=tick)
    srv.run(_fresh(js[:2]))   # warm the context (module load, first graph capture)
    packed = srv.run(_fresh(js), on_tick=_progress("packed"))
    srv.close()
    log(f"packed: {packed['makespan_s']:.1f} s; one at a time through the server")
    seq_srv = serving.Server(precision=capi.LMFF_MIXED, hbm_budget_bytes=200e9, tick=tick, max_resident=1)
    seq_srv.run(_fresh(js[:1]))
    seq = seq_srv.run(_fresh(js), on_tick=_progress("one at a time through the server"))
    seq_srv.close()
    log(f"through the server: {seq['makespan_s']:.1f} s; one at a time, plain path")
    direct = sequential_direct(_fresh(js), tick=tick)
    return packed, seq, direct


def stream_latency(js, rate, tick=10):
    from lmff import serving

    srv = serving.Server(precision=capi.LMFF_MIXED, hbm_budget_bytes=200e9, tick=tick)
    srv.run(_fresh(js[:2]))
    lat = srv.run(_fresh(js, rate=rate), on_tick=_progress("latency run"))
    srv.close()
    return lat


def _capacity_dict(packed, seq, direct):
    return {"packed_atom_steps_per_s": packed["atom_steps_per_s"],
            "sequential_atom_steps_per_s": direct["atom_steps_per_s"],
            "speedup": packed["atom_steps_per_s"] / direct["atom_steps_per_s"],
            "sequential_through_server_atom_steps_per_s": seq["atom_steps_per_s"],
            "speedup_vs_sequential_through_server": packed["atom_steps_per_s"] / seq["atom_steps_per_s"],
            "packed_makespan_s": packed["makespan_s"], "sequential_makespan_s": direct["makespan_s"],
            "packed_max_resident": packed["max_resident"],
            "packed_gpu_busy_fraction": packed["gpu_busy_s"] / packed["makespan_s"],
            "packed_ticks": packed["ticks"],
            "packed_admit_s": packed["admit_s"],
            "sequential_through_server_gpu_busy_fraction": seq["gpu_busy_s"] / seq["makespan_s"],
            "note": "sequential = each job alone through the plain single-system path (one lmff_run call "
                    "per job); the server-code sequential run (max one resident job, a host round trip "
                    "every tick) is given beside it"}


def serving_short(args):
    """A short slice of the job stream for the default line: the first --serve-jobs jobs of the
    seeded twisted-bilayer stream (2k-80k atoms, Langevin 300 K, damp 1 ps, mixed precision),
    --serve-steps steps each: capacity packed vs sequential, then p50/p99 latency under
    Poisson arrivals at half the packed capacity."""
    from lmff import serving

    jobs = serving.tbg_job_stream(args.serve_jobs, rate_hz=1e9, steps=args.serve_steps, seed=2024)
    sizes = np.array([j.natoms for j in jobs])
    log(f"job stream: {len(jobs)} jobs x {args.serve_steps} steps, packed then one at a time")
    packed, seq, direct = stream_capacity(jobs)
    rate = 0.5 * packed["jobs"] / packed["makespan_s"]
    log(f"job stream latency run at {rate:.2f} jobs/s")
    lat = stream_latency(jobs, rate)
    return {"workload": f"first {len(jobs)} jobs of the seeded twisted-bilayer stream (m,n) 1<=n<m<=12, "
                        f"{sizes.min()}-{sizes.max()} atoms (median {int(np.median(sizes))}), {args.serve_steps} "
                        f"Langevin steps each, 300 K, damp 1 ps, dt 1 fs, mixed precision, admission every 10 steps",
            "capacity": _capacity_dict(packed, seq, direct),
            "latency": {"arrival_rate_jobs_per_s": rate, "p50_s": lat["latency_p50_s"], "p99_s": lat["latency_p99_s"],
                        "atom_steps_per_s": lat["atom_steps_per_s"], "max_resident": lat["max_resident"]}}


def serving_bench(args):
    """Twisted-bilayer Langevin jobs (mixed precision, 300 K, tau 1 ps): capacity packed vs
    sequential, on the seeded 2k-80k stream and on each twist's minimal cell, then latency
    under Poisson arrivals at a stated rate."""
    from lmff import serving

    steps = args.steps
    jobs = serving.tbg_job_stream(args.jobs, rate_hz=1e9, steps=steps, seed=2024)
    sizes = np.array([j.natoms for j in jobs])
    small = serving.tbg_job_stream(args.jobs, rate_hz=1e9, steps=steps, seed=2025, sizes="minimal")
    small_sizes = np.array([j.natoms for j in small])
    log(f"capacity: {len(jobs)} jobs packed, then one at a time")
    packed, seq, direct = stream_capacity(jobs)
    log(f"capacity: {len(small)} minimal-cell jobs")
    packed_s, seq_s, direct_s = stream_capacity(small)
    rate = args.rate or 0.5 * packed["jobs"] / packed["makespan_s"]
    log(f"latency run at {rate:.2f} jobs/s")
    lat = stream_latency(jobs, rate)
    return {
        "metric": "job stream atom-steps/s packed vs sequential; job p50/p99 latency",
        "value": packed["atom_steps_per_s"], "unit": "atom-steps/s", "n_gpus": 1, "steps": steps,
        "higher_is_better": True, "dtype": "mixed (FP32 ILP pair math, FP64 elsewhere)",
        "data": "synthetic twisted-bilayer stream (seeded)",
        "config": {"workload": f"{len(jobs)} twisted-bilayer graphene jobs, (m,n) with 1<=n<m<=12, "
                               f"{sizes.min()}-{sizes.max()} atoms (median {int(np.median(sizes))}), "
                               f"{steps} Langevin steps each, 300 K, damp 1 ps, dt 1 fs, tick 10"},
        "capacity": _capacity_dict(packed, seq, direct),
        "capacity_minimal_cells": dict(
            _capacity_dict(packed_s, seq_s, direct_s),
            workload=f"{len(small)} jobs, each twist's smallest cell with >= 2000 atoms and sides >= 36.5 A "
                     f"({small_sizes.min()}-{small_sizes.max()} atoms, median {int(np.median(small_sizes))})"),
        "latency": {"arrival_rate_jobs_per_s": rate, "p50_s": lat["latency_p50_s"], "p99_s": lat["latency_p99_s"],
                    "atom_steps_per_s": lat["atom_steps_per_s"], "max_resident": lat["max_resident"]},
    }


if __name__ == "__main__":
    main()
