"""Throughput of the LMFF force path (Tersoff + ILP with normals and Tap) on one MI355X.

Workload: graphene/hBN heterobilayer, 220x130 rectangular cells per layer (228,800 atoms),
seeded 0.02 A jitter, 300 K Maxwell-Boltzmann velocities, NVE dt = 1 fs, neighbor rebuild
every 10 steps with a 2 A skin.  A "step" is one Velocity-Verlet step of the whole system
(force evaluation + integration + the amortised neighbor rebuild), with all state resident
in HBM.  The FP64 path gives `value`; the mixed-precision path (FP32 pair/bond-order math,
FP64 displacements and accumulation) is timed on the same input and reported beside it,
with its force error against FP64.

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


def build_ctx(s, precision):
    ctx = capi.Context(0)
    capi.setup_for(ctx, s, precision=precision, skin=2.0, every=10, delay=0, check=0)
    sid = ctx.add_system(s)
    ctx.integrator(sid, capi.LMFF_NVE, dt=0.001)
    return ctx, sid


def one_shot(s, precision):
    ctx = capi.Context(0)
    try:
        capi.setup_for(ctx, s, precision=precision)
        return ctx.pair_compute(s, 1, 2)
    finally:
        ctx.close()


def timed_run(ctx, sid, s, steps, warmup, profile):
    ctx.evaluate(1, 2)
    th0 = ctx.thermo(sid)
    if warmup:
        ctx.run(warmup, 0, use_graph=True)
    ctx.profile(profile)
    t0 = time.perf_counter()
    rec = ctx.run(steps, steps, use_graph=not profile)
    wall = time.perf_counter() - t0
    ms_total, fam_ms, fam_n = ctx.profile_read()
    ctx.profile(False)
    th1 = rec[0][-1]
    return {"wall_s": wall, "gpu_ms": ms_total, "fam_ms": fam_ms.tolist(), "fam_n": fam_n.tolist(),
            "e0": th0["pe"] + th0["ke"], "e1": th1["pe"] + th1["ke"], "temp_end": th1["temp"]}


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
    while t_force < seconds_budget or steps == 0:
        t1 = time.perf_counter()
        r = o.compute(1, 2)
        t_force += time.perf_counter() - t1
        steps += 1
        if f_ref is None:
            f_ref = r
        if steps >= 3:
            break
    per_step = t_force / steps + t_list / 10.0  # list rebuild amortised over 10 steps
    return {"value": s.n / per_step, "unit": "atom-steps/s", "cores": cores, "kind": "port",
            "sample": f"{steps} force evaluations + 1 list build of the same {s.n}-atom system "
                      f"(oracle/lmff_oracle.c, OpenMP), integration excluded",
            "cpu_model": _cpu_model()}, f_ref


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--nx", type=int, default=220)
    ap.add_argument("--ny", type=int, default=130)
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

    s = S.layered_stack(args.nx, args.ny, ["G", "BN"], seed=12346)
    n = s.n

    # ---- single force evaluations of the initial state (for the error figures)
    r64 = one_shot(s, capi.LMFF_FP64)
    rmx = one_shot(s, capi.LMFF_MIXED)

    # ---- FP64: the headline number
    ctx, sid = build_ctx(s, capi.LMFF_FP64)
    stats = ctx.pair_stats()
    res64 = timed_run(ctx, sid, s, args.steps, args.warmup, profile=True)
    ctx.close()

    # ---- mixed precision on the same input
    ctxm, sidm = build_ctx(s, capi.LMFF_MIXED)
    resmx = timed_run(ctxm, sidm, s, args.steps, args.warmup, profile=True)
    ctxm.close()

    df = rmx["f"] - r64["f"]
    mixed_err = {"max_abs": float(np.abs(df).max()), "rms": float(np.sqrt((df ** 2).sum(1).mean()))}

    ms64 = res64["gpu_ms"] / args.steps
    msmx = resmx["gpu_ms"] / args.steps
    value = n / (ms64 * 1e-3)

    # roofline of the dominant kernel (ILP pair): algorithmic FLOP per launch / mean launch time
    ilp_ms = res64["fam_ms"][2]
    flop = stats["ilp_pairs"] * ILP_PAIR_FLOP
    achieved = flop / (ilp_ms * 1e-3) / 1e12
    # algorithmic bytes per launch: list stream + x/type of i and j-gathers once per atom + outputs
    traffic_alg = 4.0 * stats["ilp_entries"] + n * (32 + 32 + 32 + 32)
    ilp_ms_mx = resmx["fam_ms"][2]

    cpu = None
    oracle_err = None
    if not args.no_cpu:
        cpu, fref = cpu_baseline(s)
        d = r64["f"] - fref["f"]
        rmsf = np.sqrt((fref["f"] ** 2).sum(1).mean())
        oracle_err = {"fp64_max_abs": float(np.abs(d).max()), "fp64_rel_max_over_rms": float(np.abs(d).max() / rmsf),
                      "fp64_energy_rel": float(abs(r64["energy"] - fref["energy"]) / abs(fref["energy"])),
                      "mixed_max_abs": float(np.abs(rmx["f"] - fref["f"]).max()),
                      "mixed_rms": float(np.sqrt(((rmx["f"] - fref["f"]) ** 2).sum(1).mean()))}

    out = {
        "metric": "atom-steps/s FP64 vs mixed at RMS force err<=1e-5 eV/A; job p50/p99 latency",
        "value": value,
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
        "config": {"workload": f"graphene/hBN heterobilayer {args.nx}x{args.ny} cells/layer, {n} atoms, "
                               f"NVE dt=1 fs, rebuild every 10 steps, skin 2 A, ILP 16 A with Tap",
                   "natoms": n, "precision": "fp64"},
        "mixed": {"value": n / (msmx * 1e-3), "ms_per_step": msmx, "speedup_vs_fp64": ms64 / msmx,
                  "force_err_vs_fp64": mixed_err, "ilp_kernel_ms": ilp_ms_mx},
        "drift_eV_per_atom": {"fp64": (res64["e1"] - res64["e0"]) / n, "mixed": (resmx["e1"] - resmx["e0"]) / n,
                              "steps": args.steps},
        "kernel_ms_per_launch": dict(zip(["list_build", "normals", "ilp_pair", "tersoff_zeta", "gather_tersoff",
                                          "integrate", "reduce"], res64["fam_ms"])),
        "roofline": {"bound": "valu", "kernel": "k_ilp<double>", "achieved": achieved, "peak": FP64_VALU_PEAK,
                     "unit": "TFLOP/s", "frac": achieved / FP64_VALU_PEAK, "traffic": None,
                     "algorithmic_bytes_per_launch": traffic_alg,
                     "flop_per_launch": flop, "ilp_pairs": stats["ilp_pairs"],
                     "ilp_list_entries": stats["ilp_entries"]},
        "cpu_baseline": cpu,
        "oracle_force_error": oracle_err,
        "wall_s": {"fp64": res64["wall_s"], "mixed": resmx["wall_s"]},
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
