"""Benchmark: SIMPLE cell-iterations/s of the prognostic flow solver on one MI355X.

Workload: the plain solver on the seeded random city block (400 x 400 cells,
dx = 2 m, 40 prognostic layers, cuboids at 30 % plan-area density, log inflow,
wind from 270 deg).  A step is one SIMPLE iteration of that situation with the
state resident in HBM (no host transfers in the timed region).

Prints one JSON line (see the repository's bench contract).  ``--gpus N`` with
N > 1 exits: this build targets one GPU.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# algorithmic FP32 bytes per active cell for one pass of each kernel family:
# every array the pass reads or writes counted once (pressure: DIV, DPM, DPMNEW
# read; DPM, DPMNEW written; momentum (mixing length): 7 arrays read, 1 written)
ALG_BYTES = {"pressure": 5 * 4, "u": 8 * 4, "v": 8 * 4, "w": 8 * 4}
ALG_BYTES_ITER = 188          # whole SIMPLE iteration, mixing-length model
HBM_PEAK_GBS = 8000.0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=400, help="grid cells per side (city block)")
    ap.add_argument("--direction", type=float, default=270.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-iterations", type=int, default=2)
    ap.add_argument("--no-family-profile", action="store_true",
                    help="skip the per-launch event timing (for counter-collection runs)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB five times (counter calibration) and exit")
    args = ap.parse_args()
    if args.calibrate_copy:
        import torch
        n = args.calibrate_copy * (1 << 20) // 4
        a = torch.ones(n, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"calibrate_copy_bytes_read": n * 4, "calibrate_copy_bytes_written": n * 4, "launches": 5}))
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this solver runs on one GPU; multi-GPU runs are not part of this build",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    from gral_amd import domain, pf

    case = domain.config2(direction=args.direction, iterations=args.warmup + args.steps + 16, n=args.n)
    active = case.active_cells()
    solver = pf.PrognosticFlowfield(case)
    solver.prepare(iterations=args.warmup + args.steps + 16)
    if args.warmup:
        solver.iterate(args.warmup)
    t0 = time.perf_counter()
    solver.iterate(args.steps)
    wall = time.perf_counter() - t0
    tm = solver.timing()
    ms = tm["ms_total"] / args.steps
    value = active * args.steps / (tm["ms_total"] / 1e3)

    # dominant kernel: per-launch device time of each sweep family over a few
    # instrumented iterations (events bracket single launches on the solver's stream)
    if args.no_family_profile:
        solver.close()
        print(json.dumps({"metric": "SIMPLE cell-iters/s", "value": value, "ms_per_step": ms, "steps": args.steps}))
        return
    prof = solver.profile_families(4)
    dom = max(prof, key=lambda k: prof[k]["ms"])
    pdom = prof[dom]
    avg_ms = pdom["ms"] / max(1, pdom["launches"])
    bytes_per_launch = ALG_BYTES.get(dom, 0) * pdom["cells"] / max(1, pdom["launches"])
    achieved = bytes_per_launch / (avg_ms / 1e3) / 1e9 if avg_ms > 0 else 0.0
    roofline = {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                "avg_launch_us": avg_ms * 1e3, "launches_per_iter": pdom["launches"] / 4,
                "alg_bytes_per_cell": ALG_BYTES.get(dom, 0)}
    solver.close()

    cpu = None
    if not args.no_cpu_baseline:
        from oracle import pyoracle
        c0 = case.copy()
        t = time.perf_counter(); pyoracle.run(c0, iterations=0, keep_history=False); tp = time.perf_counter() - t
        t = time.perf_counter(); pyoracle.run(c0, iterations=args.cpu_iterations, keep_history=False)
        tc = time.perf_counter() - t - tp
        cpu = {"value": active * args.cpu_iterations / tc, "unit": "cell-iter/s", "cores": 1, "kind": "port",
               "sample": f"{args.cpu_iterations} SIMPLE iterations of the same situation, deterministic "
                         f"single-thread C restatement, prologue excluded ({tc:.1f} s)"}

    line = {"metric": "SIMPLE cell-iters/s", "value": value, "unit": "cell-iter/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
            "config": {"workload": f"city block {case.nii}x{case.njj}, dx=2 m, 40 prognostic layers, "
                                   f"30% buildings (seed 20260001), wind {args.direction:g} deg, plain solver, "
                                   f"mixing-length model", "active_cells": active, "kadvmax": case.kadvmax,
                       "parallelism": "single GPU"},
            "wall_s_timed": wall, "alg_bytes_per_cell_iter": ALG_BYTES_ITER,
            "iteration_frac_of_hbm_roofline": value * ALG_BYTES_ITER / (HBM_PEAK_GBS * 1e9),
            "roofline": roofline, "cpu_baseline": cpu,
            "kernel_families_ms_per_iter": {k: v["ms"] / 4 for k, v in prof.items()}}
    print(json.dumps(line))


if __name__ == "__main__":
    main()
