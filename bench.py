"""Benchmark: SIMPLE cell-iterations/s of the prognostic flow solver on one MI355X.

Workload: the seeded random city block (400 x 400 cells, dx = 2 m, 40
prognostic layers, cuboids at 30 % plan-area density, log inflow, z0 = 0.5 m)
under its 8 wind directions (0, 45, ..., 315 deg), the 8 situations advanced
together as one lock-step batch.  Each situation is the plain solver (same
order, same arithmetic, bit-identical to a solve on its own).  A step is one
SIMPLE iteration of all situations with the state resident in HBM (no host
transfers in the timed region).  The single-direction (270 deg) plain solve is
timed beside it.

Prints one JSON line.  ``--gpus N`` with N > 1 exits: the solver runs on one
GPU (situations would be dealt to one process per GPU, with no communication).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# algorithmic FP32 bytes per active cell for one pass of each sweep family:
# every array the pass reads or writes counted once (pressure: DIV, DPM, DPMNEW
# read; DPM, DPMNEW written; momentum (mixing length): 7 arrays read, 1 written;
# k-eps momentum: 7 read (TURB instead of VEG), 1 written; TKE: 5 read, 2 written)
ALG_BYTES = {"pressure": 5 * 4, "u": 8 * 4, "v": 8 * 4, "w": 8 * 4, "tke": 7 * 4}
KERNEL_OF = {"pressure": "k_pressure_strip", "u": "k_ml_strip<0, 0>", "v": "k_ml_strip<1, 0>", "w": "k_ml_strip<2, 0>"}
KERNEL_OF_KEPS = {"pressure": "k_pressure_strip", "u": "k_ml_strip<0, 1>", "v": "k_ml_strip<1, 1>",
                  "w": "k_ml_strip<2, 1>", "tke": "k_tke_strip"}
ALG_BYTES_ITER = 188          # whole SIMPLE iteration, mixing-length model
ALG_BYTES_ITER_KEPS = 212     # k-eps model (53 array accesses)
HBM_PEAK_GBS = 8000.0


def particle_wind(n, cells):
    """The particle model's wind at n particles spread uniformly over the
    solved city field (terrain branch, every height of the column), five calls
    through the C ABI.  The rate includes the host-to-device and back copies of
    the particles (48 B each way); the kernel alone is timed by rocprofv3."""
    from gral_amd import domain, pf
    c = domain.config2(direction=270.0, iterations=4, n=cells)
    r = np.random.default_rng(7)
    ikoo, jkoo, ahmin = 600000, 5200000, 100.0
    ci = r.integers(1, c.nii + 1, n).astype(np.int32)
    cj = r.integers(1, c.njj + 1, n).astype(np.int32)
    x = ikoo + (ci - 1 + r.uniform(0, 1, n)) * c.dxk
    y = jkoo + (cj - 1 + r.uniform(0, 1, n)) * c.dyk
    z = (r.uniform(0.0, float(c.hokart[c.nkk]), n) + ahmin).astype(np.float32)
    ah = np.zeros(n, np.float32)
    with pf.PrognosticFlowfield(c) as sv:
        sv.prepare(iterations=4)
        sv.iterate(2)
        sv.wind_at_particles(1, ikoo, jkoo, ahmin, ci, cj, ah, x, y, z)
        t0 = time.perf_counter()
        for _ in range(5):
            sv.wind_at_particles(1, ikoo, jkoo, ahmin, ci, cj, ah, x, y, z)
        wall = (time.perf_counter() - t0) / 5
    return {"metric": "particle wind evaluations/s (with host transfers)", "value": n / wall, "particles": n,
            "ms_per_call": wall * 1e3, "workload": f"city block {cells}x{cells}, 270 deg, fields after 2 iterations, "
            "particles uniform over cells and heights", "alg_bytes_per_particle": 24 + 32 + 16}


def situations_per_hour(n, directions):
    """Flow situations per hour, batched against plain, at equal residual.

    Every situation is solved to the reference's convergence test with its
    default limits (100 to 500 iterations, stop test every 100), once as one
    lock-step batch and once one after another.  The batch is bit-identical
    to the plain solves, so both stop at the same iterations with the same
    residuals (checked).  Wall time covers what a caller waits for: handle
    creation and upload, the host prologue, the solve and the download of the
    fields."""
    from gral_amd import domain, pf

    def cases():
        cs = [domain.config2(direction=d, n=n) for d in directions]
        for c in cs:
            c.iter_min, c.iter_max = 100, 500
        return cs

    cs = cases()
    t0 = time.perf_counter()
    solvers = [pf.PrognosticFlowfield(c) for c in cs]
    for sv in solvers:
        sv.prepare()
    with pf.Batch(solvers) as b:
        st_b = b.solve()
    for sv in solvers:
        sv.download()
        sv.close()
    t_batch = time.perf_counter() - t0

    cs = cases()
    t0 = time.perf_counter()
    st_p = []
    for c in cs:
        with pf.PrognosticFlowfield(c) as sv:
            st_p.append(sv.solve())
            sv.download()
    t_plain = time.perf_counter() - t0
    its_b = [s["iteration_loops"] for s in st_b]
    its_p = [s["iteration_loops"] for s in st_p]
    same = its_b == its_p and [s["delta_finish"] for s in st_b] == [s["delta_finish"] for s in st_p]
    return {"situations": len(cs), "batched_per_hour": len(cs) / t_batch * 3600.0,
            "plain_per_hour": len(cs) / t_plain * 3600.0, "batched_over_plain": t_plain / t_batch,
            "batched_s": t_batch, "plain_s": t_plain, "iterations": its_b, "equal_residual": same,
            "limits": "IterationStepsMin 100, IterationStepsMax 500, stop test every 100 iterations"}


def gff_timing(sv, reps=5):
    """Flow-field file writer on a prepared situation whose fields are on the
    device: payload packed on the GPU and copied to the host (per kind), the
    zip file written (writing mode 2, parallel deflate), and beside it the
    oracle's numpy payload plus single-thread zlib deflate of the same data."""
    import tempfile
    import zipfile
    meta = dict(wind_dir=270.0, wind_vel=3.0, stab_class=4)
    out = {"kinds": {}}
    sv.gff_payload(2, **meta)                 # device copy of the upper levels, once per situation
    buf = np.empty(max(sv.gff_size(k) for k in (0, 1, 2)), dtype=np.uint8)
    buf[:] = 0                                # pages touched before timing
    for kind in (0, 1, 2):
        ts = []
        for _ in range(reps):
            t = time.perf_counter()
            n = sv.gff_payload_into(kind, buf, **meta)
            ts.append(time.perf_counter() - t)
        ms = sorted(ts)[reps // 2] * 1e3
        out["kinds"][str(kind)] = {"payload_bytes": n, "pack_and_copy_ms": ms, "gb_per_s": n / ms / 1e6}
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "00001.gff")
        ts = []
        for _ in range(3):
            t = time.perf_counter()
            sv.write_gff(p, 2, **meta)
            ts.append(time.perf_counter() - t)
        out["write_mode2_ms"] = sorted(ts)[1] * 1e3
        out["zip_bytes_mode2"] = os.path.getsize(p)
        with zipfile.ZipFile(p) as z:
            assert z.read("00001.gff") == sv.gff_payload(2, **meta)
    from oracle import gff_oracle
    f = sv.download()
    c = sv.case
    t = time.perf_counter()
    ref = gff_oracle.payload(2, f["uk"], f["vk"], f["wk"], c.kkart, c.nii, c.njj, c.nkk, c.dxk, **meta)
    t_pack = time.perf_counter() - t
    t = time.perf_counter()
    zlib_len = len(__import__("zlib").compress(ref, 6))
    t_zip = time.perf_counter() - t
    out["cpu_oracle"] = {"pack_ms": t_pack * 1e3, "deflate_ms": t_zip * 1e3, "deflated_bytes": zlib_len, "cores": 1,
                         "sample": "writing mode 2 of the same situation: numpy payload oracle, zlib level 6"}
    out["equal_to_oracle"] = ref == sv.gff_payload(2, **meta)
    return out


def batched_workload(k, reps_log=sys.stderr):
    """The batched workload generator's first k situations (directions in 10 deg
    steps, speeds 0.5-6 m/s, stability classes 1-7, windows of 200-600 cells,
    30-50 prognostic layers), each set up on the GPU and solved to the stop
    test (100-500 iterations): once as one lock-step batch, once one after
    another with the plain solver.  Wall time covers the set-up, handle
    creation, prologue, solve and download."""
    from gral_amd import domain, pf
    sits = domain.config3_situations()[:k]

    def log(msg):
        print(f"[batched-workload] {msg}", file=reps_log, flush=True)

    t0 = time.perf_counter()
    solvers = []
    for s in sits:
        c = domain.config3(s)
        sv = pf.PrognosticFlowfield(c)
        sv.prepare()
        solvers.append(sv)
    t_setup = time.perf_counter() - t0
    log(f"batch set up ({t_setup:.1f} s), solving {len(solvers)} situations in lock step")
    with pf.Batch(solvers) as b:
        st_b = b.solve()
        tb = b.timing()
    for sv in solvers:
        sv.download()
        sv.close()
    t_batch = time.perf_counter() - t0
    log(f"batched {t_batch:.1f} s")
    st_p = []
    t0 = time.perf_counter()
    for q, s in enumerate(sits):
        c = domain.config3(s)
        with pf.PrognosticFlowfield(c) as sv:
            st_p.append(sv.solve())
            sv.download()
        log(f"plain situation {q + 1}/{len(sits)} ({time.perf_counter() - t0:.1f} s)")
    t_plain = time.perf_counter() - t0
    its_b = [x["iteration_loops"] for x in st_b]
    same = its_b == [x["iteration_loops"] for x in st_p] and \
        [x["delta_finish"] for x in st_b] == [x["delta_finish"] for x in st_p]
    return {"situations": len(sits), "sizes": [s["n"] for s in sits], "layers": [s["vertcellsff"] for s in sits],
            "batched_per_hour": len(sits) / t_batch * 3600.0, "plain_per_hour": len(sits) / t_plain * 3600.0,
            "batched_over_plain": t_plain / t_batch, "batched_s": t_batch, "plain_s": t_plain,
            "batched_setup_s": t_setup, "batched_device_ms": tb["ms_total"], "iterations": its_b,
            "equal_residual": same}


def diagnostic_timing(n, reps=3):
    """The mass-conservation post-pass (DiagnosticFlowfield: 10 pressure
    iterations over the full column height, then the W mass step) on the city
    block after 2 prognostic iterations, wall time of gral_pf_diagnostic."""
    from gral_amd import domain, pf
    c = domain.config2(direction=270.0, iterations=4, n=n)
    ts = []
    with pf.PrognosticFlowfield(c) as sv:
        sv.prepare(iterations=4)
        sv.iterate(2)
        sv.diagnostic(flow_field_level=2)               # warm-up (device buffers, host copies)
        for _ in range(reps):
            t = time.perf_counter()
            sv.diagnostic(flow_field_level=2)
            ts.append(time.perf_counter() - t)
    return {"ms": sorted(ts)[reps // 2] * 1e3, "reps": reps,
            "workload": f"city block {n}x{n}, 270 deg, after 2 iterations, flow_field_level 2 (10 pressure iterations)"}


def adjoint_timing(steps=20, reps=3):
    """The adjoint of the pressure-correction steps on the adjoint configuration's
    grid (256 x 256 city, 32 layers): device time of a forward window of
    ``steps`` steps (gral_pf_forward_ckpt) and of its VJP (gral_pf_vjp, the
    gradient of the mean wind speed at 64 receptors 3 m above ground), beside
    a full SIMPLE iteration of the same situation."""
    from gral_amd import adjoint, domain, pf
    c = domain.config4(iterations=steps + 8)
    rec = adjoint.receptors(c)
    fw, vj = [], []
    with pf.PrognosticFlowfield(c) as sv:
        sv.prepare(iterations=steps + 8)
        sv.iterate(4)
        it_ms = sv.timing()["ms_total"] / 4
        for _ in range(reps):
            sv.forward_ckpt(steps)
            fw.append(sv.timing()["ms_total"])
            du, dv, dw = adjoint.mean_receptor_speed_grad(sv.download(), rec)
            sv.vjp(du, dv, dw)
            vj.append(sv.timing()["ms_total"])
            sv.restore_ckpt()
        # one mixing-length momentum sweep per component (gral_pf_momentum_sweep,
        # per-diagonal launches) and its VJP (gral_pf_momentum_vjp) from the same state
        mom = {}
        rng = np.random.default_rng(5)
        gbar = rng.standard_normal(c.uk.shape).astype(np.float32)
        base = sv.download()
        for comp, name in enumerate("UVW"):
            sw, mv = [], []
            for _ in range(reps):
                sv.set_velocity(base["uk"], base["vk"], base["wk"])
                sv.momentum_sweep(comp)
                sw.append(sv.timing()["ms_total"])
                sv.momentum_vjp(gbar)
                mv.append(sv.timing()["ms_total"])
            a, b = sorted(sw)[reps // 2], sorted(mv)[reps // 2]
            mom[name] = {"sweep_ms": a, "vjp_ms": b, "vjp_over_sweep": b / a}
        sv.set_velocity(base["uk"], base["vk"], base["wk"])
    f, v = sorted(fw)[reps // 2], sorted(vj)[reps // 2]
    return {"workload": f"city 256x256, 32 layers, {steps} pressure-correction steps, VJP of the mean speed at "
                        f"{len(rec)} receptors", "active_cells": c.active_cells(), "forward_ms": f, "vjp_ms": v,
            "vjp_over_forward": v / f, "simple_iteration_ms": it_ms, "vjp_step_over_simple_iteration": v / steps / it_ms,
            "momentum": mom}


def graph_timing(iters=100):
    """Launch replay on the small plain case (100 x 100, one building): device
    time per SIMPLE iteration with every kernel launched, and with the
    iteration chunks replayed as captured HIP graphs (GRAL_PF_GRAPH=1)."""
    from gral_amd import domain, pf
    out = {}
    for mode in ("0", "1"):
        os.environ["GRAL_PF_GRAPH"] = mode
        c = domain.config1(iterations=iters + 20)
        with pf.PrognosticFlowfield(c) as sv:
            sv.prepare(iterations=iters + 20)
            sv.iterate(20)                               # the first chunk is captured
            sv.iterate(iters)
            out["graphs" if mode == "1" else "launches"] = sv.timing()["ms_total"] / iters
    os.environ.pop("GRAL_PF_GRAPH", None)
    return {"workload": "config 1 (100x100, 30 layers, one building), plain", "ms_per_iter_launches": out["launches"],
            "ms_per_iter_graphs": out["graphs"], "speedup": out["launches"] / out["graphs"]}


def keps_timing(n, directions, steps=5, warmup=12):
    """The same 8-direction city batch with the k-eps model (turbulence model
    2: k-eps momentum and the TKE/dissipation sweep), timed the same way as
    the main line: device time per SIMPLE iteration over ``steps`` iterations
    after ``warmup``, and the per-family launch times of two instrumented
    iterations with the dominant sweep's share of the HBM peak."""
    from gral_amd import domain, pf
    total = warmup + steps + 8
    cases = [domain.config2(direction=d, iterations=total, n=n, turbulence_model=2) for d in directions]
    solvers = [pf.PrognosticFlowfield(c) for c in cases]
    for sv in solvers:
        sv.prepare(iterations=total)
    with pf.Batch(solvers) as batch:
        if warmup:
            batch.iterate(warmup)
        batch.iterate(steps)
        tm = batch.timing()
        prof = batch.profile_families(2)
    for sv in solvers:
        sv.close()
    sweeps = {k: v for k, v in prof.items() if k in ALG_BYTES}
    dom = max(sweeps, key=lambda k: sweeps[k]["ms"])
    pdom = sweeps[dom]
    avg_ms = pdom["ms"] / max(1, pdom["launches"])
    achieved = ALG_BYTES[dom] * pdom["cells"] / max(1, pdom["launches"]) / (avg_ms / 1e3) / 1e9 if avg_ms > 0 else 0.0
    return {"workload": f"city block {n}x{n}, the same {len(cases)} directions as one lock-step batch, k-eps model",
            "ms_per_step": tm["ms_total"] / steps, "value": tm["cell_iterations"] / (tm["ms_total"] / 1e3),
            "unit": "cell-iter/s", "alg_bytes_per_cell_iter": ALG_BYTES_ITER_KEPS,
            "dominant_kernel": KERNEL_OF_KEPS[dom], "dominant_avg_launch_us": avg_ms * 1e3,
            "dominant_frac_of_hbm_peak": achieved / HBM_PEAK_GBS,
            "kernel_families_ms_per_iter": {k: v["ms"] / 2 for k, v in prof.items()}}


def convergence_study(log=sys.stderr):
    """The reference's stop test on city situations: delta = (sum of DeltaUMax
    over 100 iterations) / topwind^2 at each test (iterations 100..500;
    PrognosticFlowfield.cs:1149-1166, stop below 0.012) for three situations of
    the batched workload's generator on a 200 x 200 window and for cities of
    2 %, 10 % and 30 % plan-area density (buildings 6-20 m) at 1, 3 and 6 m/s."""
    from gral_amd import domain, pf
    F32 = np.float32

    def deltas(c):
        c.iter_min, c.iter_max = 100, 500
        ds = []
        with pf.PrognosticFlowfield(c) as sv:
            sv.prepare()
            for _ in range(5):
                st = sv.iterate(100, check_convergence=True)
                ds.append(round(st["last_delta"], 4))
                if st["iteration_loops"] % 100 != 1 or st["iteration_loops"] > 500:
                    break
        return {"iteration_loops": st["iteration_loops"], "delta_at_tests": ds}

    out = []
    for s in domain.config3_situations()[:3]:
        s = dict(s, n=200)
        row = {"case": f"batched workload situation {s['index']} on a 200x200 window", "speed": s["speed"],
               "stability": s["stability"], "vertcellsff": s["vertcellsff"], **deltas(domain.config3(s))}
        print(f"[convergence] {row}", file=log, flush=True)
        out.append(row)
    for density in (0.02, 0.1, 0.3):
        cut = np.zeros((202, 202), dtype=F32)
        for b in domain.city_buildings(200, 200, 2.0, 77, density=density, hmin=6.0, hmax=20.0):
            sl = (slice(b.i0, b.i1 + 1), slice(b.j0, b.j1 + 1))
            cut[sl] = np.maximum(cut[sl], F32(b.height))
        for speed in (1.0, 3.0, 6.0):
            c = domain.flat_case(f"d{density}", 200, 200, cut, vertcellsff=40, speed=speed, direction=270.0)
            row = {"case": f"city 200x200, {int(100 * density)} % density, buildings 6-20 m", "speed": speed,
                   **deltas(c)}
            print(f"[convergence] {row}", file=log, flush=True)
            out.append(row)
    return out


def batch_sweep(sizes, n, steps=6, warmup=12, log=sys.stderr):
    """Throughput against batch size on the city block: S situations (the 8
    directions repeated) advanced in lock step; per size the device time per
    iteration, cell-iterations/s and the W sweep's share of the HBM roofline
    (algorithmic bytes over its event-timed launch)."""
    from gral_amd import domain, pf
    out = []
    for S in sizes:
        cases = [domain.config2(direction=45.0 * (q % 8), iterations=warmup + steps + 4, n=n) for q in range(S)]
        solvers = [pf.PrognosticFlowfield(c) for c in cases]
        for sv in solvers:
            sv.prepare(iterations=warmup + steps + 4)
        with pf.Batch(solvers) as b:
            b.iterate(warmup)
            b.iterate(steps)
            tm = b.timing()
            prof = b.profile_families(1)
        for sv in solvers:
            sv.close()
        w = prof["w"]
        wfrac = ALG_BYTES["w"] * w["cells"] / (w["ms"] / w["launches"] / 1e3) / 1e9 / HBM_PEAK_GBS / w["launches"]
        row = {"situations": S, "ms_per_iter": tm["ms_total"] / steps,
               "cell_iter_per_s": tm["cell_iterations"] / (tm["ms_total"] / 1e3), "w_sweep_frac": wfrac,
               "w_sweep_ms": w["ms"] / w["launches"], "stream_groups": tm["groups"]}
        print(f"[batch-sweep] {row}", file=log, flush=True)
        out.append(row)
    return out


def stream_workload(k, capacity, plain, log=sys.stderr):
    """The first k situations of the batched workload generator, each set up on
    the GPU and solved to the reference's stop test (100-500 iterations), with
    one batch fed while it runs (at most ``capacity`` situations in the batch,
    the next ones set up on host threads meanwhile, largest first), against
    ``plain`` of them solved one after another with the plain solver."""
    from gral_amd import domain, pf
    sits = domain.config3_situations()[:k]
    cells = {}
    mass = []

    def make(s):
        def f():
            c = domain.config3(s)
            cells[c.name] = c.active_cells()
            return c
        return f

    done = [0]
    ci = [0]
    groups = {}

    def on_step(tm):
        groups[tm["groups"]] = groups.get(tm["groups"], 0) + 1

    def on_done(case, sv, st):
        done[0] += 1
        ci[0] += cells[case.name] * st["iterations_run"]
        mass.append(st["mass_div"] / st["mass_flux"] if st["mass_flux"] > 0 else 0.0)
        if done[0] % 16 == 0:
            print(f"[stream] {done[0]}/{k} situations done ({time.perf_counter() - t0:.0f} s)", file=log, flush=True)

    t0 = time.perf_counter()
    order = sorted(sits, key=lambda s: -s["n"] * s["n"] * s["vertcellsff"])      # longest first
    stats = pf.solve_stream([make(s) for s in order], capacity=capacity, on_done=on_done, on_step=on_step)
    t_stream = time.perf_counter() - t0
    loops = [st["iteration_loops"] for st in stats]
    runs = [st["iterations_run"] for st in stats]
    tot_cells = sum(cells.values())
    out = {"situations": k, "capacity": capacity, "wall_s": t_stream, "per_hour": k / t_stream * 3600.0,
           "stopped_before_limit": int(sum(1 for x in loops if x < 500)), "iteration_loops_min": min(loops),
           "iteration_loops_max": max(loops), "iterations_run_total": int(sum(runs)),
           "active_cells_total": int(tot_cells), "cell_iterations": int(ci[0]),
           "cell_iter_per_s_wall": ci[0] / t_stream,
           "mass_imbalance_median": float(np.median(mass)) if mass else None,
           "steps_by_stream_groups": {str(g): n for g, n in sorted(groups.items())}}
    if plain:
        # plain solves of a sample, one after another (set-up, prologue, solve,
        # download); extrapolated to the whole set by cell-iterations
        t1 = time.perf_counter()
        ci_p = 0
        for q, s in enumerate(sits[:plain]):
            c = domain.config3(s)
            with pf.PrognosticFlowfield(c) as sv:
                st = sv.solve()
                sv.download()
            ci_p += c.active_cells() * st["iterations_run"]
            print(f"[stream] plain {q + 1}/{plain} ({time.perf_counter() - t1:.1f} s)", file=log, flush=True)
        t_plain = time.perf_counter() - t1
        t_plain_all = t_plain / ci_p * ci[0]
        out["plain_sample"] = plain
        out["plain_sample_s"] = t_plain
        out["plain_cell_iter_per_s_wall"] = ci_p / t_plain
        out["plain_per_hour"] = k / t_plain_all * 3600.0
        out["batched_over_plain"] = t_plain_all / t_stream
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=12,
                    help="untimed iterations first (a batch times its stream-group counts in the first 12)")
    ap.add_argument("--n", type=int, default=400, help="grid cells per side (city block)")
    ap.add_argument("--directions", type=int, default=8, help="wind directions 0, 45, ... solved as one batch")
    ap.add_argument("--no-plain", action="store_true", help="skip the single-situation plain solve")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-iterations", type=int, default=2)
    ap.add_argument("--no-family-profile", action="store_true",
                    help="skip the per-launch event timing (for counter-collection runs)")
    ap.add_argument("--situations-hour", action="store_true",
                    help="also solve every situation to convergence, batched and plain (about 1-2 minutes)")
    ap.add_argument("--gff", action="store_true",
                    help="also time the flow-field file writer (GPU pack + zip) on the first situation")
    ap.add_argument("--batched-workload", type=int, default=0, metavar="K",
                    help="also solve the first K situations of the batched workload, batched and plain")
    ap.add_argument("--graph", action="store_true",
                    help="also time the small plain case with launches against replayed HIP graphs")
    ap.add_argument("--adjoint", action="store_true",
                    help="also time the adjoint of the pressure-correction steps (forward window and VJP)")
    ap.add_argument("--diagnostic", action="store_true",
                    help="also time the diagnostic post-pass (gral_pf_diagnostic) on the 270 deg city")
    ap.add_argument("--convergence-study", action="store_true",
                    help="also report the reference's stop-test delta every 100 iterations on city situations")
    ap.add_argument("--batch-sweep", type=str, default="", metavar="S1,S2,...",
                    help="also time lock-step batches of these sizes on the city block")
    ap.add_argument("--stream", type=int, default=0, metavar="K",
                    help="also solve the first K situations of the batched workload with one batch fed while it "
                         "runs (see --capacity), to the stop test")
    ap.add_argument("--capacity", type=int, default=32, help="situations on the device at a time in --stream")
    ap.add_argument("--stream-plain", type=int, default=4, metavar="P",
                    help="plain solves of the first P situations beside --stream")
    ap.add_argument("--no-keps", action="store_true",
                    help="skip the k-eps leg (the same batch with turbulence model 2)")
    ap.add_argument("--turbulence-model", type=int, default=1, choices=(1, 2),
                    help="1: mixing length (the benchmark's model), 2: k-eps")
    ap.add_argument("--particle-wind", type=int, default=0, metavar="N",
                    help="only evaluate the particle model's wind (gral_pf_wind_at_particles) at N particles on the "
                         "solved 270 deg city field, 5 calls, and exit")
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
        print(json.dumps({"error": "this solver runs on one GPU; runs over several GPUs are not supported",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    if args.particle_wind:
        print(json.dumps(particle_wind(args.particle_wind, args.n)))
        return

    from gral_amd import domain, pf

    total = args.warmup + args.steps + 16
    directions = [45.0 * q for q in range(args.directions)]
    model = args.turbulence_model
    kernel_of = KERNEL_OF if model == 1 else KERNEL_OF_KEPS
    alg_iter = ALG_BYTES_ITER if model == 1 else ALG_BYTES_ITER_KEPS
    cases = [domain.config2(direction=d, iterations=total, n=args.n, turbulence_model=model) for d in directions]
    solvers = [pf.PrognosticFlowfield(c) for c in cases]
    for sv in solvers:
        sv.prepare(iterations=total)
    batch = pf.Batch(solvers)
    if args.warmup:
        batch.iterate(args.warmup)
    t0 = time.perf_counter()
    batch.iterate(args.steps)
    wall = time.perf_counter() - t0
    tm = batch.timing()
    ms = tm["ms_total"] / args.steps
    stream_groups = tm["groups"]
    value = tm["cell_iterations"] / (tm["ms_total"] / 1e3)

    if args.no_family_profile:
        batch.close()
        print(json.dumps({"metric": "SIMPLE cell-iters/s", "value": value, "ms_per_step": ms, "steps": args.steps}))
        return
    # dominant kernel: per-launch device time of each sweep family over a few
    # instrumented iterations (events bracket single launches on the batch's stream)
    prof = batch.profile_families(2)
    batch.close()
    gff = gff_timing(solvers[0]) if args.gff else None
    for sv in solvers:
        sv.close()
    sweeps = {k: v for k, v in prof.items() if k in ALG_BYTES}
    dom = max(sweeps, key=lambda k: sweeps[k]["ms"])
    pdom = sweeps[dom]
    avg_ms = pdom["ms"] / max(1, pdom["launches"])
    bytes_per_launch = ALG_BYTES[dom] * pdom["cells"] / max(1, pdom["launches"])
    achieved = bytes_per_launch / (avg_ms / 1e3) / 1e9 if avg_ms > 0 else 0.0
    # HBM bytes per launch of that kernel, from the calibrated FETCH_SIZE / WRITE_SIZE
    # counter passes committed with the profiles (same workload and kernel version)
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "hbm_traffic_city400_8dir.json" if model == 1 else
                         "hbm_traffic_city400_8dir_keps.json")
    if os.path.exists(tpath) and args.n == 400 and args.directions == 8:
        kt = json.load(open(tpath))["kernels"].get(kernel_of[dom])
        if kt:
            traffic = kt["traffic_bytes"]
    roofline = {"bound": "hbm", "kernel": kernel_of[dom], "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                "avg_launch_us": avg_ms * 1e3, "launches_per_iter": pdom["launches"] / 2,
                "alg_bytes_per_cell": ALG_BYTES[dom], "alg_bytes_per_launch": bytes_per_launch}

    # the plain single-situation solve (270 deg), timed the same way
    plain = None
    if not args.no_plain:
        c270 = domain.config2(direction=270.0, iterations=total, n=args.n, turbulence_model=model)
        with pf.PrognosticFlowfield(c270) as sv:
            sv.prepare(iterations=total)
            if args.warmup:
                sv.iterate(args.warmup)
            sv.iterate(args.steps)
            tp = sv.timing()
        plain = {"value": tp["cell_iterations"] / (tp["ms_total"] / 1e3), "ms_per_step": tp["ms_total"] / args.steps,
                 "direction_deg": 270.0, "active_cells": tp["active_cells"]}

    cpu = None
    if not args.no_cpu_baseline and model == 1:   # the timing mode threads the mixing-length sweeps only
        from oracle import pyoracle
        c0 = domain.config2(direction=270.0, iterations=args.cpu_iterations, n=args.n)
        # the AVX2 build evaluates the coefficient lanes 8 at a time (Vector<float>), bit for bit the scalar build
        simd = pyoracle.simd_available()
        lib = pyoracle.SIMD_PATH if simd else None

        def loop_time(iters, threads):
            # the SIMPLE loop alone: a run of `iters` iterations minus a run of none (the prologue)
            t = time.perf_counter(); pyoracle.run(c0, iterations=0, keep_history=False, threads=threads, lib_path=lib)
            t_pro = time.perf_counter() - t
            t = time.perf_counter(); pyoracle.run(c0, iterations=iters, keep_history=False, threads=threads, lib_path=lib)
            return time.perf_counter() - t - t_pro

        # the host share of one GPU on the box is 16 cores (os.cpu_count() may show the whole machine)
        threads = max(1, min(16, os.cpu_count() or 1))
        n_par = 4 * args.cpu_iterations
        reps = sorted(loop_time(n_par, threads) for _ in range(5))
        t_par = reps[2]
        t_one = loop_time(args.cpu_iterations, 1)
        width = ("8-wide AVX2 vectors for the coefficient lanes (the reference's Vector<float> width on AVX2 "
                 "hosts), scalar recurrences as in the reference" if simd else
                 "scalar C (no AVX2 on this host): the reference vectorises the coefficient lanes 8-wide")
        cpu = {"value": c0.active_cells() * n_par / t_par, "unit": "cell-iter/s", "cores": threads, "kind": "port",
               "simd": "avx2-8" if simd else "scalar",
               "sample": f"{n_par} SIMPLE iterations of the 270 deg situation, median of 5 runs "
                         f"({', '.join(f'{x:.2f}' for x in reps)} s), the C restatement run with the reference's "
                         f"default multi-threaded schedule on {threads} threads, the host share of one GPU (rows of "
                         f"each sweep split across threads, U and V concurrent), prologue excluded; {width}",
               "single_core": {"value": c0.active_cells() * args.cpu_iterations / t_one, "cores": 1,
                               "sample": f"{args.cpu_iterations} iterations on one thread, same build "
                                         f"({t_one:.1f} s)"}}

    line = {"metric": "SIMPLE cell-iters/s", "value": value, "unit": "cell-iter/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
            "config": {"workload": f"city block {cases[0].nii}x{cases[0].njj}, dx=2 m, 40 prognostic layers, "
                                   f"30% buildings (seed 20260001), {len(cases)} wind directions "
                                   f"({', '.join(f'{d:g}' for d in directions)} deg) as one lock-step batch, "
                                   f"plain solver per situation, {'mixing-length' if model == 1 else 'k-eps'} model",
                       "situations": len(cases), "active_cells": tm["active_cells"], "kadvmax": cases[0].kadvmax,
                       "stream_groups": stream_groups, "parallelism": "single GPU"},
            "wall_s_timed": wall, "alg_bytes_per_cell_iter": alg_iter,
            "iteration_frac_of_hbm_roofline": value * alg_iter / (HBM_PEAK_GBS * 1e9),
            "roofline": roofline, "cpu_baseline": cpu, "plain_single_situation": plain,
            "kernel_families_ms_per_iter": {k: v["ms"] / 2 for k, v in prof.items()}}
    if gff is not None:
        line["gff_writer"] = gff
    if args.batched_workload:
        line["batched_workload"] = batched_workload(args.batched_workload)
    if args.situations_hour:
        line["situations_per_hour"] = situations_per_hour(args.n, directions)
    if model == 1 and not args.no_keps:
        line["keps"] = keps_timing(args.n, directions)
    if args.diagnostic:
        line["diagnostic_post_pass"] = diagnostic_timing(args.n)
    if args.adjoint:
        line["adjoint"] = adjoint_timing()
    if args.graph:
        line["graph_replay"] = graph_timing()
    if args.batch_sweep:
        line["batch_sweep"] = batch_sweep([int(x) for x in args.batch_sweep.split(",")], args.n)
    if args.convergence_study:
        line["convergence_study"] = convergence_study()
    if args.stream:
        line["stream_workload"] = stream_workload(args.stream, args.capacity, args.stream_plain)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
