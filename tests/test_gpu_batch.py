"""Batched solves: several situations advanced in lock step by the same launches.

Every situation of a batch must come out bit-identical to the oracle's plain
solve of that situation alone (same fields, same iteration count, same
convergence figures), whatever the other situations are: different grid
sizes, wind directions, iteration limits, and stop iterations of the
reference's convergence test (early retirement).
"""
import numpy as np
import pytest

from gral_amd import domain
from tests.helpers import FIELDS, KE_FIELDS, diff_report

STAT_KEYS = ("iteration_loops", "iterations_run", "delta_first", "delta_finish", "topwind", "dtime", "frund",
             "geschw_max", "unreal_u", "unreal_v", "unreal_w", "convergence_problem", "ug", "vg")


def _city(n, model, direction, seed=11):
    bl = domain.city_buildings(n, n, 2.0, seed, density=0.25, fmin=6.0, fmax=16.0, hmin=4.0, hmax=14.0, margin=3)
    return domain.build_case(f"city{n}", n, n, bl, dx=2.0, vertcellsff=12, speed10=4.0, direction=direction, z0=0.5,
                             turbulence_model=model, iterations=10)


def _check_against_oracle(cases, got_fields, got_stats, iterations=None):
    from oracle import pyoracle
    for q, c in enumerate(cases):
        ref = pyoracle.run(c, iterations=None if iterations is None else iterations[q])
        fields = FIELDS + (KE_FIELDS if c.turbulence_model != 1 else ())
        rep = diff_report(got_fields[q], ref, fields)
        for f, r in rep.items():
            assert r["n_bits_differ"] == 0, (q, c.name, f, r)
        for k in STAT_KEYS:
            assert ref["stats"][k] == got_stats[q][k], (q, c.name, k, ref["stats"][k], got_stats[q][k])


def _run_batch(cases, iterations=None, solve=False):
    from gral_amd import pf
    solvers = [pf.PrognosticFlowfield(c) for c in cases]
    try:
        for q, s in enumerate(solvers):
            s.prepare(None if iterations is None else iterations[q])
        with pf.Batch(solvers) as b:
            if solve:
                stats = b.solve()
            else:
                b.iterate(max(iterations))
                stats = [s.finish() for s in solvers]
            t = b.timing()
        fields = [s.download() for s in solvers]
    finally:
        for s in solvers:
            s.close()
    return fields, stats, t


@pytest.mark.gpu
def test_batch_mixed_sizes_fixed_iterations():
    cases = [domain.tiny_case(n=9, vertcellsff=8, turbulence_model=1, iterations=5),
             domain.tiny_case(n=12, vertcellsff=9, turbulence_model=1, iterations=7, direction=200.0),
             _city(24, 1, 135.0)]
    its = [5, 7, 6]
    print(f"[batch] fixed iterations {its}, sizes {[c.nii for c in cases]}", flush=True)
    fields, stats, t = _run_batch(cases, its)
    _check_against_oracle(cases, fields, stats, its)
    assert t["situations"] == 3
    assert t["cell_iterations"] == sum(c.active_cells() * n for c, n in zip(cases, its))


@pytest.mark.gpu
def test_batch_early_retirement():
    # with IterationStepsMin 100 and Max 400, the reference's stop test ends the
    # 9x9 and 16x16 situations at 300 iterations and runs the 12x12 one to 400
    cases = []
    for n, d in ((9, 250.0), (12, 200.0), (16, 100.0)):
        c = domain.tiny_case(n=n, vertcellsff=8, turbulence_model=1, iterations=400, direction=d)
        c.iter_min, c.iter_max = 100, 400
        cases.append(c)
    print("[batch] early retirement, 3 situations, up to 400 iterations", flush=True)
    fields, stats, _ = _run_batch(cases, solve=True)
    _check_against_oracle(cases, fields, stats)
    loops = [s["iteration_loops"] for s in stats]
    assert len(set(loops)) > 1, loops


@pytest.mark.gpu
@pytest.mark.parametrize("model", [2, 0])
def test_batch_other_models(model):
    cases = [domain.tiny_case(n=9, vertcellsff=8, turbulence_model=model, iterations=3),
             domain.tiny_case(n=11, vertcellsff=8, turbulence_model=model, iterations=4, direction=40.0)]
    its = [3, 4]
    print(f"[batch] model {model}, iterations {its}", flush=True)
    fields, stats, _ = _run_batch(cases, its)
    _check_against_oracle(cases, fields, stats, its)


@pytest.mark.gpu
def test_batch_equals_plain_gpu_solve():
    # eight wind directions over one city, as the multi-direction runs of a project
    from gral_amd import pf
    cases = [_city(32, 1, d) for d in range(0, 360, 45)]
    its = [6] * len(cases)
    print("[batch] 8 directions of a 32x32 city vs plain GPU solves", flush=True)
    fields, stats, _ = _run_batch(cases, its)
    for q, c in enumerate(cases):
        plain = pf.run(c, iterations=6)
        for f in FIELDS:
            assert np.array_equal(plain[f].view(np.uint32), fields[q][f].view(np.uint32)), (q, f)
        for k in STAT_KEYS:
            assert plain["stats"][k] == stats[q][k], (q, k)
