"""A batch fed while it runs (gral_pf_batch_add / gral_pf_batch_step): a
situation admitted after 100 iterations runs 100 iterations behind the
others in the same launches, situations leave at their stop test or
iteration limit, and every one of them equals the oracle's plain solve bit
for bit (fields and convergence figures)."""
import numpy as np
import pytest

from gral_amd import domain
from tests.helpers import FIELDS, diff_report

STAT_KEYS = ("iteration_loops", "iterations_run", "delta_first", "delta_finish", "topwind", "dtime", "frund",
             "geschw_max", "unreal_u", "unreal_v", "unreal_w", "convergence_problem", "ug", "vg")


def _cases():
    a = domain.tiny_case(n=9, vertcellsff=8, turbulence_model=1, iterations=200)
    b = domain.tiny_case(n=12, vertcellsff=9, turbulence_model=1, iterations=250, direction=200.0)
    c = domain.tiny_case(n=10, vertcellsff=8, turbulence_model=1, iterations=150, direction=300.0)
    for x, lo in ((a, 100), (b, 100), (c, 100)):
        x.iter_min = lo
    return a, b, c


@pytest.mark.gpu
def test_late_admission_equals_plain_solves():
    from gral_amd import pf
    from oracle import pyoracle
    a, b, c = _cases()
    print("[stream] a, b from the start; c admitted after 100 iterations", flush=True)
    got = {}
    sa, sb, sc = pf.PrognosticFlowfield(a), pf.PrognosticFlowfield(b), pf.PrognosticFlowfield(c)
    try:
        for s in (sa, sb, sc):
            s.prepare()
        with pf.Batch() as batch:
            batch.add(sa)
            batch.add(sb)
            done = batch.step()                         # iterations 1..100 of a and b
            assert batch.size() == (2 - len(done), 0)
            for sv, st in done:
                got[sv.case.name + str(sv.case.nii)] = (sv.download(), st)
            batch.add(sc)                               # joins at a's and b's iteration 101
            while batch.size() != (0, 0):
                for sv, st in batch.step():
                    got[sv.case.name + str(sv.case.nii)] = (sv.download(), st)
    finally:
        for s in (sa, sb, sc):
            s.close()
    assert len(got) == 3
    for case in (a, b, c):
        ref = pyoracle.run(case)
        g, st = got[case.name + str(case.nii)]
        for f, r in diff_report(g, ref, FIELDS).items():
            assert r["n_bits_differ"] == 0, (case.nii, f, r)
        for k in STAT_KEYS:
            assert ref["stats"][k] == st[k], (case.nii, k, ref["stats"][k], st[k])


@pytest.mark.gpu
def test_solve_stream_capacity_two():
    from gral_amd import pf
    from oracle import pyoracle
    cases = list(_cases())
    print("[stream] solve_stream, capacity 2", flush=True)
    res = {}
    stats = pf.solve_stream(cases, capacity=2,
                            on_done=lambda case, sv, st: res.__setitem__(case.nii, (sv.download(), st)))
    assert len(stats) == 3
    for case in cases:
        ref = pyoracle.run(case)
        g, st = res[case.nii]
        assert np.array_equal(g["uk"].view(np.uint32), ref["uk"].view(np.uint32)), case.nii
        assert st["iteration_loops"] == ref["stats"]["iteration_loops"]
