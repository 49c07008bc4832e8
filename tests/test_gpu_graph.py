"""Replayed iteration chunks (GRAL_PF_GRAPH=1: the first chunk of a set of
situations runs eagerly and is captured as a hipGraph, later chunks replay it).
Every replay must give the eager launches' results bit for bit: a batch of
three situations of different sizes is advanced 20 iterations per call (one
replay per call after the first, with the host's residual bookkeeping and
synchronisations between replays), its fields and convergence figures
compared with an eager batch of the same situations after every call; the
same for a single situation, and for a batch fed while it runs."""
import os

import numpy as np
import pytest

from gral_amd import domain
from tests.helpers import FIELDS, diff_report

STAT_KEYS = ("iteration_loops", "iterations_run", "delta_first", "delta_finish", "last_delta")


def _cases():
    a = domain.tiny_case(n=9, vertcellsff=8, turbulence_model=1, iterations=120)
    b = domain.tiny_case(n=14, vertcellsff=9, turbulence_model=1, iterations=120, direction=200.0)
    bl = domain.city_buildings(30, 30, 2.0, 3, density=0.25, fmin=6.0, fmax=16.0, hmin=4.0, hmax=14.0, margin=3)
    c = domain.build_case("city30", 30, 30, bl, dx=2.0, vertcellsff=12, speed10=4.0, direction=235.0, z0=0.5,
                          iterations=120)
    return [a, b, c]


def _with_graphs(on, fn):
    old = os.environ.get("GRAL_PF_GRAPH")
    os.environ["GRAL_PF_GRAPH"] = "1" if on else "0"
    try:
        return fn()
    finally:
        if old is None:
            os.environ.pop("GRAL_PF_GRAPH", None)
        else:
            os.environ["GRAL_PF_GRAPH"] = old


def _compare(tag, ga, ea):
    for g, e in zip(ga, ea):
        for f, r in diff_report(g, e, FIELDS).items():
            assert r["n_bits_differ"] == 0, (tag, f, r)


@pytest.mark.gpu
def test_batch_replays_equal_eager():
    from gral_amd import pf
    cs = _cases()
    sg = [pf.PrognosticFlowfield(c) for c in cs]
    se = [pf.PrognosticFlowfield(c) for c in cs]
    try:
        for s in sg + se:
            s.prepare()
        with pf.Batch(sg) as bg, pf.Batch(se) as be:
            for call in range(6):                      # 120 iterations: 1 eager + captured chunk, 5 replays
                print(f"[graph] batch call {call}", flush=True)
                stg = _with_graphs(True, lambda: bg.iterate(20, check_convergence=True))
                ste = _with_graphs(False, lambda: be.iterate(20, check_convergence=True))
                _compare(f"call {call}", [s.download() for s in sg], [s.download() for s in se])
                for x, y in zip(stg, ste):
                    for k in STAT_KEYS:
                        assert x[k] == y[k], (call, k, x[k], y[k])
    finally:
        for s in sg + se:
            s.close()


@pytest.mark.gpu
def test_single_situation_replays_equal_oracle():
    from gral_amd import pf
    from oracle import pyoracle
    c = _cases()[2]
    with pf.PrognosticFlowfield(c) as sv:
        sv.prepare(iterations=60)
        for call in range(3):
            print(f"[graph] single call {call}", flush=True)
            _with_graphs(True, lambda: sv.iterate(20))
            ref = pyoracle.run(c, iterations=20 * (call + 1))
            for f, r in diff_report(sv.download(), ref, FIELDS).items():
                assert r["n_bits_differ"] == 0, (call, f, r)


@pytest.mark.gpu
def test_stream_with_replays_equals_oracle():
    from gral_amd import pf
    from oracle import pyoracle
    cs = _cases()
    for c in cs:
        c.iter_min = 100
        c.iter_max = 200
    print("[graph] solve_stream with replays", flush=True)
    res = {}
    _with_graphs(True, lambda: pf.solve_stream(cs, capacity=2,
                                               on_done=lambda case, sv, st: res.__setitem__(case.nii, sv.download())))
    for c in cs:
        ref = pyoracle.run(c)
        assert np.array_equal(res[c.nii]["uk"].view(np.uint32), ref["uk"].view(np.uint32)), c.nii
