"""GPU parity at the configurations' own sizes, bit-exact against the oracle.

* The 400 x 400 x 40 city block (30 % density, heights 6-40 m, KADVMAX 74,
  27 strips per sweep): two wind directions, 3 iterations (sweep phases 1-3,
  PrognosticFlowfield.cs:770-802), plain and as a two-situation batch.
* The batched workload's situations: four of the 256 generated situations,
  set up on the GPU (the largest window, a 30-layer and a 50-layer one), as
  one batch of 2 iterations.
* The terrain configuration: 1000 x 1000 cells at 4 m over 12 Gaussian hills
  with buildings, vegetation and k-eps, 30 prognostic layers, 2 iterations.
"""
import time

import numpy as np
import pytest

from gral_amd import domain
from tests.helpers import FIELDS, KE_FIELDS, diff_report

_ORACLE = {}


def _oracle(key, case, iterations):
    from oracle import pyoracle
    if key not in _ORACLE:
        t0 = time.time()
        _ORACLE[key] = pyoracle.run(case, iterations=iterations)
        print(f"[configs] oracle {key}: {time.time() - t0:.1f} s", flush=True)
    return _ORACLE[key]


def _assert_equal(case, got, ref, tag):
    fields = FIELDS + (KE_FIELDS if case.turbulence_model != 1 else ())
    rep = diff_report(got, ref, fields)
    for f, r in rep.items():
        assert r["n_bits_differ"] == 0, (tag, f, r)
    rs, gs = ref["stats"], got["stats"]
    for k in ("iterations_run", "topwind", "dtime", "frund", "geschw_max", "ug", "vg"):
        assert rs[k] == gs[k], (tag, k, rs[k], gs[k])


def _plain(case, iterations):
    from gral_amd import pf
    return pf.run(case, iterations=iterations)


C2_DIRS = (270.0, 45.0)


@pytest.mark.gpu
@pytest.mark.parametrize("direction", C2_DIRS)
def test_city400_plain(direction):
    c = domain.config2(direction, iterations=3)
    assert (c.nii, c.njj, c.vertcellsff, c.kadvmax) == (400, 400, 40, 74)
    print(f"[configs] city400 dir={direction} plain, 3 iterations", flush=True)
    got = _plain(c, 3)
    _assert_equal(c, got, _oracle(("c2", direction), c, 3), f"city400-{direction}")


@pytest.mark.gpu
def test_city400_batch_two_directions():
    from gral_amd import pf
    cases = [domain.config2(d, iterations=3) for d in C2_DIRS]
    print("[configs] city400 batch of 2 directions, 3 iterations", flush=True)
    got = pf.run_batch(cases, iterations=3)
    for c, g, d in zip(cases, got, C2_DIRS):
        _assert_equal(c, g, _oracle(("c2", d), c, 3), f"city400-batch-{d}")


@pytest.mark.gpu
def test_batched_workload_situations():
    from gral_amd import pf
    sits = domain.config3_situations()
    big = max(sits, key=lambda s: s["n"])
    l30 = next(s for s in sits if s["vertcellsff"] == 30)
    l50 = next(s for s in sits if s["vertcellsff"] == 50)
    pick = [big, l30, l50, sits[0]]
    print("[configs] batched workload:",
          [(s["index"], s["n"], s["vertcellsff"], s["direction"], s["stability"]) for s in pick], flush=True)
    cases = [domain.config3(s, iterations=2) for s in pick]
    got = pf.run_batch(cases, iterations=2)
    for s, c, g in zip(pick, cases, got):
        print(f"[configs] situation {s['index']}: {c.nii}x{c.njj} kadvmax={c.kadvmax}", flush=True)
        _assert_equal(c, g, _oracle(("c3", s["index"]), c, 2), f"situation-{s['index']}")


@pytest.mark.gpu
def test_terrain_configuration_30_layers():
    c = domain.config5(direction=200.0, speed=4.0, stability=2, iterations=2, vertcellsff=30)
    vi = np.minimum(c.nkk - 1, c.kkart.astype(np.int64) + 30)
    print(f"[configs] terrain {c.nii}x{c.njj} nkk={c.nkk} kadvmax={c.kadvmax} "
          f"active={c.active_cells()} ({c.active_cells() / (c.nii - 2) ** 2:.1f} per column)", flush=True)
    assert c.kadvmax + 2 <= 128 and c.vertcellsff == 30
    assert vi.max() <= c.kadvmax
    got = _plain(c, 2)
    _assert_equal(c, got, _oracle(("c5", 200.0), c, 2), "terrain")
