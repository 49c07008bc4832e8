"""Golden fixtures (tests/golden/, written by tests/golden/make_golden.py).

CPU: the input generator and the oracle still reproduce the fixtures bit for
bit.  GPU: the HIP solver reproduces them bit for bit (config 1 at 200
iterations included) without running the oracle.
"""
import json
import os

import numpy as np
import pytest

from tests.golden import make_golden as mg

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
META = json.load(open(os.path.join(GOLDEN, "golden.json")))
SOLVE = [n for n, m in META.items() if "diagnostic" not in m]
DIAG = [n for n, m in META.items() if "diagnostic" in m]
CPU_CASES = [n for n in SOLVE if n != "c1_m1_it200"]


def _check(name, out):
    m = META[name]
    for f, h in m["field_sha256"].items():
        a = mg.crop(out[f], m["kadvmax"])
        if mg.digest(a) != h:
            path = os.path.join(GOLDEN, f"{name}.npz")
            extra = ""
            if os.path.exists(path):
                ref = np.load(path)[f]
                extra = f" max abs diff {float(np.abs(a.astype(np.float64) - ref).max()):.3e}"
            raise AssertionError(f"{name}: field {f} differs from the fixture{extra}")
    for k, v in m["stats"].items():
        assert out["stats"][k] == pytest.approx(v, rel=0, abs=0), (name, k, out["stats"][k], v)


@pytest.mark.parametrize("name", list(META))
def test_generator_reproduces_inputs(name):
    m = META[name]
    c = mg.make_case(m["generator"], m["kwargs"], max(1, m["iterations"]))
    assert mg.input_digest(c) == m["input_sha256"]
    assert c.kadvmax == m["kadvmax"]


@pytest.mark.parametrize("name", CPU_CASES)
def test_oracle_reproduces_fixture(name):
    from oracle import pyoracle
    m = META[name]
    c = mg.make_case(m["generator"], m["kwargs"], m["iterations"])
    _check(name, pyoracle.run(c, iterations=m["iterations"], keep_history=False))


def test_fixture_arrays_match_digests():
    for name in SOLVE:
        m = META[name]
        path = os.path.join(GOLDEN, f"{name}.npz")
        if not os.path.exists(path):
            continue
        z = np.load(path)
        for f, h in m["field_sha256"].items():
            assert mg.digest(z[f]) == h, (name, f)


def _diag_oracle(name):
    from oracle import pyoracle
    m = META[name]
    c = mg.make_case(m["generator"], m["kwargs"], max(1, m["iterations"]))
    f = None
    if m["iterations"]:
        r = pyoracle.run(c, iterations=m["iterations"], keep_history=False)
        f = {k: r[k] for k in mg.DIAG_FIELDS}
    return pyoracle.diagnostic(c, f, flow_field_level=m["diagnostic"], terrain=m["terrain"])


@pytest.mark.parametrize("name", DIAG)
def test_oracle_reproduces_diagnostic_fixture(name):
    d = _diag_oracle(name)
    for k, h in META[name]["field_sha256"].items():
        assert mg.digest(d[k]) == h, (name, k)
    assert d["dtime"] == META[name]["dtime"]


@pytest.mark.gpu
@pytest.mark.parametrize("name", DIAG)
def test_gpu_reproduces_diagnostic_fixture(name):
    from gral_amd import pf
    m = META[name]
    c = mg.make_case(m["generator"], m["kwargs"], max(1, m["iterations"]))
    print(f"[golden] {name}: {m['generator']} {m['kwargs']} it={m['iterations']} level={m['diagnostic']}", flush=True)
    with pf.PrognosticFlowfield(c) as s:
        if m["iterations"]:
            s.prepare(m["iterations"])
            s.iterate(m["iterations"])
            s.finish()
        s.diagnostic(m["diagnostic"], m["terrain"])
        out = s.download()
    for k, h in m["field_sha256"].items():
        assert mg.digest(out[k]) == h, (name, k)


@pytest.mark.gpu
@pytest.mark.parametrize("name", SOLVE)
def test_gpu_reproduces_fixture(name):
    from gral_amd import pf
    m = META[name]
    c = mg.make_case(m["generator"], m["kwargs"], m["iterations"])
    print(f"[golden] {name}: {m['generator']} {m['kwargs']} it={m['iterations']}", flush=True)
    _check(name, pf.run(c, iterations=m["iterations"]))
