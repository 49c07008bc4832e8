"""The FP64 shadow of the oracle (the same restatement with every float
widened to double), used to state the FP32 solve's error per field.

Checks that the shadow is the same algorithm (the quiescent state stays an
exact fixed point; the FP32 run lies within the target tolerance of it,
relative L2 <= 1e-4 and max-abs <= 1e-3 m/s) and that it is a different
computation (its fields are not the FP32 bits)."""
import numpy as np
import pytest

from gral_amd import domain


@pytest.mark.parametrize("model", [1, 2])
def test_fp32_within_target_of_fp64(model):
    from oracle.fp64_error import field_errors
    c = domain.tiny_case(n=10, vertcellsff=9, turbulence_model=model, iterations=2)
    err = field_errors(c, 2)
    for f in ("uk", "vk", "wk"):
        assert err[f]["rel_l2"] <= 1e-4 and err[f]["max_abs"] <= 1e-3, (f, err[f])
    assert any(e["max_abs"] > 0 for e in err.values())


def test_quiescent_fixed_point_fp64():
    from oracle import pyoracle
    c = domain.tiny_case(n=9, vertcellsff=8, turbulence_model=1, iterations=3)
    for a in (c.uk, c.vk, c.wk):
        a[:] = 0
    r = pyoracle.run(c, iterations=3, fp64=True)
    assert r["uk"].dtype == np.float64
    for f in ("uk", "vk", "wk", "dpm"):
        assert not r[f].any(), f
