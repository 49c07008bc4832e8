"""The boundary's less common inputs and error paths on the GPU, through the C ABI:

* the TKE initialisation from a measured sigma profile (IStatistics zr /
  sonic, PrognosticFlowfield.cs:452-506), bit-exact against the oracle after
  the prologue and after 3 k-eps iterations;
* the unrealistic-speed counters of the epilogue (:1192-1243) on planted
  values at, below and above the thresholds and the counted level range;
* a prepare that fails (device memory budget) leaves the handle unprepared:
  iterate and the flow-field packer refuse with GRAL_PF_ESTATE.
"""
import ctypes as C
import os

import numpy as np
import pytest

from tests.helpers import FIELDS, KE_FIELDS, diff_report
from tests.test_met_profile import _case, unreal_cases


@pytest.mark.gpu
@pytest.mark.parametrize("istat", [0, 3])
def test_met_profile_tke_init(istat):
    from gral_amd import pf
    from oracle import pyoracle
    c = _case(istat)
    print(f"[boundary] met profile istat={istat}", flush=True)
    ref0 = pyoracle.run(c, iterations=0)
    ref3 = pyoracle.run(c, iterations=3)
    with pf.PrognosticFlowfield(c) as sv:
        sv.prepare(iterations=3)
        got0 = sv.download()
        sv.iterate(3)
        sv.finish()
        got3 = sv.download()
    for got, ref, tag in ((got0, ref0, "prologue"), (got3, ref3, "3 iterations")):
        rep = diff_report(got, ref, FIELDS + KE_FIELDS)
        for f, r in rep.items():
            assert r["n_bits_differ"] == 0, (tag, f, r)


@pytest.mark.gpu
@pytest.mark.parametrize("which", [0, 1])
def test_unrealistic_speed_counters(which):
    from gral_amd import pf
    from oracle import pyoracle
    name, c, want = unreal_cases()[which]
    print(f"[boundary] unrealistic speeds: {name}", flush=True)
    ref = pyoracle.run(c, iterations=0)["stats"]
    with pf.PrognosticFlowfield(c) as sv:
        sv.prepare(iterations=1)
        st = sv.finish()
    assert st["geschw_max"] == ref["geschw_max"]
    assert (st["unreal_u"], st["unreal_v"], st["unreal_w"]) == want == (ref["unreal_u"], ref["unreal_v"],
                                                                        ref["unreal_w"])


@pytest.mark.gpu
def test_failed_prepare_leaves_handle_unprepared():
    from gral_amd import domain, pf
    c = domain.config1(iterations=2)             # 2 MB per device field: over a 1 MiB budget
    print("[boundary] failed prepare", flush=True)
    with pf.PrognosticFlowfield(c) as sv:
        sv.prepare(iterations=2)                  # a good prepare first
        sv.iterate(1)
        os.environ["GRAL_PF_DEVICE_MB"] = "1"
        try:
            small = pf.PrognosticFlowfield(c)
        finally:
            os.environ.pop("GRAL_PF_DEVICE_MB", None)
        with small:
            with pytest.raises(RuntimeError, match=r"\(-3\)"):
                small.prepare(iterations=2)
            st = pf.Stats()
            assert small.lib.gral_pf_iterate(small.h, 1, 0, C.byref(st)) == -4
            assert small.lib.gral_pf_finish(small.h, C.byref(st)) == -4
        # the budget is per handle: the first handle still works
        sv.iterate(1)
        assert np.isfinite(sv.download()["uk"]).all()
