"""gral_pf_mass_imbalance_noalloc (the reduction on the library's fixed device
scratch, no allocation) against gral_pf_mass_imbalance (the same kernels on
scratch the handle allocates), through the C ABI: two calls in a row, the
second on a smaller grid (fewer partial sums than the first left in the
scratch), then the first grid again.  The results must be the same bits."""
import ctypes as C

import numpy as np
import pytest

from gral_amd import domain


def _both(sv):
    lib = sv.lib
    a, b = C.c_double(0), C.c_double(0)
    assert lib.gral_pf_mass_imbalance_noalloc(sv.h, C.byref(a), C.byref(b)) == 0
    ref = sv.mass_imbalance()
    return (a.value, b.value), ref


@pytest.mark.gpu
def test_noalloc_matches_allocating_entry_point():
    from gral_amd import pf
    lib = pf.load()
    lib.gral_pf_mass_imbalance_noalloc.argtypes = [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double)]
    lib.gral_pf_mass_imbalance_noalloc.restype = C.c_int
    big = domain.config1(turbulence_model=1, iterations=2)
    small = domain.tiny_case(n=9, vertcellsff=8, turbulence_model=1, iterations=2)
    print("[no-alloc] config1 then tiny then config1", flush=True)
    with pf.PrognosticFlowfield(big) as sb, pf.PrognosticFlowfield(small) as ss:
        sb.prepare(iterations=2)
        ss.prepare(iterations=2)
        sb.iterate(2)
        ss.iterate(2)
        got, ref = _both(sb)
        assert got == ref and got[1] > 0
        got_s, ref_s = _both(ss)
        assert got_s == ref_s and got_s[1] > 0
        assert got_s != got
        got2, _ = _both(sb)
        assert got2 == ref
