"""The oracle's CPU timing build (liboracle_pf_simd.so: the mixing-length
coefficient lanes evaluated 8 at a time in AVX2 vectors, the reference's
Vector<float> width) against the default scalar build: the whole solve must be
bit-identical, since every vector operation is the lane-wise IEEE operation of
the scalar form in the same order.  The cases cover all four sweep phases, the
block that holds KKART+1 (the wall term and the masked cross terms), columns
with CUTK > 0 (KSTART = 1), vegetation, and sub-domain edges."""
import numpy as np
import pytest

from gral_amd import domain
from oracle import pyoracle

FIELDS = ("uk", "vk", "wk", "dpm", "dpmnew", "uks", "vks", "wks")

pytestmark = pytest.mark.skipif(not pyoracle.simd_available(), reason="no AVX2 timing build or no AVX2 on this CPU")


def _city():
    bl = domain.city_buildings(40, 40, 2.0, 5, density=0.25, fmin=6.0, fmax=16.0, hmin=4.0, hmax=14.0, margin=3)
    return domain.build_case("city40", 40, 40, bl, dx=2.0, vertcellsff=12, speed10=4.0, direction=235.0, z0=0.5,
                             iterations=5)


@pytest.mark.parametrize("name,make,iters", [
    ("config1", lambda: domain.config1(iterations=6), 6),
    ("city", _city, 5),
    ("vegetation", lambda: domain.tiny_case(n=12, vertcellsff=10, veg=True, iterations=4), 4),
])
def test_simd_build_bit_identical(name, make, iters):
    c = make()
    ref = pyoracle.run(c, iterations=iters, keep_history=False)
    got = pyoracle.run(c, iterations=iters, keep_history=False, lib_path=pyoracle.SIMD_PATH)
    for f in FIELDS:
        if f not in ref:
            continue
        a, b = np.asarray(ref[f]), np.asarray(got[f])
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), (name, f, np.abs(a - b).max())


def test_vegetation_case_has_vegetation_columns():
    c = domain.tiny_case(n=12, vertcellsff=10, veg=True, iterations=4)
    assert np.asarray(c.veg_present).any() and np.asarray(c.veg).any()
