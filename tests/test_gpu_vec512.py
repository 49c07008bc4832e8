"""GPU parity of the 512-bit mixing-length solvers (simd_width 16,
U/V/W-prognostic-microscale_1_Vec512.cs): 16-lane blocks and their fused
multiply-adds, bit for bit against the oracle, on both sweep schedules and in
a batch next to width-8 situations."""
import numpy as np
import pytest

from gral_amd import domain
from tests.test_gpu_parity import _compare


def _w16(c):
    c.simd_width = 16
    return c


@pytest.mark.gpu
@pytest.mark.parametrize("iterations", [1, 5])
def test_tiny_building_w16(iterations):
    _compare(_w16(domain.tiny_case(n=9, vertcellsff=16, turbulence_model=1, iterations=iterations)), iterations)


@pytest.mark.gpu
def test_tiny_vegetation_w16():
    _compare(_w16(domain.tiny_case(n=10, vertcellsff=17, turbulence_model=1, iterations=4, veg=True)), 4)


@pytest.mark.gpu
def test_two_buildings_w16():
    # building columns start at k=1: 16-wide blocks straddle KKART+1 and the
    # moved-down last block of a column overlaps levels already solved
    bl = [domain.Building(3, 4, 3, 5, 6.5), domain.Building(6, 7, 4, 6, 3.2)]
    c = _w16(domain.build_case("two", 10, 10, bl, vertcellsff=18, direction=235.0, iterations=4))
    _compare(c, 4)


@pytest.mark.gpu
def test_small_city_w16():
    # the reference needs at least 16 prognostic layers for 16-lane blocks
    bl = domain.city_buildings(48, 48, 2.0, 11, density=0.25, fmin=6.0, fmax=16.0, hmin=4.0, hmax=14.0, margin=3)
    c = domain.build_case("city48", 48, 48, bl, dx=2.0, vertcellsff=16, speed10=4.0, direction=235.0, z0=0.5,
                          turbulence_model=1, iterations=9)
    _compare(_w16(c), 9)


@pytest.mark.gpu
def test_config1_w16():
    _compare(_w16(domain.config1(turbulence_model=1, iterations=8)), 8)


@pytest.mark.gpu
def test_batch_mixed_widths():
    # one lock-step batch holding a width-16 and a width-8 situation: each
    # situation keeps its own solver
    from tests.test_gpu_batch import _check_against_oracle, _run_batch
    its = [6, 6]
    cases = [_w16(domain.tiny_case(n=12, vertcellsff=16, turbulence_model=1, iterations=6)),
             domain.tiny_case(n=12, vertcellsff=16, turbulence_model=1, iterations=6)]
    print(f"[vec512] batch of widths {[c.simd_width for c in cases]} it={its}", flush=True)
    fields, stats, _ = _run_batch(cases, its)
    _check_against_oracle(cases, fields, stats, its)
    assert not np.array_equal(fields[0]["uk"].view(np.uint32), fields[1]["uk"].view(np.uint32))
