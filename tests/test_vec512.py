"""The 512-bit mixing-length solvers (U/V/W-prognostic-microscale_1_Vec512.cs,
selected by simd_width 16; PrognosticFlowfield.cs:1072-1091).

They differ from the Vector<float> solvers in their block width (16 lanes) and
in fused multiply-adds: the ground height offset, the power-law coefficients,
the neighbour part of the source term (not for V without vegetation) and, for
U without vegetation, the pressure term.  The C oracle (C library fmaf) is
pinned against the numpy transcription (fma32: float64 product and TwoSum),
and fma32 itself against exact rational arithmetic.
"""
from fractions import Fraction

import numpy as np
import pytest

from gral_amd import domain
from oracle import pyoracle
from tests import np_transcription
from tests.np_transcription import fma32

F = np.float32


def _round_f32(x: Fraction) -> np.float32:
    """x rounded to the nearest float32, ties to even, by exact comparison."""
    c = F(float(x))
    cands = {c, np.nextafter(c, F(np.inf)), np.nextafter(c, F(-np.inf))}
    best = None
    for v in cands:
        if not np.isfinite(v):
            continue
        err = abs(Fraction(float(v)) - x)
        key = (err, int(np.array(v, dtype=F).view(np.uint32)) & 1)   # ties: even mantissa first
        if best is None or key < best[0]:
            best = (key, v)
    return best[1]


def _exact_fma(a, b, c):
    return _round_f32(Fraction(float(a)) * Fraction(float(b)) + Fraction(float(c)))


def test_fma32_random_against_exact():
    rng = np.random.default_rng(11)
    n = 3000
    a = (rng.standard_normal(n) * 10.0 ** rng.integers(-3, 4, n)).astype(F)
    b = (rng.standard_normal(n) * 10.0 ** rng.integers(-3, 4, n)).astype(F)
    c = (-(a.astype(np.float64) * b) * (1 + rng.standard_normal(n) * 1e-6)).astype(F)   # heavy cancellation
    c[::3] = rng.standard_normal(n)[::3].astype(F)
    got = fma32(a, b, c)
    for t in range(n):
        want = _exact_fma(a[t], b[t], c[t])
        assert got[t].view(np.uint32) == np.array(want, dtype=F).view(np.uint32), (t, a[t], b[t], c[t], got[t], want)


def test_fma32_double_rounding_ties():
    # a*b = 1 + 2^-24 + 2^-47 exactly: a float32 tie plus a tail below half a
    # float64 ulp of the sum, so rounding the float64 sum would land on the tie
    a = F(1 + 2.0 ** -23)
    b = F(1 + 2.0 ** -23)
    # exact a*b = 1 + 2^-22 + 2^-46; choose c so that p + c = 1 + 2^-24 + tiny
    for c in (F(-(2.0 ** -22) + 2.0 ** -24), F(-(2.0 ** -22) + 2.0 ** -24 + 2.0 ** -40),
              F(-(2.0 ** -22)), F(-(2.0 ** -23))):
        want = _exact_fma(a, b, c)
        got = fma32(a, b, c)
        assert np.array(got, dtype=F).view(np.uint32) == np.array(want, dtype=F).view(np.uint32), (c, got, want)
    # the plain product-then-sum must differ from the fused result somewhere
    assert F(F(a * b) + F(-(2.0 ** -22) + 2.0 ** -24)) != fma32(a, b, F(-(2.0 ** -22) + 2.0 ** -24))


def _case(**kw):
    c = domain.tiny_case(**kw)
    c.simd_width = 16
    return c


def _agree(case, iterations):
    ref = pyoracle.run(case, iterations=iterations)
    npy = np_transcription.run(case, iterations)
    for f in ("uk", "vk", "wk", "uks", "vks", "wks", "dpm", "dpmnew"):
        bad = np.argwhere(ref[f].view(np.uint32) != npy[f].view(np.uint32))
        assert bad.size == 0, (f, bad[:5].tolist())
    assert np.array_equal(ref["stats"]["delta_umax_hist"], npy["delta_umax_hist"])
    return ref


def test_oracle_vec512_matches_transcription_all_phases():
    _agree(_case(n=8, vertcellsff=16, turbulence_model=1, iterations=5), 5)


def test_oracle_vec512_vegetation():
    # vegetation: the fused V branch and the unfused U pressure term
    _agree(_case(n=8, vertcellsff=17, turbulence_model=1, iterations=3, veg=True), 3)


def test_oracle_vec512_two_buildings():
    # building columns start at k=1: the 16-wide blocks straddle KKART+1 and
    # the last block of a column is moved down over levels already solved
    bl = [domain.Building(3, 4, 3, 5, 6.5), domain.Building(6, 7, 4, 6, 3.2)]
    c = domain.build_case("two", 10, 10, bl, vertcellsff=18, direction=235.0, iterations=4)
    c.simd_width = 16
    _agree(c, 4)


def test_vec512_differs_from_vector256():
    # the fused forms must change the result (a test of W=16 that passed with
    # the unfused arithmetic would pin nothing)
    c = _case(n=8, vertcellsff=16, turbulence_model=1, iterations=4)
    a = pyoracle.run(c, iterations=4)
    c8 = c.copy()
    c8.simd_width = 8
    b = pyoracle.run(c8, iterations=4)
    assert any(not np.array_equal(a[f].view(np.uint32), b[f].view(np.uint32)) for f in ("uk", "vk", "wk"))


@pytest.mark.parametrize("model", [0, 2])
def test_width_only_affects_mixing_length(model):
    # the k-eps and no-diffusion solvers have no 512-bit variant: width 16 must
    # reproduce width 8 bit for bit
    c = domain.tiny_case(n=8, vertcellsff=16, turbulence_model=model, iterations=3)
    a = pyoracle.run(c, iterations=3)
    c.simd_width = 16
    b = pyoracle.run(c, iterations=3)
    for f in ("uk", "vk", "wk", "dpm"):
        assert np.array_equal(a[f].view(np.uint32), b[f].view(np.uint32)), f
