"""The particle model's wind interpolation (Intwind.cs:31-170, BinarySearch
:174-210), CPU side: the oracle pinned by closed forms.  The reference ships
no fixtures for it; the checks restate what it fixes in another form: the
level searches against a sorted search, and the interpolation of fields that
are linear in x, y and z, at coordinates 600 km from the origin, where an
offset evaluated in float instead of double is off by centimetres."""
import numpy as np
import pytest

from oracle import intwind_oracle as W
from tests import flat_cases as FC

F = np.float32


def _first_top_at_or_above(hok, nkk, h):
    k = int(np.searchsorted(hok[1:nkk + 1], F(h), side="left")) + 1
    return min(k, nkk)


@pytest.mark.parametrize("nkk", [6, 40, 221])
def test_binary_search_is_the_first_level_reaching_the_height(nkk):
    dzk, hok = FC.vertical_grid(nkk, 1.0, 1.03)
    r = np.random.default_rng(nkk)
    hs = np.concatenate([r.uniform(-3.0, float(hok[nkk]) + 5.0, 4000).astype(F), hok[:nkk + 2].astype(F)])
    for h in hs:
        assert W.binary_search(hok, nkk, h) == _first_top_at_or_above(hok, nkk, h), h


def _fields(nii, njj, nkk, hok, a=0.0, b=1.0, c=0.5):
    # u linear in x across the faces, v linear in y, w linear in the height of
    # the level's bottom face
    i = np.arange(nii + 2, dtype=F)[:, None, None]
    j = np.arange(njj + 2, dtype=F)[None, :, None]
    uk = np.broadcast_to(F(a) + F(b) * (i - 1), (nii + 2, njj + 2, nkk + 2)).astype(F)
    vk = np.broadcast_to(F(a) - F(b) * (j - 1), (nii + 2, njj + 2, nkk + 2)).astype(F)
    wk = np.zeros((nii + 2, njj + 2, nkk + 2), F)
    wk[:, :, 1:] = (F(c) * hok[None, None, 0:nkk + 1]).astype(F)
    return uk, vk, wk


@pytest.mark.parametrize("topo", [0, 1])
def test_linear_fields_interpolate_exactly(topo):
    nii, njj, nkk, dx = 12, 10, 30, 2.0
    dzk, hok = FC.vertical_grid(nkk, 2.0, 1.04)
    uk, vk, wk = _fields(nii, njj, nkk, hok)
    ikoo, jkoo, ahmin = 600000, 5200000, 100.0
    r = np.random.default_rng(3 + topo)
    n = 3000
    ci, cj = r.integers(1, nii + 1, n), r.integers(1, njj + 1, n)
    tx, ty = r.uniform(0, 1, n), r.uniform(0, 1, n)
    x = ikoo + (ci - 1 + tx) * dx
    y = jkoo + (cj - 1 + ty) * dx
    h = r.uniform(0.01, float(hok[nkk]) - 0.01, n).astype(F)
    z = (h + F(ahmin)).astype(F) if topo else h
    ahint = np.zeros(n, F)
    ux, uy, uz, kk = W.wind_at_particles(uk, vk, wk, hok, dzk, nkk, dx, dx, ikoo, jkoo, ahmin, topo, ci, cj, ahint,
                                         x, y, z)
    # u = i - 1 + t exactly on the faces: the double offset keeps the error at
    # the float rounding of the result (a float offset would err by ~0.03 m / DX)
    assert np.max(np.abs(ux - (ci - 1 + tx))) < 2e-6 * (nii + 1)
    assert np.max(np.abs(uy + (cj - 1 + ty))) < 2e-6 * (njj + 1)
    hh = (z - F(ahmin)).astype(F) if topo else z
    assert np.allclose(uz, 0.5 * hh.astype(np.float64), rtol=2e-6, atol=2e-5)
    want_k = [_first_top_at_or_above(hok, nkk, v) for v in hh]
    assert np.array_equal(kk, want_k)


def test_flat_branch_adds_the_cell_ground_height():
    # flat terrain with buildings: the level comes from z, the w offset from
    # z + AHint (the reference's mix of the two)
    nii, njj, nkk = 6, 6, 20
    dzk, hok = FC.vertical_grid(nkk, 2.0, 1.04)
    uk, vk, wk = _fields(nii, njj, nkk, hok, c=1.0)
    z = np.array([3.0, 3.0, 7.5], F)
    ah = np.array([0.0, 2.0, 1.0], F)
    ux, uy, uz, kk = W.wind_at_particles(uk, vk, wk, hok, dzk, nkk, 2.0, 2.0, 0, 0, 0.0, 0, [2, 2, 3], [2, 2, 3], ah,
                                         [3.0, 3.0, 5.0], [3.0, 3.0, 5.0], z)
    assert kk[0] == kk[1]
    assert uz[1] - uz[0] == pytest.approx(2.0, rel=1e-5)
