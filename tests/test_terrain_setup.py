"""Terrain set-up from a GRAMM field (MicroscaleTerrain.cs:57-316), CPU side:
the serial oracle pinned by closed forms.  The reference ships no fixtures
for this path; the checks below restate what the reference fixes in another
form: the blocked levels (GRAL level centre at or below the interpolated
ground plus CUTK), the inverse-distance weights summing to one, the vertical
interpolation of a horizontally uniform GRAMM wind, the face and ghost rules,
and the 10 m mean-wind count."""
import numpy as np
import pytest

from oracle import terrain_oracle as T
from tests import flat_cases as FC
from tests import terrain_cases as TC

F = np.float32


def _setup(seed=0, nii=40, njj=36, nkk=24, uniform_wind=False, delta=(25, 15), near_wall=0, hill=30.0, city=True):
    dzk, hok = FC.vertical_grid(nkk, 2.0, 1.04)
    gr = TC.gramm(seed=seed, uniform_wind=uniform_wind, delta=delta, hill=hill)
    cutk = FC.city(nii, njj, seed + 1, density=0.05, hmax=8.0) if city else np.zeros((nii + 2, njj + 2), F)
    adv = np.zeros((nii + 2, njj + 2), np.uint8)
    adv[8:30, 6:28] = 1
    shp3, shp2 = (nii + 2, njj + 2, nkk + 2), (nii + 2, njj + 2)
    z3 = np.zeros(shp3, F)
    res = T.terrain_init(nii, njj, nkk, 6, 2.0, 2.0, dzk, hok, cutk, adv, gr, z3, z3, z3, np.zeros(shp2, np.int16),
                         np.zeros(shp2, F), np.zeros(shp2, F), F(100.0), near_wall=near_wall)
    return dzk, hok, gr, cutk, adv, res


def test_weights_are_inverse_distance_to_the_16_points():
    r = T.weights(F(7.5), F(3.25), F(20), F(20))
    # the 16 points of the 4 x 4 GRAMM stencil around the cell, from the offsets
    pts = [(x, y) for (x, y) in [(7.5, 3.25), (7.5, 16.75), (12.5, 16.75), (12.5, 3.25), (27.5, 23.25), (7.5, 23.25),
                                 (12.5, 23.25), (32.5, 23.25), (32.5, 3.25), (32.5, 16.75), (32.5, 36.75),
                                 (12.5, 36.75), (7.5, 36.75), (27.5, 36.75), (27.5, 16.75), (27.5, 3.25)]]
    want = [1.0 / (x * x + y * y + 0.1) ** 2 for x, y in pts]
    assert np.allclose(r, want, rtol=1e-6)
    assert len(T.NEIGH) == 16 and len(set(T.NEIGH)) == 16


def test_flat_gramm_blocking_closed_form():
    # no hill: every GRAMM level lies at the same height, so a GRAL level is
    # blocked exactly where its centre is at or below 100 m + CUTK
    dzk, hok, gr, cutk, adv, (uk, vk, wk, kk, ahk, bui, ahmin, kadv, vec) = _setup(hill=0.0)
    nii, njj, nkk = 40, 36, 24
    vertk = (hok[0:nkk] + dzk[1:nkk + 1] * 0.5).astype(np.float64)
    for i in range(1, nii + 1):
        for j in range(1, njj + 1):
            want = int((vertk <= cutk[i, j] + 1e-4).sum())
            assert kk[i, j] == want, (i, j)
            assert ahk[i, j] == (F(hok[want] + F(100)) if want else F(100))
            assert not uk[i, j, 1:want + 1].any()
            if want:
                assert not wk[i, j, 1:want + 2].any()          # w of the first free level too
    assert ahmin == F(100) + (F(hok[kk[1:nii + 1, 1:njj + 1].min()]) if kk[1:, 1:].min() else 0)
    assert kadv == min(nkk, max(1, int(kk[adv > 0].max())) + 6)


def test_uniform_gramm_wind_vertical_interpolation():
    # a horizontally uniform GRAMM wind: u is the linear interpolation between
    # the two GRAMM levels around the GRAL level centre
    dzk, hok, gr, cutk, adv, (uk, vk, wk, kk, *_r) = _setup(hill=0.0, uniform_wind=True, city=False)
    lev = gr["ahe"][3, 3, :] - F(100)
    prof = gr["uwin"][0, 0, :]
    for k in range(1, 25):
        z = float(hok[k - 1]) + float(dzk[k]) * 0.5
        ik = max(q for q in range(1, gr["nz"] + 1) if z > lev[q])
        ikp = min(ik + 1, gr["nz"])
        want = prof[ik] + (prof[ikp] - prof[ik]) / max(lev[ikp] - lev[ik], 0.1) * (z - lev[ik])
        assert uk[20, 20, k] == pytest.approx(want, rel=2e-6, abs=1e-6)
        assert vk[20, 20, k] == pytest.approx(-0.5 * want, rel=2e-6, abs=1e-6)


def test_faces_ghosts_and_mean_wind_count():
    nii, njj = 40, 36
    dzk, hok, gr, cutk, adv, (uk, vk, wk, kk, ahk, bui, ahmin, kadv, vec) = _setup(seed=3, delta=(5, 3))
    for i in range(2, nii + 1):
        for j in range(1, njj + 1):
            assert not uk[i, j, 1:kk[i - 1, j] + 1].any()
    for i in range(1, nii + 1):
        for j in range(2, njj + 1):
            assert not vk[i, j, 1:kk[i, j - 1] + 1].any()
    free = np.arange(1, 25)[None, :] > kk[nii, 1:njj + 1, None]
    assert np.array_equal(uk[nii + 1, 1:njj + 1, 1:25][free], uk[nii, 1:njj + 1, 1:25][free])
    # every column has free levels whose GRAMM level is within 10 m of the ground
    assert vec[2] == nii * njj
    assert bui[cutk > 0].max() > 0 and not bui[cutk <= 0].any()


# --- the variant with the original GRAL topography (MicroscaleTerrain.cs:318-598)

def _topo_setup(seed=0, nii=40, njj=36, nkk=24, delta=(25, 15), hill=30.0, uniform_wind=False, city=True,
                ahko=None, pre=None):
    dzk, hok = FC.vertical_grid(nkk, 2.0, 1.04)
    gr = TC.gramm(seed=seed, uniform_wind=uniform_wind, delta=delta, hill=hill)
    cutk = FC.city(nii, njj, seed + 1, density=0.05, hmax=8.0) if city else np.zeros((nii + 2, njj + 2), F)
    adv = np.zeros((nii + 2, njj + 2), np.uint8)
    adv[8:30, 6:28] = 1
    if ahko is None:
        rr = np.random.default_rng(seed + 5)
        ahko = (F(100.0) + rr.uniform(0.0, 12.0, (nii + 2, njj + 2))).astype(F)
    shp3, shp2 = (nii + 2, njj + 2, nkk + 2), (nii + 2, njj + 2)
    z3 = np.zeros(shp3, F)
    u0, v0, w0 = (pre if pre is not None else (z3, z3, z3))
    aom = F(ahko[1:nii + 1, 1:njj + 1].min())
    res = T.terrain_init_topo(nii, njj, nkk, 6, 2.0, 2.0, dzk, hok, cutk, adv, gr, u0, v0, w0,
                              np.zeros(shp2, np.int16), np.zeros(shp2, F), np.zeros(shp2, F), F(100.0), ahko, aom)
    return dzk, hok, gr, cutk, adv, ahko, aom, res


def _vertk_f(hok, dzk, aom, k):
    return F(F(hok[k - 1] + F(dzk[k] * F(0.5))) + aom)


def test_topo_blocking_closed_form():
    # a level is blocked exactly when its (float) centre is at or below AHKOri + CUTK
    dzk, hok, gr, cutk, adv, ahko, aom, (uk, vk, wk, kk, ahk, bui, ahmin, kadv, vec) = _topo_setup()
    for i in range(1, 41):
        for j in range(1, 37):
            want = sum(1 for k in range(1, 25) if _vertk_f(hok, dzk, aom, k) <= F(ahko[i, j] + cutk[i, j]))
            assert kk[i, j] == want, (i, j)
            assert ahk[i, j] == (F(hok[want] + aom) if want else aom)
            assert not uk[i, j, 1:want + 1].any()
            if want:
                assert not wk[i, j, 1:want + 2].any()
            # building height: AHK minus the bottom of the lowest blocked level at or above AHKOri
            ks = [k for k in range(2, want + 1) if _vertk_f(hok, dzk, aom, k) >= ahko[i, j]]
            if cutk[i, j] > 0 and ks:
                assert bui[i, j] == F(ahk[i, j] - F(hok[min(ks) - 1] + aom))
            else:
                assert bui[i, j] == 0
    assert kadv == min(24, max(1, int(kk[adv > 0].max())) + 6)


def test_topo_border_gramm_heights_stay_relative():
    # at the GRAMM border the heights stay relative to the incoming AHMIN
    # while the level centre is absolute (the reference's mix): with a
    # horizontally uniform GRAMM wind, u is linear between the GRAMM levels
    # around the absolute centre on the relative heights, which differs from
    # what absolute heights would give
    dzk, hok, gr, cutk, adv, ahko, aom, (uk, vk, wk, kk, *_r) = _topo_setup(delta=(5, 3), uniform_wind=True, city=False)
    prof = gr["uwin"][0, 0, :]
    nz = gr["nz"]

    def lin(z, h):
        ik = max((q for q in range(1, nz + 1) if z > h[q]), default=None)
        if ik is None:
            return None
        ikp = min(ik + 1, nz)
        return prof[ik] + (prof[ikp] - prof[ik]) / max(h[ikp] - h[ik], 0.1) * (z - h[ik])

    nb = ndiff = 0
    for i in range(1, 41):
        for j in range(1, 37):
            indi = int(F(F(F(F(5) + F(F(2) * F(i))) - F(1)) / F(20))) + 1
            indj = int(F(F(F(F(3) + F(F(2) * F(j))) - F(1)) / F(20))) + 1
            if 2 <= indi <= gr["nx"] - 1 and 2 <= indj <= gr["ny"] - 1:
                continue
            col = gr["ahe"][min(max(indi, 1), gr["nx"]), min(max(indj, 1), gr["ny"])]
            rel = (col - F(100.0)).astype(F)
            for k in range(int(kk[i, j]) + 1, 25):
                if i > 1 and k <= kk[i - 1, j]:
                    continue
                z = float(_vertk_f(hok, dzk, aom, k))
                want = lin(z, rel)
                if want is None:
                    continue
                assert uk[i, j, k] == pytest.approx(want, rel=2e-6, abs=1e-6), (i, j, k)
                nb += 1
                alt = lin(z, col)
                ndiff += alt is None or abs(alt - want) > 1e-3
    assert nb > 100 and ndiff > 10


def test_topo_interior_vertical_interpolation():
    # flat GRAMM terrain (AHE = 100 m + level height everywhere) and a uniform
    # wind: in the interior u is linear between the GRAMM levels around the
    # absolute level centre
    dzk, hok, gr, cutk, adv, ahko, aom, (uk, *_r) = _topo_setup(hill=0.0, uniform_wind=True, city=False,
                                                                 ahko=np.full((42, 38), F(100.0), F))
    lev = gr["ahe"][3, 3, :]
    prof = gr["uwin"][0, 0, :]
    for k in range(1, 25):
        z = float(_vertk_f(hok, dzk, aom, k))
        ik = max(q for q in range(1, gr["nz"] + 1) if z > lev[q])
        ikp = min(ik + 1, gr["nz"])
        want = prof[ik] + (prof[ikp] - prof[ik]) / max(lev[ikp] - lev[ik], 0.1) * (z - lev[ik])
        assert uk[20, 20, k] == pytest.approx(want, rel=2e-6, abs=1e-6)


def test_topo_free_level_below_every_gramm_level():
    # a GRAL topography 20 m below the GRAMM surface: free levels under the
    # lowest GRAMM level take GRAMM level 1 (no vertical offset)
    dzk, hok, gr, cutk, adv, ahko, aom, (uk, *_r) = _topo_setup(hill=0.0, uniform_wind=True, city=False,
                                                                 ahko=np.full((42, 38), F(80.0), F))
    nlow = 0
    for k in range(1, 25):
        if _vertk_f(hok, dzk, aom, k) <= F(100.0):
            assert uk[20, 20, k] == gr["uwin"][0, 0, 1]
            nlow += 1
    assert nlow >= 3


def test_topo_ghosts_and_mean_wind_count():
    # the ghosts are copied while column 1 (row 1) is set: they take the value
    # row NII (column NJJ) had before the set-up at the levels where column 1
    # is free, and keep their own elsewhere
    nii, njj, nkk = 40, 36, 24
    rr = np.random.default_rng(3)
    pre = tuple(rr.uniform(-3, 3, (nii + 2, njj + 2, nkk + 2)).astype(F) for _ in range(3))
    dzk, hok, gr, cutk, adv, ahko, aom, (uk, vk, wk, kk, ahk, bui, ahmin, kadv, vec) = _topo_setup(seed=3, pre=pre)
    for j in range(1, njj + 1):
        for k in range(1, nkk + 1):
            want = pre[0][nii, j, k] if k > kk[1, j] else pre[0][nii + 1, j, k]
            assert uk[nii + 1, j, k] == want, (j, k)
    for i in range(1, nii + 1):
        for k in range(1, nkk + 1):
            want = pre[1][i, njj, k] if k > kk[i, 1] else pre[1][i, njj + 1, k]
            assert vk[i, njj + 1, k] == want, (i, k)
    n = 0
    for i in range(1, nii + 1):
        for j in range(1, njj + 1):
            lim = [F(F(ahko[i, j] + F(10)) + F(F(0.5) * dzk[k])) for k in range(1, nkk + 1)]
            n += any(k > kk[i, j] and _vertk_f(hok, dzk, aom, k) < lim[k - 1] for k in range(1, nkk + 1))
    assert vec[2] == n and n > 0
