"""Flat-terrain set-up of a situation (MicroscaleFlat.cs:29-345,
ReadGeometryAndDomains.cs:321-349, MicroscaleTerrainSearchRefPoint.cs:27-103),
CPU side: the oracle pinned by closed forms, hand-computed values and
brute-force properties; the host-only reference-point search through the C
ABI against the oracle.

The reference ships no fixtures for this path.  Each check below is a second
formulation of what the reference fixes (the blocked-level definition, the
face rule, the power law at its anchor height, the interpolation at its
nodes, the damping factor of one obstacle), so that a shifted index, a
swapped face or a wrong branch in the serial oracle fails one of them."""
import numpy as np
import pytest

from oracle import flat_oracle as O
from tests import flat_cases as FC

F = np.float32


def _run(nii, njj, nkk, seed, metkind="single", near_wall=0, dz1=2.0, stretch=1.05, vcf=6):
    dzk, hok = FC.vertical_grid(nkk, dz1, stretch)
    cutk = FC.city(nii, njj, seed)
    adv = O.subdomains(cutk, nii, njj, 2.0, 2.0, factor=2)
    m = FC.met(metkind, seed, nii, njj, near_wall=near_wall)
    z = np.zeros((nii + 2, njj + 2, nkk + 2), F)
    res = O.flat_init(nii, njj, nkk, vcf, 2.0, 2.0, dzk, hok, cutk, m, z, z, z, advdom=adv)
    return dzk, hok, cutk, adv, m, res


def test_subdomains_match_box_union():
    nii, njj = 23, 19
    cutk = FC.city(nii, njj, 3, density=0.05)
    for factor, dx, dy in ((2, 2.0, 2.0), (1, 3.0, 2.0), (15, 4.0, 4.0)):
        got = O.subdomains(cutk, nii, njj, dx, dy, factor)
        want = np.zeros_like(got)
        for i, j in zip(*np.nonzero(cutk[1:nii + 1, 1:njj + 1] > 0)):
            i, j = i + 1, j + 1
            ip = int(F(F(factor) * cutk[i, j]) / F(dx))
            jp = int(F(F(factor) * cutk[i, j]) / F(dy))
            want[max(1, i - ip):min(nii, i + ip) + 1, max(1, j - jp):min(njj, j + jp) + 1] = 1
        assert np.array_equal(got, want)
    # one 4 m column, factor 15, dx 2 m: (int)(60/2) = 30 cells each side, clipped
    c = np.zeros((12, 12), F)
    c[5, 6] = 4.0
    a = O.subdomains(c, 10, 10, 2.0, 2.0, 15)
    assert a[1:11, 1:11].all() and not a[0].any() and not a[:, 11].any()


def _ring(i, j, xmax, ymax):
    return min(i, j, xmax - 1 - i, ymax - 1 - j)


@pytest.mark.parametrize("seed", range(6))
def test_reference_point_properties(seed):
    r = np.random.default_rng(seed)
    adv = np.zeros((14, 11), np.uint8)
    for _ in range(r.integers(1, 4)):
        adv[r.integers(2, 12), r.integers(2, 9)] = 1
    i, j = O.reference_point(adv)
    assert adv[i, j] == 1
    ring = _ring(i, j, *adv.shape)
    outer = [(a, b) for a, b in zip(*np.nonzero(adv)) if _ring(a, b, *adv.shape) < ring]
    assert not outer


def test_reference_point_by_hand():
    adv = np.zeros((12, 12), np.uint8)
    adv[3, 5] = 1
    adv[7, 3] = 1                      # also on ring 3, but later on the spiral
    assert O.reference_point(adv) == (3, 5)
    assert O.reference_point(np.zeros((12, 12), np.uint8)) == (1, 1)


def test_reference_point_c_abi_matches_oracle():
    from gral_amd import pf
    for seed in range(20):
        r = np.random.default_rng(100 + seed)
        nii, njj = int(r.integers(5, 30)), int(r.integers(5, 30))
        adv = (r.random((nii + 2, njj + 2)) < 0.02).astype(np.uint8)
        want = O.reference_point(adv)
        got, ok = pf.reference_point(adv)
        assert got == want and ok == bool(adv[want] == 1 and adv.any())


@pytest.mark.parametrize("metkind", ["single", "profile3", "columns"])
def test_blocking_and_faces_closed_form(metkind):
    nii, njj, nkk = 14, 12, 18
    dzk, hok, cutk, adv, m, (uk, vk, wk, kk, ahk, bui, kadv) = _run(nii, njj, nkk, 5, metkind)
    vertk, blocked, kk2, ahk2 = FC.closed_form_blocking(nkk, dzk, hok, cutk)
    inner = (slice(1, nii + 1), slice(1, njj + 1))
    assert np.array_equal(kk[inner], kk2[inner])
    assert np.array_equal(ahk[inner], ahk2[inner])
    assert np.array_equal(bui[inner][kk2[inner] > 0], ahk2[inner][kk2[inner] > 0])
    big = cutk[inner] > 0
    assert kadv == min(nkk, max(1, int(kk2[inner][big].max()) if big.any() else 1) + 6)
    L = slice(1, nkk + 1)
    # u is zero where the cell or its western neighbour is blocked, v where the cell or its southern one is
    bu = blocked[1:nii + 1, 1:njj + 1, L].copy()
    bu[1:] |= blocked[1:nii, 1:njj + 1, L]
    bv = blocked[1:nii + 1, 1:njj + 1, L].copy()
    bv[:, 1:] |= blocked[1:nii + 1, 1:njj, L]
    assert not uk[1:nii + 1, 1:njj + 1, L][bu].any()
    assert not vk[1:nii + 1, 1:njj + 1, L][bv].any()
    if metkind != "columns":
        # away from the faces every column carries the same profile
        free_u = ~bu
        col = uk[1:nii + 1, 1:njj + 1, L]
        ref = np.where(free_u, col, np.nan)
        for k in range(nkk):
            vals = ref[..., k][~np.isnan(ref[..., k])]
            assert vals.size == 0 or np.all(vals == vals[0])
    # ghost copies and zeros
    assert np.array_equal(uk[nii + 1, 1:njj + 1, L], uk[nii, 1:njj + 1, L])
    assert np.array_equal(vk[1:nii + 1, njj + 1, L], vk[1:nii + 1, njj, L])
    assert not uk[1:, njj + 1, L].any() and not vk[nii + 1, 1:, L].any() and not wk.any()


def test_power_law_at_anchor_and_between():
    nii, njj, nkk = 6, 6, 8
    dzk = np.full(nkk + 2, F(4.0), F)
    hok = np.arange(nkk + 2, dtype=F) * F(4.0)       # vertk = 2, 6, 10, 14, ...
    cutk = np.zeros((nii + 2, njj + 2), F)
    z = np.zeros((nii + 2, njj + 2, nkk + 2), F)
    for ob in (-50.0, 0.0, 300.0):
        m = {"n_profile": 1, "meas_height": [0.0, 10.0], "obs_u": [0.0, 3.25], "obs_v": [0.0, -1.5],
             "ob_length": ob, "ol_column": None, "near_wall": 0}
        uk, vk, *_ = O.flat_init(nii, njj, nkk, 2, 2.0, 2.0, dzk, hok, cutk, m, z, z, z)
        assert uk[3, 3, 3] == F(3.25) and vk[3, 3, 3] == F(-1.5)        # vertk = 10: (10/10)^e = 1
        e = 0.56 * ob ** -0.15 if ob > 0 else max(0.35 - 0.4 * abs(ob) ** -0.15, 0.05) if ob != 0 else 0.05
        for k in (1, 2, 5, 8):
            vz = 4.0 * (k - 1) + 2.0
            assert uk[3, 3, k] == pytest.approx(3.25 * (vz / 10.0) ** e, rel=3e-7)
    # three measurement heights: linear between them, the top value above them
    m = {"n_profile": 3, "meas_height": [0.0, 6.0, 14.0, 26.0], "obs_u": [0.0, 1.0, 2.0, 4.0],
         "obs_v": [0.0, 0.0, 0.0, 0.0], "ob_length": 100.0, "ol_column": None, "near_wall": 0}
    uk, *_ = O.flat_init(nii, njj, nkk, 2, 2.0, 2.0, dzk, hok, cutk, m, z, z, z)
    assert uk[2, 2, 2] == F(1.0) and uk[2, 2, 4] == F(2.0)               # at 6 m and 14 m
    assert uk[2, 2, 3] == F(1.5)                                         # 10 m, half way
    assert uk[2, 2, 5] == pytest.approx(2.0 + 2.0 * 4.0 / 12.0, rel=1e-7)  # 18 m
    assert uk[2, 2, 8] == F(4.0)                                         # 30 m, above the top


def test_damping_of_one_obstacle():
    # one 8 m column; a free column g cells east of it sees it to the west only
    # (damped columns: IGEB+1..NII-IGEB = 11..40)
    nii = njj = 50
    nkk = 10
    dzk = np.full(nkk + 2, F(2.0), F)
    hok = np.arange(nkk + 2, dtype=F) * F(2.0)
    cutk = np.zeros((nii + 2, njj + 2), F)
    cutk[25, 25] = 8.0                                                  # blocks levels 1..4 (vertk 1, 3, 5, 7)
    m = {"n_profile": 1, "meas_height": [0.0, 10.0], "obs_u": [0.0, 2.0], "obs_v": [0.0, 1.0],
         "ob_length": 500.0, "ol_column": None, "near_wall": 2}
    z = np.zeros((nii + 2, njj + 2, nkk + 2), F)
    uk, vk, _, kk, ahk, _, _ = O.flat_init(nii, njj, nkk, 2, 2.0, 2.0, dzk, hok, cutk, m, z, z, z)
    m0 = dict(m, near_wall=0)
    u0, v0, *_ = O.flat_init(nii, njj, nkk, 2, 2.0, 2.0, dzk, hok, cutk, m0, z, z, z)
    assert ahk[25, 25] == F(8.0) and kk[25, 25] == 4
    for g in range(1, 11):
        f = 0.19 * np.log((g * 2.0 + 0.5) * 10)
        for k in (1, 4):                                                 # vertk 1 and 7 <= AHK 8
            assert uk[25 + g, 25, k] == pytest.approx(u0[25 + g, 25, k] * f, rel=3e-7)
            assert vk[25 - g, 25, k] == pytest.approx(v0[25 - g, 25, k] * f, rel=3e-7)
        assert uk[25 + g, 25, 5] == u0[25 + g, 25, 5]                     # vertk 9 > AHK: no obstacle
        d = np.hypot(g * 2.0, g * 2.0)
        fd = 0.19 * np.log((d + 0.5) * 10) if d <= 20 else 1.0
        assert uk[25 + g, 25 + g, 2] == pytest.approx(u0[25 + g, 25 + g, 2] * fd, rel=3e-7)
    assert uk[37, 25, 2] == u0[37, 25, 2]                                # 12 cells away: beyond IGEB = 10


def test_near_wall_outside_subdomains_only():
    nii, njj, nkk = 30, 28, 10
    dzk, hok, cutk, adv, m, (u1, *_r) = _run(nii, njj, nkk, 9, "single", near_wall=1)
    *_a, (u0, *_s) = _run(nii, njj, nkk, 9, "single", near_wall=0)
    *_b, (u2, *_t) = _run(nii, njj, nkk, 9, "single", near_wall=2)
    inside = adv.astype(bool)
    assert np.array_equal(u1[inside], u0[inside])
    assert np.array_equal(u1[~inside], u2[~inside])
    assert not np.array_equal(u2, u0)
