"""GPU parity of the flat-terrain set-up (gral_pf_subdomains,
gral_pf_flat_init) against the serial oracle, bit for bit, on small random
cities; at the 400 x 400 city size through the closed forms (box union of
the sub-domains, blocked levels, faces, ghost copies); and a situation set
up on the GPU and then solved, against the prognostic oracle."""
import numpy as np
import pytest

from gral_amd import domain
from oracle import flat_oracle as O
from tests import flat_cases as FC

F = np.float32


def _garbage(shape, seed):
    return np.random.default_rng(seed).uniform(-9, 9, size=shape).astype(F)


@pytest.mark.gpu
@pytest.mark.parametrize("factor,dx,dy", [(2, 2.0, 2.0), (1, 3.0, 2.0), (15, 4.0, 5.0)])
def test_subdomains_small(factor, dx, dy):
    from gral_amd import pf
    for seed in range(3):
        nii, njj = 31 + seed, 27 - seed
        cutk = FC.city(nii, njj, seed, density=0.04)
        pre = (np.random.default_rng(seed).random((nii + 2, njj + 2)) < 0.05).astype(np.uint8)
        print(f"[setup] subdomains {nii}x{njj} factor={factor} dx={dx} dy={dy} seed={seed}", flush=True)
        want = O.subdomains(cutk, nii, njj, dx, dy, factor, advdom=pre)
        got = pf.subdomains(cutk, nii, njj, dx, dy, factor, advdom=pre)
        assert np.array_equal(got, want)


@pytest.mark.gpu
@pytest.mark.parametrize("metkind", ["single", "profile3", "columns"])
@pytest.mark.parametrize("near_wall", [0, 1, 2])
def test_flat_init_small(metkind, near_wall):
    from gral_amd import pf
    nii, njj, nkk = 34, 30, 22
    dzk, hok = FC.vertical_grid(nkk, 1.5, 1.06)
    cutk = FC.city(nii, njj, 11, density=0.1)
    adv = O.subdomains(cutk, nii, njj, 2.0, 2.0, factor=1)
    m = FC.met(metkind, 3, nii, njj, near_wall=near_wall)
    shp3, shp2 = (nii + 2, njj + 2, nkk + 2), (nii + 2, njj + 2)
    u0, v0, w0 = _garbage(shp3, 1), _garbage(shp3, 2), _garbage(shp3, 3)
    kk0 = np.random.default_rng(4).integers(0, 9, size=shp2).astype(np.int16)
    ah0, bu0 = _garbage(shp2, 5), _garbage(shp2, 6)
    print(f"[setup] flat_init {nii}x{njj}x{nkk} met={metkind} near_wall={near_wall}", flush=True)
    want = O.flat_init(nii, njj, nkk, 7, 2.0, 2.0, dzk, hok, cutk, m, u0, v0, w0, advdom=adv, ahk=ah0, kkart=kk0,
                       bui=bu0)
    got = pf.flat_init(nii, njj, nkk, 7, 2.0, 2.0, dzk, hok, cutk, m, u0, v0, w0, advdom=adv, kkart=kk0, ahk=ah0,
                       bui_height=bu0)
    for name, a, b in zip(("uk", "vk", "wk", "kkart", "ahk", "bui_height"), want[:6],
                          (got["uk"], got["vk"], got["wk"], got["kkart"], got["ahk"], got["bui_height"])):
        bad = np.argwhere(a.view(np.uint32 if a.dtype == F else np.uint16) !=
                          b.view(np.uint32 if b.dtype == F else np.uint16))
        assert bad.size == 0, (name, bad[:5].tolist())
    assert got["kadvmax"] == want[6]


@pytest.mark.gpu
def test_city400_closed_form():
    from gral_amd import pf
    n = 400
    nkk, dzk, hok = domain.vertical_grid_flat()
    cutk = np.zeros((n + 2, n + 2), F)
    for b in domain.city_buildings(n, n, 2.0, 20260001):
        cutk[b.i0:b.i1 + 1, b.j0:b.j1 + 1] = np.maximum(cutk[b.i0:b.i1 + 1, b.j0:b.j1 + 1], F(b.height))
    print("[setup] city 400 sub-domains and initial field", flush=True)
    adv = pf.subdomains(cutk, n, n, 2.0, 2.0, 15)
    want = np.zeros_like(adv)
    for i, j in zip(*np.nonzero(cutk > 0)):
        r = int(F(F(15) * cutk[i, j]) / F(2.0))
        want[max(1, i - r):min(n, i + r) + 1, max(1, j - r):min(n, j + r) + 1] = 1
    assert np.array_equal(adv, want)
    u, v = domain.wind_from_met(3.7, 230.0)
    m = {"n_profile": 1, "meas_height": [0.0, 10.0], "obs_u": [0.0, u], "obs_v": [0.0, v], "ob_length": -60.0,
         "ol_column": None, "near_wall": 0}
    r = pf.flat_init(n, n, nkk, 40, 2.0, 2.0, dzk, hok, cutk, m, advdom=adv)
    vertk, blocked, kk2, ahk2 = FC.closed_form_blocking(nkk, dzk, hok, cutk)
    inner = (slice(1, n + 1), slice(1, n + 1))
    assert np.array_equal(r["kkart"][inner], kk2[inner]) and np.array_equal(r["ahk"][inner], ahk2[inner])
    assert r["kadvmax"] == min(nkk, int(kk2[inner][cutk[inner] > 0].max()) + 40)
    L = slice(1, nkk + 1)
    bu = blocked[1:n + 1, 1:n + 1, L].copy()
    bu[1:] |= blocked[1:n, 1:n + 1, L]
    uk = r["uk"][1:n + 1, 1:n + 1, L]
    assert not uk[bu].any()
    # every free cell of a level carries that level's profile value (one Obukhov length)
    top = uk[~bu.any(-1)]
    assert top.shape[0] > 0 and np.all(top == top[0])
    assert np.all((uk == top[0][None, None, :]) | bu)
    assert np.array_equal(r["uk"][n + 1, 1:n + 1, L], r["uk"][n, 1:n + 1, L])


@pytest.mark.gpu
def test_situation_set_up_on_gpu_then_solved():
    # the set-up feeds the solver: a small situation of the batched workload
    # generator, solved a few iterations, against the prognostic oracle
    from tests.test_gpu_parity import _compare
    nii = njj = 40
    cutk = FC.city(nii, njj, 21, density=0.08, hmax=9.0)
    c = domain.flat_case("setup40", nii, njj, cutk, vertcellsff=12, speed=2.5, direction=200.0, stability=2,
                         iterations=3)
    print(f"[setup] solve the GPU-set-up situation refp={c.refp} kadvmax={c.kadvmax}", flush=True)
    _compare(c, 3, sweeps=("strip",))
