"""GPU parity of the terrain set-up (gral_pf_terrain_init) against the serial
oracle, bit for bit: interior and border GRAMM cells, hills, the three
near-wall modes, and input values outside the written set preserved; then a
situation set up this way, solved a few iterations, against the prognostic
oracle."""
import numpy as np
import pytest

from gral_amd import domain
from oracle import terrain_oracle as T
from tests import flat_cases as FC
from tests import terrain_cases as TC

F = np.float32


def _inputs(seed, nii, njj, nkk, delta, hill):
    dzk, hok = FC.vertical_grid(nkk, 2.0, 1.04)
    gr = TC.gramm(seed=seed, delta=delta, hill=hill)
    cutk = FC.city(nii, njj, seed + 1, density=0.06, hmax=9.0)
    adv = np.zeros((nii + 2, njj + 2), np.uint8)
    adv[6:nii - 4, 5:njj - 5] = 1
    r = np.random.default_rng(seed + 7)
    shp3, shp2 = (nii + 2, njj + 2, nkk + 2), (nii + 2, njj + 2)
    pre = {"uk": r.uniform(-5, 5, shp3).astype(F), "vk": r.uniform(-5, 5, shp3).astype(F),
           "wk": r.uniform(-1, 1, shp3).astype(F), "kkart": r.integers(0, 5, shp2).astype(np.int16),
           "ahk": r.uniform(0, 50, shp2).astype(F), "bui": r.uniform(0, 3, shp2).astype(F)}
    return dzk, hok, gr, cutk, adv, pre


@pytest.mark.gpu
@pytest.mark.parametrize("delta,hill,near_wall", [((25, 15), 30.0, 0), ((5, 3), 45.0, 1), ((25, 15), 0.0, 2),
                                                   ((61, 47), 30.0, 0)])
def test_terrain_init_matches_oracle(delta, hill, near_wall):
    from gral_amd import pf
    nii, njj, nkk = 44, 38, 26
    dzk, hok, gr, cutk, adv, pre = _inputs(5, nii, njj, nkk, delta, hill)
    print(f"[terrain] {nii}x{njj}x{nkk} delta={delta} hill={hill} near_wall={near_wall}", flush=True)
    want = T.terrain_init(nii, njj, nkk, 6, 2.0, 2.0, dzk, hok, cutk, adv, gr, pre["uk"], pre["vk"], pre["wk"],
                          pre["kkart"], pre["ahk"], pre["bui"], F(100.0), near_wall=near_wall)
    got = pf.terrain_init(nii, njj, nkk, 6, 2.0, 2.0, dzk, hok, cutk, adv, gr, F(100.0), pre["uk"], pre["vk"],
                          pre["wk"], pre["kkart"], pre["ahk"], pre["bui"], near_wall=near_wall)
    for name, a, b in zip(("uk", "vk", "wk", "kkart", "ahk", "bui_height"), want[:6],
                          (got["uk"], got["vk"], got["wk"], got["kkart"], got["ahk"], got["bui_height"])):
        view = np.uint32 if a.dtype == F else np.uint16
        bad = np.argwhere(a.view(view) != b.view(view))
        assert bad.size == 0, (name, bad[:5].tolist())
    assert got["ahmin"] == want[6] and got["kadvmax"] == want[7]
    assert got["wind10"][2] == want[8][2] and got["wind10"][0] == want[8][0] and got["wind10"][1] == want[8][1]


@pytest.mark.gpu
def test_terrain_situation_solved():
    from gral_amd import pf
    from tests.test_gpu_parity import _compare
    nii, njj, nkk = 30, 28, 40
    dzk, hok = FC.vertical_grid(nkk, 1.0, 1.05)
    gr = TC.gramm(seed=2, delta=(20, 10), hill=12.0)
    cutk = FC.city(nii, njj, 9, density=0.08, hmax=6.0)
    adv = pf.subdomains(cutk, nii, njj, 2.0, 2.0, 3)
    shp3, shp2 = (nii + 2, njj + 2, nkk + 2), (nii + 2, njj + 2)
    z = np.zeros(shp3, F)
    r = pf.terrain_init(nii, njj, nkk, 12, 2.0, 2.0, dzk, hok, cutk, adv, gr, F(100.0), z, z, z,
                        np.zeros(shp2, np.int16), np.zeros(shp2, F), np.zeros(shp2, F))
    refp, ok = pf.reference_point(adv)
    assert ok
    z0 = np.full(shp2, F(0.3), F)
    c = domain.Case(name="terrain30", nii=nii, njj=njj, nkk=nkk, kadvmax=r["kadvmax"], vertcellsff=12, dxk=2.0,
                    dyk=2.0, dzk=dzk, hokart=hok, advdom=adv, kkart=r["kkart"], cutk=cutk,
                    cov=np.zeros(shp2, F), ahk=r["ahk"], z0_wall=z0, z0_solver=z0.copy(),
                    ob_init=np.full(shp2, F(500.0), F), ust_init=np.full(shp2, F(0.3), F), z0_init=z0.copy(),
                    uk=r["uk"], vk=r["vk"], wk=r["wk"], refp=refp, iter_min=3, iter_max=3)
    print(f"[terrain] solve the GPU-set-up terrain situation kadvmax={c.kadvmax} refp={refp}", flush=True)
    _compare(c, 3, sweeps=("strip",))


@pytest.mark.gpu
@pytest.mark.parametrize("delta,hill,near_wall,low", [((25, 15), 30.0, 0, False), ((5, 3), 45.0, 1, False),
                                                      ((25, 15), 0.0, 2, True), ((61, 47), 30.0, 0, True)])
def test_terrain_init_topo_matches_oracle(delta, hill, near_wall, low):
    # the original-GRAL-topography variant; low: a GRAL topography partly below
    # the GRAMM surface, so that free levels lie under every GRAMM level
    from gral_amd import pf
    nii, njj, nkk = 44, 38, 26
    dzk, hok, gr, cutk, adv, pre = _inputs(5, nii, njj, nkk, delta, hill)
    r = np.random.default_rng(17)
    ahko = (F(100.0) + r.uniform(-25.0 if low else 0.0, 15.0, (nii + 2, njj + 2))).astype(F)
    aom = F(ahko[1:nii + 1, 1:njj + 1].min())
    print(f"[terrain-topo] {nii}x{njj}x{nkk} delta={delta} hill={hill} near_wall={near_wall} low={low}", flush=True)
    want = T.terrain_init_topo(nii, njj, nkk, 6, 2.0, 2.0, dzk, hok, cutk, adv, gr, pre["uk"], pre["vk"], pre["wk"],
                               pre["kkart"], pre["ahk"], pre["bui"], F(100.0), ahko, aom, near_wall=near_wall)
    got = pf.terrain_init_topo(nii, njj, nkk, 6, 2.0, 2.0, dzk, hok, cutk, adv, gr, F(100.0), ahko, aom, pre["uk"],
                               pre["vk"], pre["wk"], pre["kkart"], pre["ahk"], pre["bui"], near_wall=near_wall)
    for name, a, b in zip(("uk", "vk", "wk", "kkart", "ahk", "bui_height"), want[:6],
                          (got["uk"], got["vk"], got["wk"], got["kkart"], got["ahk"], got["bui_height"])):
        view = np.uint32 if a.dtype == F else np.uint16
        bad = np.argwhere(a.view(view) != b.view(view))
        assert bad.size == 0, (name, bad[:5].tolist())
    assert got["ahmin"] == want[6] and got["kadvmax"] == want[7]
    assert got["wind10"][2] == want[8][2] and got["wind10"][0] == want[8][0] and got["wind10"][1] == want[8][1]
    assert want[8][2] > 0

