"""Pins the C oracle against an independent numpy transcription of the
reference (tests/np_transcription.py): both must agree bit for bit on tiny
domains, in every turbulence model, with and without vegetation, over enough
iterations to pass through all four sweep phases (IterationLoops % 4)."""
import numpy as np
import pytest

from gral_amd import domain
from oracle import pyoracle
from tests import np_transcription

FIELDS = ("uk", "vk", "wk", "uks", "vks", "wks", "dpm", "dpmnew")


def _agree(case, iterations):
    ref = pyoracle.run(case, iterations=iterations)
    npy = np_transcription.run(case, iterations)
    fields = FIELDS + (("turb", "tdiss") if case.turbulence_model != 1 else ())
    for f in fields:
        a, b = ref[f], npy[f]
        bad = np.argwhere(a.view(np.uint32) != b.view(np.uint32))
        assert bad.size == 0, (f, bad[:5].tolist(), float(np.abs(a.astype(np.float64) - b).max()))
    assert np.array_equal(ref["stats"]["delta_umax_hist"], npy["delta_umax_hist"])


@pytest.mark.parametrize("model", [1, 2, 0])
def test_tiny_building_all_phases(model):
    _agree(domain.tiny_case(n=8, vertcellsff=8, turbulence_model=model, iterations=5), 5)


@pytest.mark.parametrize("model", [1, 2])
def test_tiny_vegetation(model):
    _agree(domain.tiny_case(n=8, vertcellsff=9, turbulence_model=model, iterations=3, veg=True), 3)


def test_two_buildings_tall_columns():
    # two buildings of different height in one row: building columns start at
    # k=1, the SIMD blocks straddle KKART+1 and the stale PIMP top reaches the
    # lower columns of the row
    bl = [domain.Building(3, 4, 3, 5, 6.5), domain.Building(6, 7, 4, 6, 3.2)]
    c = domain.build_case("two", 10, 10, bl, vertcellsff=10, direction=235.0, iterations=4)
    _agree(c, 4)


@pytest.mark.parametrize("level,model,after,terrain", [(2, 1, 3, False), (1, 1, 0, False), (2, 2, 2, False),
                                                     (2, 1, 2, True)])
def test_diagnostic_pass(level, model, after, terrain):
    # DiagnosticFlowfield + final W adjustment: on the prognostic result (level 2,
    # all four sweep orders over 10 iterations) and on the initial field (level 1)
    bl = [domain.Building(3, 4, 3, 5, 6.5), domain.Building(6, 7, 4, 6, 3.2)]
    c = domain.build_case("two", 10, 10, bl, vertcellsff=10, direction=235.0, turbulence_model=model, iterations=3)
    if after:
        r = pyoracle.run(c, iterations=after)
        f = {k: r[k] for k in ("uk", "vk", "wk", "dpm")}
    else:
        f = {"uk": c.uk, "vk": c.vk, "wk": c.wk, "dpm": np.zeros_like(c.uk)}
    if terrain:
        c.dyk = np.float32(2.5)                      # the terrain variant's DYK on the j faces must show
    ref = pyoracle.diagnostic(c, f, flow_field_level=level, terrain=terrain)
    npy = np_transcription.diagnostic(c, f, level, terrain=terrain)
    assert ref["dtime"] == npy["dtime"]
    for k in ("uk", "vk", "wk", "dpm"):
        bad = np.argwhere(ref[k].view(np.uint32) != npy[k].view(np.uint32))
        assert bad.size == 0, (k, bad[:5].tolist())
