"""Synthetic-domain generator: grid, masks and indices follow the reference's formulas."""
import math

import numpy as np

from gral_amd import domain


def test_vertical_grid_flat_config1():
    nkk, dzk, hok = domain.vertical_grid_flat(1.0, 1.01)
    # sum_{m<k} 1.01^m reaches 800 m at level 221 (ProgramFunctions.cs:69-73)
    assert nkk == 221
    assert hok[nkk] >= 800 and hok[nkk - 1] < 800
    assert dzk.dtype == np.float32 and hok[0] == 0 and dzk[1] == 1
    # closed form of the geometric series within float32 accumulation error
    assert abs(float(hok[100]) - (1.01 ** 100 - 1) / 0.01) < 1e-3


def test_config1_building_indices():
    c = domain.config1(iterations=1)
    # vertk(k) = HOKART[k-1] + DZK[k]/2 <= 15 m holds up to k = 14 (MicroscaleFlat.cs:94-96)
    assert int(c.kkart[50, 50]) == 14 and int(c.kkart[45, 50]) == 0 and int(c.kkart[46, 46]) == 14
    assert c.kadvmax == 44
    # factor 15 * 15 m / 2 m = 112 cells: the sub-domain covers the whole grid
    assert c.advdom[1:101, 1:101].all() and not c.advdom[0].any() and not c.advdom[101].any()
    assert c.refp == (1, 1)
    # 98 x 98 interior columns x 30 prognostic levels each
    assert c.active_cells() == 98 * 98 * 30
    f = 2 * 7.29e-5 * math.sin(47 * 3.1415 / 180)
    assert abs(c.coriolis - f) < 1e-9


def test_wind_direction_convention():
    u, v = domain.wind_components(3.0, 270.0)
    assert abs(u - 3.0) < 1e-6 and abs(v) < 1e-5
    u, v = domain.wind_components(3.0, 180.0)       # from the south blows to +y
    assert abs(u) < 1e-5 and abs(v - 3.0) < 1e-6


def test_reference_point_spiral():
    adv = np.zeros((12, 12), dtype=np.uint8)
    adv[5, 7] = 1
    assert domain.search_reference_point(adv) == (5, 7)
    adv[:] = 0
    assert domain.search_reference_point(adv) == (0, 0)


def test_initial_field_blocked_faces():
    c = domain.tiny_case(n=9)
    b = (3, 5, 4, 6)
    k = 2
    assert c.kkart[4, 5] == 4
    assert (c.uk[3:6, 4:7, 1:5] == 0).all()       # cells inside the building
    assert (c.uk[6, 4:7, 1:5] == 0).all()         # east faces of the last blocked cells
    assert (c.vk[3:6, 7, 1:5] == 0).all()         # north faces
    assert c.uk[7, 5, k] != 0
