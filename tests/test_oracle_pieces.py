"""Closed-form checks of the oracle's momentum and TKE assembly, its
divergence, cell averages and sub-domain boundary rule.

The oracle's checking hook records, per level of one column in one sweep, the
pieces a mixing-length (or k-eps TKE) equation is assembled from: eddy
viscosity, conductances, face fluxes, power-law coefficients and the source
terms (pressure, Coriolis/geostrophic, wall drag, cross terms).  The inputs
here are constructed so that each piece has a closed form that follows from
the physics the reference discretises, not from its code:

* a horizontally uniform shear flow u = a z, v = b z over flat ground on a
  uniform 1 m grid, with values on the solver's quantisation grid (1/Frund)
  so that the pressure-correction step leaves them unchanged: the mixing
  length is l = 0.071 z (z the cell-centre height above ground), reduced by
  (1 - 0.99 COV) in a canopy, so the eddy viscosity is l^2 |S| with the
  reference's shear measure |S| = sqrt((a^2 + b^2)/2), capped at 15 m^2/s;
* faces carry the advecting velocity at the face: u dy dz through the x
  faces, v dx dz through the y faces, and for the w equation the u of the
  w-point height; conductances are Gamma A / delta, floored at 1e-4;
* the power-law scheme (Patankar): a_nb = D A(|F/D|) + max(-+F, 0) with
  A(P) = max(0, (1 - 0.1 P)^5), which implies a_W - a_E = F for equal
  west and east faces whatever A is;
* the wall drag in the first cell above ground follows the log law:
  u*^2 = (kappa |U| / ln(z1/z0))^2 at z1 = dz/2, acting along -u/|U|;
* the geostrophic term is f (UG - u) dV; TKE shear production is
  nu_t dV (2 u_x^2 + ... + (u_z + w_x)^2 + (v_z + w_y)^2 - eps) = nu_t dV
  (a^2 + b^2 - eps) for this flow, buoyancy -dtheta/dz g/283 1.35 nu_t dV
  (negative for stable stratification);
* DIV of a linear field u = c x is -c dx dy dz per cell; the cell averages
  of a divergence-free linear field are its values at the cell centres;
* the sub-domain boundary copies a face velocity from outside only for
  inflow (u >= 0 on the west side, u <= 0 on the east side).

Each test also evaluates a plausible misreading (a dropped factor, the wrong
height, a swapped sign) and checks that it would fail.
"""
import math

import numpy as np
import pytest

from gral_amd import domain
from oracle import pyoracle

F = np.float32
N = 12
LIB_PATH = None                    # another build of the oracle (the mutant builds of test_oracle_mutants.py)
I0, J0 = 8, 8                      # the captured column (interior, ground level 0)


def shear_case(a, b, model=1, z0=0.05, cov=0.0, relax=0.0, potdtdz=0.0):
    """u = a z, v = b z below 10 m, u = 3.2, v = 0 above (so topwind = 3.2 and
    Frund = 2500 for model 1, 3125 for model 2: the values below lie on the
    quantisation grid when a z and b z are multiples of 0.0004 and 0.00032).
    relax = 0 keeps the momentum updates from changing the fields, so the V, W
    and TKE sweeps of the iteration see the same flow as the U sweep."""
    c = domain.build_case("shear", N, N, [domain.Building(3, 3, 3, 3, 0.4)], dx=2.0, dz1=1.0, stretch=1.0,
                          vertcellsff=12, z0=z0, turbulence_model=model, iterations=1)
    c.advdom[1:N + 1, 1:N + 1] = 1
    assert c.kkart.max() == 0
    nkk = c.nkk
    zc = (c.hokart[0:nkk] + F(0.5) * c.dzk[1:nkk + 1]).astype(F)     # centre heights of levels 1..nkk
    u = np.where(zc < 10, F(a) * zc, F(3.2)).astype(F)
    v = np.where(zc < 10, F(b) * zc, F(0)).astype(F)
    c.uk[:] = 0
    c.vk[:] = 0
    c.wk[:] = 0
    c.uk[1:N + 2, 1:N + 2, 1:nkk + 1] = u
    c.vk[1:N + 2, 1:N + 2, 1:nkk + 1] = v
    c.cov[I0, J0] = F(cov)
    c.z0_wall[:] = F(z0)
    c.z0_solver[:] = F(z0)
    c.relax = relax
    c.potdtdz = potdtdz
    return c


def capture(c, comp, i=I0, j=J0):
    r = pyoracle.run(c, iterations=1, capture=(comp, i, j, 1), lib_path=LIB_PATH)
    cap = r["capture"]
    assert cap["captured"] or comp == 4
    return cap["pieces"], r


def expected_vis(k, a, b, cov=0.0):
    l = 0.071 * (k - 0.5) * (1 - 0.99 * cov)
    return min(l * l * math.sqrt(0.5 * (a * a + b * b)), 15.0)


@pytest.mark.parametrize("comp", [0, 1, 2])
@pytest.mark.parametrize("cov", [0.0, 0.5])
def test_eddy_viscosity_of_linear_shear(comp, cov):
    a, b = 0.16, 0.08
    p, _ = capture(shear_case(a, b, cov=cov), comp)
    vis = p["vis"]
    assert vis[1] == 0                                # masked at the first level above ground
    for k in range(2, 9):
        assert vis[k] == pytest.approx(expected_vis(k, a, b, cov), rel=2e-6), k
    # misreadings: no 1/2 under the root, the level top instead of its centre,
    # no canopy reduction
    k = 5
    assert abs(vis[k] - (0.071 * (k - 0.5)) ** 2 * math.sqrt(a * a + b * b) * (1 - 0.99 * cov) ** 2) > 1e-3 * vis[k]
    assert abs(vis[k] - expected_vis(k + 0.5, a, b, cov)) > 1e-3 * vis[k]
    if cov:
        assert abs(vis[k] - expected_vis(k, a, b, 0.0)) > 1e-3 * vis[k]


def test_eddy_viscosity_cap():
    a = 80.0
    p, _ = capture(shear_case(a, 0.0), 0)
    uncapped = lambda k: (0.071 * (k - 0.5)) ** 2 * math.sqrt(0.5) * a
    assert uncapped(8) > 15.0 and p["vis"][8] == 15.0
    assert uncapped(7) < 15.0 and p["vis"][7] == pytest.approx(uncapped(7), rel=2e-6)


@pytest.mark.parametrize("sign", [1.0, -1.0])
def test_fluxes_conductances_and_power_law(sign):
    a, b = 0.16 * sign, 0.08
    c = shear_case(a, b)
    p, _ = capture(c, 0)
    dx = dy = 2.0
    for k in range(1, 9):
        dz = 1.0
        z = k - 0.5
        u, v = a * z, b * z
        assert p["fe"][k] == pytest.approx(u * dy * dz, rel=1e-6) and p["fw"][k] == pytest.approx(u * dy * dz, rel=1e-6)
        assert p["fs"][k] == pytest.approx(v * dx * dz, rel=1e-6) and p["fn"][k] == pytest.approx(v * dx * dz, rel=1e-6)
        assert p["ft"][k] == 0 and p["fb"][k] == 0
        vis = max(float(p["vis"][k]), 0.0)
        assert p["de"][k] == pytest.approx(max(vis * dy * dz / dx, 1e-4), rel=1e-6)
        assert p["ds"][k] == pytest.approx(max(vis * dx * dz / dy, 1e-4), rel=1e-6)
        assert p["dt"][k] == pytest.approx(max(vis * dx * dy / dz, 1e-4), rel=1e-6)
        assert p["db"][k] == pytest.approx(max(vis * dx * dy / dz, 1e-4) if k > 1 else 1e-4, rel=1e-6)
        # Patankar's power-law coefficients, each face with its own upwind side
        A = lambda d, f: d * max(0.0, (1 - 0.1 * abs(f / d)) ** 5)
        for nb, d, f, up in (("ae", "de", "fe", -1), ("aw", "dw", "fw", 1), ("an", "dn", "fn", -1),
                             ("as", "ds", "fs", 1), ("at", "dt", "ft", -1), ("ab", "db", "fb", 1)):
            want = A(float(p[d][k]), float(p[f][k])) + max(up * float(p[f][k]), 0.0)
            assert p[nb][k] == pytest.approx(want, rel=2e-5, abs=1e-9), (nb, k)
        # the upwind structure alone: a_W - a_E = F for equal faces
        assert float(p["aw"][k]) - float(p["ae"][k]) == pytest.approx(u * dy * dz, rel=1e-5)
        # a swapped upwind side would give a_W - a_E = -F
        assert abs((float(p["aw"][k]) - float(p["ae"][k])) + u * dy * dz) > 0.5 * abs(u * dy * dz)


def test_w_equation_fluxes_at_the_w_point():
    a, b = 0.16, 0.08
    p, _ = capture(shear_case(a, b), 2)
    for k in range(2, 9):
        zw = k - 1.0                                  # the w point: bottom face of level k
        assert p["fe"][k] == pytest.approx(a * zw * 2.0, rel=1e-6)
        assert p["fs"][k] == pytest.approx(b * zw * 2.0, rel=1e-6)
        assert abs(p["fe"][k] - a * (zw + 1.0) * 2.0) > 0.1 * abs(a * 2.0)   # the top face would be off by a dz


@pytest.mark.parametrize("comp", [0, 1])
def test_wall_drag_follows_the_log_law(comp):
    a, b, z0 = 0.4, 0.2, 0.05
    p, _ = capture(shear_case(a, b, z0=z0), comp)
    u1, v1 = a * 0.5, b * 0.5
    speed = math.hypot(u1, v1)
    ustar2 = (0.4 * speed / math.log(0.5 / z0)) ** 2  # log law at z1 = dz/2
    along = (u1 if comp == 0 else v1) / speed
    assert p["wall"][1] == pytest.approx(ustar2 * along * 4.0, rel=2e-5)
    assert not p["wall"][2:].any()                    # only the first level above ground
    wrong = (0.4 * speed / math.log(1.0 / z0)) ** 2 * along * 4.0   # z1 = dz
    assert abs(p["wall"][1] - wrong) > 0.1 * abs(wrong)


@pytest.mark.parametrize("comp", [0, 1])
def test_geostrophic_term(comp):
    a, b = 0.16, 0.08
    c = shear_case(a, b)
    c.coriolis = 1.0e-4
    p, r = capture(c, comp)
    ug, vg = r["stats"]["ug"], r["stats"]["vg"]
    assert ug == pytest.approx(3.2) and vg == 0
    for k in range(1, 9):
        vel = (a if comp == 0 else b) * (k - 0.5)
        g = ug if comp == 0 else vg
        assert p["cor"][k] == pytest.approx(1.0e-4 * (g - vel) * 4.0 * 1.0, rel=1e-5), k
        assert p["press"][k] == 0                     # no pressure gradient in this flow


def test_cross_terms_vanish_for_horizontally_uniform_flow():
    p, _ = capture(shear_case(0.16, 0.08), 0)
    assert not p["cross"].any()


def test_cross_term_gate_follows_the_vector_blocks():
    """In a building column (levels from k = 1) the cross terms are added only
    in vector blocks that start above KKART+1 (U-prognostic-microscale_1.cs:453):
    the lanes above the roof in the block that starts at k = 1 miss them."""
    c = domain.tiny_case(n=12, vertcellsff=14, building=(5, 7, 5, 7, 4.5), turbulence_model=1, iterations=2)
    i, j = 6, 6
    kk = int(c.kkart[i, j])
    assert c.cutk[i, j] > 0 and 1 <= kk < 7
    r = pyoracle.run(c, iterations=2, capture=(0, i, j, 2), lib_path=LIB_PATH)
    cross = r["capture"]["pieces"]["cross"]
    vi = min(c.nkk - 1, kk + c.vertcellsff)
    for k in range(1, vi + 1):
        block = 1 + 8 * ((k - 1) // 8)
        if block > vi + 1 - 8:
            block = vi + 1 - 8                          # the last block is moved down to end at VI
        if block <= kk + 1 or k <= kk + 1:
            assert cross[k] == 0, k
    assert np.count_nonzero(cross[kk + 2:vi + 1]) > 0
    assert not cross[kk + 2:9].any()                    # above the roof, yet in the first block


def test_tke_production_and_buoyancy():
    a, b = 0.16, 0.08
    c = shear_case(a, b, model=2, relax=0.0, potdtdz=0.01)
    p, r = capture(c, 4)
    eps = 1e-7                                          # TDISS after the initialisation
    for k in range(2, 9):
        vis = float(p["vis"][k])
        assert vis > 0
        assert p["prod"][k] == pytest.approx(vis * 4.0 * 1.0 * (a * a + b * b - eps), rel=1e-5), k
        buoy = -0.01 * 9.81 / 283.0 * vis * 1.35 * 4.0
        assert p["buoy"][k] == pytest.approx(buoy, rel=1e-6) and p["buoy"][k] < 0
    # a production without the 2x of the normal strains would be the same here,
    # a production from u_z alone would not
    k = 5
    assert abs(p["prod"][k] - float(p["vis"][k]) * 4.0 * a * a) > 0.1 * p["prod"][k]


def _linear_case(c_u, c_v, dyk=2.0):
    c = domain.build_case("lin", N, N, [domain.Building(3, 3, 3, 3, 0.4)], dx=2.0, dz1=1.0, stretch=1.0,
                          vertcellsff=10, turbulence_model=1, iterations=1)
    c.advdom[1:N + 1, 1:N + 1] = 1
    c.dyk = float(F(dyk))
    nkk = c.nkk
    c.uk[:] = 0
    c.vk[:] = 0
    c.wk[:] = 0
    i = np.arange(N + 2)
    c.uk[1:N + 2, 1:N + 2, 1:nkk + 1] = (F(c_u) * (i[1:N + 2] - 1)).astype(F)[:, None, None]
    c.vk[1:N + 2, 1:N + 2, 1:nkk + 1] = (F(c_v) * (i[1:N + 2] - 1)).astype(F)[None, :, None]
    c.relax = 0.0
    return c


def test_divergence_of_a_linear_field():
    c = _linear_case(0.04, 0.0, dyk=3.0)
    r = pyoracle.run(c, iterations=1, lib_path=LIB_PATH)
    div = r["div"]
    for i in range(4, N - 2):
        for j in range(4, N - 2):
            for k in range(1, 11):
                assert div[i, j, k] == pytest.approx(-0.04 * 3.0 * 1.0, rel=1e-5), (i, j, k)
    assert abs(div[6, 6, 3] - (-0.04 * 2.0)) > 0.01          # dx in place of dy would be off


def test_cell_averages_of_a_divergence_free_linear_field():
    c = _linear_case(0.04, -0.04)
    r = pyoracle.run(c, iterations=1, lib_path=LIB_PATH)
    assert np.abs(r["div"][4:N - 2, 4:N - 2, 1:11]).max() < 1e-6
    for i in range(4, N - 2):
        for j in range(4, N - 2):
            for k in range(2, 10):
                assert r["uks"][i, j, k] == pytest.approx(0.04 * (i - 0.5), rel=1e-5), (i, j, k)
                assert r["vks"][i, j, k] == pytest.approx(-0.04 * (j - 0.5), rel=1e-5), (i, j, k)
                assert r["wks"][i, j, k] == 0


def test_subdomain_boundary_copies_inflow_only():
    c = _linear_case(0.0, 0.0)
    c.advdom[1:5, 3:11] = 0                             # outside the sub-domain west of i = 5
    c.advdom[9:N + 1, 3:11] = 0                         # and east of i = 8 (rows 3-10; the rows
                                                        # checked are 2 away from the hole's N/S edges,
                                                        # where the N/S step copies u along the edge)
    nkk = c.nkk
    sign = np.where(np.arange(N + 2) % 2 == 0, 1.0, -1.0).astype(F)
    c.uk[3, :, 1:nkk + 1] = (F(0.8) * sign)[:, None]    # inflow on even rows, outflow on odd ones
    c.uk[4, :, 1:nkk + 1] = F(0.3)
    c.uk[10, :, 1:nkk + 1] = (F(-0.6) * sign)[:, None]  # east side: inflow is u <= 0
    c.uk[9, :, 1:nkk + 1] = F(0.2)
    r = pyoracle.run(c, iterations=1, lib_path=LIB_PATH)
    vi = 10
    for j in range(5, 9):
        # west: the face outside the sub-domain is neither corrected nor swept
        w_in = sign[j] > 0
        assert np.all(r["uk"][4, j, 1:vi + 1] == (F(0.8) if w_in else F(0.3))), j
        assert np.all(r["uk"][4, j, vi + 1:vi + 3] == F(0.3))          # above the column's top: untouched
        # east: the face of the last cell is corrected afterwards; its value after
        # the boundary step shows in that cell's divergence (u elsewhere is 0)
        e_in = -0.6 * sign[j] <= 0
        u_bc = -r["div"][8, j, 1:vi + 1] / (2.0 * 1.0)
        want = -0.6 * sign[j] if e_in else 0.2
        assert np.allclose(u_bc, want, rtol=1e-6), (j, u_bc[:3], want)
