"""Closed-form and textbook checks of the CPU oracle.

These pin parts of the oracle independently of how the reference's C# was
read: single formulas against their closed forms, and the column solves
against a banded FP64 solver (scipy.linalg.solve_banded) applied to the
coefficients the oracle itself assembled (captured through the oracle's
checking hook).  A swapped P/Q convention, a sign error in the recurrences,
an off-by-one level or a wrong relaxation form fails here.
"""
import numpy as np
import pytest
from scipy.linalg import solve_banded

from gral_amd import domain
from oracle import pyoracle

F = np.float32


@pytest.mark.parametrize("peclet,factor", [(0.0, 1.0), (1.0, 0.59049), (5.0, 0.03125), (10.0, 0.0), (20.0, 0.0),
                                           (2.5, 0.75 ** 5)])
def test_power_law_closed_form(peclet, factor):
    # Patankar's power-law scheme: A = D * max(0, (1 - 0.1 |P|)^5)
    lib = pyoracle.load()
    for d in (1.0, 3.7, 250.0):
        for fn in (lib.orc_powlaw_vector, lib.orc_powlaw_scalar):
            assert fn(d, peclet) == pytest.approx(d * factor, rel=2e-6, abs=1e-30)


@pytest.mark.parametrize("x,n", [(2.5, 3), (-2.5, -3), (2.4999999, 2), (-2.4999999, -2), (0.49, 0), (-0.49, 0),
                                 (-0.5, -1), (0.5, 1), (1e9 + 0.5, 1000000001), (-7.0, -7)])
def test_conv_to_int_rounds_half_away_from_zero(x, n):
    assert pyoracle.load().orc_conv_to_int(x) == n


def _tridiag_solve(lower, diag, upper, rhs):
    """Solve diag[k] x[k] + lower[k] x[k-1] + upper[k] x[k+1] = rhs[k] in FP64."""
    n = len(diag)
    ab = np.zeros((3, n))
    ab[0, 1:] = upper[:-1]
    ab[1, :] = diag
    ab[2, :-1] = lower[1:]
    return solve_banded((1, 1), ab, rhs)


def _case(relax):
    c = domain.tiny_case(n=12, vertcellsff=9, turbulence_model=1, iterations=1)
    c.relax = relax
    return c


@pytest.mark.parametrize("comp", [0, 1, 2])
@pytest.mark.parametrize("relax", [1.0, 0.1])
def test_momentum_column_solve_against_banded_solver(comp, relax):
    # a ground column (KKART = 0, KSTART = 1) with ground neighbours; for W one
    # next to the building, where the first iteration produces vertical flow
    c = _case(relax)
    i, j = (6, 5) if comp == 2 else (9, 9)
    assert c.kkart[i, j] == 0 and c.cutk[i, j] == 0
    if comp != 2:                                  # U/V skip levels below the upstream neighbour's KKART
        assert c.kkart[i - 1, j] == 0 and c.kkart[i, j - 1] == 0
    r = pyoracle.run(c, iterations=1, capture=(comp, i, j, 1))
    cap = r["capture"]
    assert cap["captured"]
    vi = int(min(c.nkk - 1, c.kkart[i, j] + c.vertcellsff))
    ks = np.arange(1, vi + 1)
    a, b, cc, d = (cap[n][ks].astype(np.float64) for n in ("a", "b", "c", "d"))
    xo = cap["x_old"].astype(np.float64)
    cc[0] = 0.0                                    # recurrence reset at k <= KKART+1
    # Thomas with the reference's convention x[k] = P[k] x[k+1] + Q[k]
    P = np.zeros(vi + 2)
    Q = np.zeros(vi + 2)
    for q, k in enumerate(ks):
        den = a[q] - cc[q] * P[k - 1]
        P[k] = b[q] / den
        Q[k] = (d[q] + cc[q] * Q[k - 1]) / den
    rhs = d.copy()
    rhs[-1] += b[-1] * xo[vi + 1]                  # fixed top value
    exact = _tridiag_solve(-cc, a, -b, rhs)
    # relaxed back-substitution from the top: z[k] = z_old + relax (P z[k+1] + Q - z_old)
    z = xo.copy()
    for k in range(vi, 0, -1):
        z[k] = xo[k] + relax * (P[k] * z[k + 1] + Q[k] - xo[k])
    got = cap["x_new"][1:vi + 1].astype(np.float64)
    scale = np.abs(got).max()
    assert scale > 0
    # FP32 recurrences over 9 levels stay within a few ulps; exchanging the
    # upper and lower couplings (equal inside the column, different at the
    # reset level) moves the result by ~2e-6 at relax = 1
    assert np.abs(got - z[1:vi + 1]).max() <= 5e-7 * scale
    if relax == 1.0:
        assert np.abs(got - exact).max() <= 5e-7 * scale
    assert np.array_equal(cap["x_new"][vi + 1:], cap["x_old"][vi + 1:])


def test_pressure_column_solve_against_banded_solver():
    # first iteration, sweep (IS, JS) = (-1, -1): the row through the building is
    # swept from i = NII-1 down, so a ground column east of the building comes
    # before every taller column of its row and the row scratch above its top
    # (PIMP/QIMP[VI+1]) is still zero; the building makes its divergence non-zero
    c = _case(0.1)
    i, j = 6, 5
    assert (c.kkart[i:, j] == 0).all() and c.kkart[i - 1, j] > 0
    r = pyoracle.run(c, iterations=1, capture=(3, i, j, 1))
    cap = r["capture"]
    assert cap["captured"]
    vi = int(c.vertcellsff)
    ks = np.arange(1, vi + 1)
    app, abp, atp, dim = (cap[n][ks].astype(np.float64) for n in ("a", "b", "c", "d"))
    dx, dy, dz = float(c.dxk), float(c.dyk), c.dzk.astype(np.float64)
    # closed-form coefficients of the 7-point pressure-correction equation
    area = dx * dy
    app_cf = 2 * (dy * dz[ks] / dx + dx * dz[ks] / dy + area / dz[ks])
    abp_cf = np.where(ks > 1, area / (0.5 * (dz[ks - 1] + dz[ks])), area / dz[ks])
    atp_cf = area / (0.5 * (dz[ks + 1] + dz[ks]))
    for got, ref in ((app, app_cf), (abp, abp_cf), (atp, atp_cf)):
        assert np.allclose(got, ref, rtol=3e-6, atol=0)
    # APP x[k] = ABP x[k-1] + ATP x[k+1] + DIM, x[0] = DPM[0], x[VI+1] = 0
    rhs = dim.copy()
    rhs[0] += abp[0] * float(cap["x_old"][0])
    exact = _tridiag_solve(-abp, app, -atp, rhs)
    got = cap["x_new"][1:vi + 1].astype(np.float64)
    assert np.abs(exact).max() > 0
    assert np.abs(got - exact).max() <= 5e-7 * np.abs(exact).max()


def _imbalance(case, f):
    """FP64 net volume flux out of every interior cell above the ground (blocked
    faces closed), i.e. the mass balance the diagnostic pass restores."""
    c = case
    dx, dy = float(c.dxk), float(c.dyk)
    dz = c.dzk.astype(np.float64)
    U, V, W = (f[n].astype(np.float64) for n in ("uk", "vk", "wk"))
    kk = c.kkart.astype(int)
    out = []
    for i in range(2, c.nii):
        for j in range(2, c.njj):
            for k in range(kk[i, j] + 1, c.nkk):
                fw = 0.0 if k <= kk[i - 1, j] else dz[k] * dy * U[i, j, k]
                fe = 0.0 if k <= kk[i + 1, j] else dz[k] * dy * U[i + 1, j, k]
                fs = 0.0 if k <= kk[i, j - 1] else dz[k] * dx * V[i, j, k]
                fn = 0.0 if k <= kk[i, j + 1] else dz[k] * dx * V[i, j + 1, k]
                fb = 0.0 if (k <= kk[i, j] + 1 or k == 1) else dx * dy * W[i, j, k]
                out.append(fw - fe + fs - fn + fb - dx * dy * W[i, j, k + 1])
    return np.abs(np.array(out))


def test_diagnostic_w_adjustment_closes_every_cell():
    # in diagnostic mode (level 1) the final step sets W at each cell top so that
    # the cell's net flux vanishes: a property of the result, not of the code path
    c = domain.tiny_case(n=10, vertcellsff=9, turbulence_model=1, iterations=1)
    r = pyoracle.diagnostic(c, None, flow_field_level=1)
    scale = np.abs(r["uk"]).max() * float(c.dzk.max()) * float(c.dxk)
    assert _imbalance(c, r).max() <= 1e-5 * scale


def test_diagnostic_pressure_iterations_reduce_the_mass_imbalance():
    # each pressure-correction iteration moves the field towards zero divergence;
    # measured over the cells of the sub-domain, which the final W step leaves alone
    c = domain.tiny_case(n=10, vertcellsff=9, turbulence_model=1, iterations=1)
    f0 = {"uk": c.uk, "vk": c.vk, "wk": c.wk, "dpm": np.zeros_like(c.uk)}
    imb = [_imbalance(c, f0).sum()]
    for n in (1, 4, 16):
        imb.append(_imbalance(c, pyoracle.diagnostic(c, f0, flow_field_level=2, iterations=n)).sum())
    assert imb[0] > imb[1] > imb[2] > imb[3], imb
