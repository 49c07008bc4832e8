"""The adjoint of one mixing-length momentum sweep (gral_pf_momentum_sweep /
gral_pf_momentum_vjp), through the C ABI, with UKS, VKS, WKS, DPMNEW and the
other components held fixed (gral_pf_set_velocity replaces the velocities of a
prepared situation without redoing its prologue, so the perturbed sweeps see
exactly those fixed fields).

* The sweep changes the swept component, and restore_ckpt returns the fields
  to their values before it bit for bit.
* Central differences: for random directions d of the swept component and a
  random adjoint g_bar on the levels the solver holds,
  (<g_bar, M(x + h d)> - <g_bar, M(x - h d)>) / 2h against <VJP(g_bar), d_eff>
  (d_eff: the direction the float32 fields x +- h d realise), within 1e-3
  relative, for U, V and W on config 1 (one building, 100 x 100) after 3
  iterations and on a 40 x 40 city with several buildings, at the configured
  under-relaxation (0.1: the identity part (1 - relax) x dominates) and at
  relax = 1, where the new values are the Thomas solutions alone and the
  derivative is all coupling, eddy viscosity and power law.
* The transposed sweep is linear in g_bar: VJP(a g1 + g2) = a VJP(g1) + VJP(g2)
  to float rounding.
* Single cells in a building's wake whose derivative runs mostly through the
  eddy viscosity: with that path removed from the transpose the U cell at
  (60, 50, 14) gives -6.7e-5 against a central difference of 1.05e-3, and the V
  cell at (56, 52, 10) -8e-9 against 4.9e-4, so a dropped viscosity path
  fails here even where the random directions above average it away.
"""
import numpy as np
import pytest

from gral_amd import domain


def _city():
    bl = domain.city_buildings(40, 40, 2.0, 5, density=0.25, fmin=6.0, fmax=16.0, hmin=4.0, hmax=14.0, margin=3)
    return domain.build_case("city40", 40, 40, bl, dx=2.0, vertcellsff=12, speed10=4.0, direction=235.0, z0=0.5,
                             iterations=12)


CASES = {"config1": lambda: domain.config1(iterations=12), "city": _city}
FIELD = ("uk", "vk", "wk")


def _levels(c):
    m = np.zeros(c.uk.shape, dtype=bool)
    m[:, :, :c.kadvmax + 2] = True
    return m


def _dot(a, b):
    return float(np.dot(a.astype(np.float64).ravel(), b.astype(np.float64).ravel()))


def _setup(sv, iters):
    sv.prepare(iterations=12)
    sv.iterate(iters)
    return sv.download()


@pytest.mark.gpu
@pytest.mark.parametrize("relax", [None, 1.0])
@pytest.mark.parametrize("which", ["config1", "city"])
@pytest.mark.parametrize("comp", [0, 1, 2])
def test_momentum_vjp_central_differences(which, comp, relax):
    from gral_amd import pf
    c = CASES[which]()
    if relax is not None:
        c.relax = relax
    f = FIELD[comp]
    print(f"[momentum adjoint] {which} component {f} relax {c.relax}: central differences", flush=True)
    lev = _levels(c)
    rng = np.random.default_rng(100 + comp)
    gbar = np.where(lev, rng.standard_normal(c.uk.shape), 0.0).astype(np.float32)
    with pf.PrognosticFlowfield(c) as sv:
        x0 = _setup(sv, 3)
        base = {k: x0[k].copy() for k in FIELD}
        sv.momentum_sweep(comp)
        after = sv.download()
        g = sv.momentum_vjp(gbar)
        sv.restore_ckpt()
        back = sv.download()
        for k in FIELD:
            assert np.array_equal(back[k], base[k]), k
        assert not np.array_equal(after[f], base[f])
        scale = float(np.abs(base[f][lev]).max())
        h = 1e-3 * scale
        for seed in (1, 2):
            d = np.where(lev, np.random.default_rng(seed).standard_normal(c.uk.shape), 0.0)
            out = []
            xs = []
            for sgn in (1.0, -1.0):
                x = {k: base[k].copy() for k in FIELD}
                x[f] = (base[f].astype(np.float64) + sgn * h * d).astype(np.float32)
                xs.append(x[f])
                sv.set_velocity(x["uk"], x["vk"], x["wk"])
                sv.momentum_sweep(comp)
                out.append(sv.download()[f])
            fd = (_dot(gbar, out[0]) - _dot(gbar, out[1])) / (2 * h)
            d_eff = (xs[0].astype(np.float64) - xs[1]) / (2 * h)
            ad = _dot(g, d_eff)
            print(f"[momentum adjoint] {which} {f} relax {c.relax} seed {seed}: central difference {fd:.8g}, "
                  f"adjoint {ad:.8g}, relative {abs(ad - fd) / abs(fd):.2e}", flush=True)
            assert ad == pytest.approx(fd, rel=1e-3)


@pytest.mark.gpu
@pytest.mark.parametrize("comp", [0, 2])
def test_momentum_vjp_linear_in_gbar(comp):
    from gral_amd import pf
    c = CASES["config1"]()
    lev = _levels(c)
    rng = np.random.default_rng(7)
    g1 = np.where(lev, rng.standard_normal(c.uk.shape), 0.0).astype(np.float32)
    g2 = np.where(lev, rng.standard_normal(c.uk.shape), 0.0).astype(np.float32)
    print(f"[momentum adjoint] linearity, component {FIELD[comp]}", flush=True)
    with pf.PrognosticFlowfield(c) as sv:
        _setup(sv, 2)
        sv.momentum_sweep(comp)
        r1 = sv.momentum_vjp(g1).astype(np.float64)
        r2 = sv.momentum_vjp(g2).astype(np.float64)
        r12 = sv.momentum_vjp((2.0 * g1 + g2).astype(np.float32)).astype(np.float64)
    err = np.abs(r12 - (2.0 * r1 + r2)).max()
    assert err <= 1e-5 * max(1.0, np.abs(r12).max()), err


# wake cells of config 1 after 3 iterations at relax 1 (the building spans
# i, j = 46..55, up to level 15), and the step of each central difference: at
# these cells most of d X'[k0] / d X[k0+1] runs through the eddy viscosity of
# level k0 (removing that path from the transpose leaves 1-3 % of the U values)
WAKE = {0: [((60, 50, 14), 1e-2), ((58, 50, 14), 1e-2), ((60, 50, 16), 1e-2)],
        1: [((60, 50, 14), 1e-2), ((58, 50, 14), 1e-2), ((56, 52, 10), 3e-3)]}


@pytest.mark.gpu
@pytest.mark.parametrize("comp", [0, 1])
def test_momentum_vjp_wake_cells(comp):
    """Single-cell derivatives through the eddy viscosity: L = the new value of
    one cell at level k0 in the building's wake, relax = 1, its derivative with
    respect to the old value one level above (which enters VIS at k0, hence the
    conductances and power-law coefficients of the level and the cross term), by
    central differences against the VJP, within 1e-2 relative (steps of 1e-2 or
    3e-3 of the value: float32 resolves the change of one new value)."""
    from gral_amd import pf
    c = CASES["config1"]()
    c.relax = 1.0
    f = FIELD[comp]
    print(f"[momentum adjoint] wake cells, component {f}", flush=True)
    with pf.PrognosticFlowfield(c) as sv:
        x0 = _setup(sv, 3)
        base = {k: x0[k].copy() for k in FIELD}
        for (i, j, k0), step in WAKE[comp]:
            gbar = np.zeros(c.uk.shape, dtype=np.float32)
            gbar[i, j, k0] = 1.0
            sv.set_velocity(base["uk"], base["vk"], base["wk"])
            sv.momentum_sweep(comp)
            g = sv.momentum_vjp(gbar)
            kp = k0 + 1
            h = step * max(0.1, abs(float(base[f][i, j, kp])))
            vals = []
            for sgn in (1.0, -1.0):
                x = {k: base[k].copy() for k in FIELD}
                x[f][i, j, kp] = np.float32(float(base[f][i, j, kp]) + sgn * h)
                sv.set_velocity(x["uk"], x["vk"], x["wk"])
                sv.momentum_sweep(comp)
                vals.append((float(sv.download()[f][i, j, k0]), float(x[f][i, j, kp])))
            fd = (vals[0][0] - vals[1][0]) / (vals[0][1] - vals[1][1])
            ad = float(g[i, j, kp])
            print(f"[momentum adjoint] {f} ({i}, {j}): d X'[{k0}] / d X[{kp}]: central difference {fd:.7g}, "
                  f"adjoint {ad:.7g}", flush=True)
            assert ad == pytest.approx(fd, rel=1e-2)


@pytest.mark.gpu
def test_momentum_entry_point_errors():
    """Call order and arguments: a VJP before any sweep and a sweep before
    prepare return GRAL_PF_ESTATE, a component outside 0..2 and the k-eps model
    GRAL_PF_EINVAL; a failed call changes no field."""
    from gral_amd import pf
    print("[momentum adjoint] error returns", flush=True)
    c = CASES["config1"]()
    with pf.PrognosticFlowfield(c) as sv:
        with pytest.raises(RuntimeError, match=r"\(-4\)"):
            sv.momentum_sweep(0)
        with pytest.raises(RuntimeError, match=r"\(-4\)"):
            sv.set_velocity(c.uk, c.vk, c.wk)
        x0 = _setup(sv, 1)
        with pytest.raises(RuntimeError, match=r"\(-4\)"):
            sv.momentum_vjp(np.zeros(c.uk.shape, dtype=np.float32))
        for comp in (-1, 3):
            with pytest.raises(RuntimeError, match=r"\(-1\)"):
                sv.momentum_sweep(comp)
        after = sv.download()
        for k in FIELD:
            assert np.array_equal(after[k], x0[k]), k
        # a VJP refers to the fields its sweep left: any change since voids it
        for change in (lambda: sv.iterate(1), lambda: sv.set_velocity(x0["uk"], x0["vk"], x0["wk"]),
                       sv.restore_ckpt):
            sv.momentum_sweep(1)
            sv.momentum_vjp(np.ones(c.uk.shape, dtype=np.float32))
            change()
            with pytest.raises(RuntimeError, match=r"\(-4\)"):
                sv.momentum_vjp(np.ones(c.uk.shape, dtype=np.float32))
    with pf.PrognosticFlowfield(domain.config1(turbulence_model=2, iterations=12)) as sv:
        sv.prepare(iterations=12)
        with pytest.raises(RuntimeError, match=r"\(-1\)"):
            sv.momentum_sweep(0)
