"""The adjoint of the pressure-correction steps on the GPU (gral_pf_forward_ckpt,
gral_pf_vjp, gral_pf_restore_ckpt), through the C ABI.

* The forward window equals the oracle's orc_project bit for bit, with and
  without the quantisation, and restore_ckpt returns the starting fields.
* Against an FP64 CPU reference: the window is linear, so for any direction
  d of the inputs <g_bar, F(x + d) - F(x)> = <VJP(g_bar), d>; the left side
  is evaluated with the oracle's FP64 shadow, the right side with the GPU
  VJP (FP32), for random g_bar and d on config 1 (5 steps) and a small city.
* Central finite differences at h = 1e-3 of the mean wind speed at 64 seeded
  receptors about 3 m above ground, forward on the GPU, against <grad, d>:
  within 1e-3 relative.
Directions d are zero where the prologue rewrites the inputs (building cells,
ghost rows) and on the reference-point column (whose speed sets DTIME), so
that the perturbed solve sees exactly x + h d with the same DTIME.
"""
import numpy as np
import pytest

from gral_amd import adjoint, domain
from tests.helpers import diff_report

NSTEP = 5


def _mask(c):
    m = np.zeros(c.uk.shape, dtype=bool)
    m[2:c.nii, 2:c.njj, 1:c.kadvmax + 1] = True
    k = np.arange(c.uk.shape[2])[None, None, :]
    m &= k > c.kkart[:, :, None]
    m[c.refp[0], c.refp[1], :] = False
    return m


def _direction(c, seed):
    rng = np.random.default_rng(seed)
    m = _mask(c)
    return tuple(np.where(m, rng.standard_normal(m.shape), 0.0).astype(np.float32) for _ in range(3))


def _shifted(c, d, h):
    """The case with fields x + h d, kept in float64: the oracle's FP64 shadow
    then runs on the exact sum; the GPU rounds it to float32 on upload."""
    s = c.copy()
    s.uk = c.uk.astype(np.float64) + h * d[0].astype(np.float64)
    s.vk = c.vk.astype(np.float64) + h * d[1].astype(np.float64)
    s.wk = c.wk.astype(np.float64) + h * d[2].astype(np.float64)
    return s


def _forward_gpu(c, n=NSTEP, quantise=False):
    from gral_amd import pf
    with pf.PrognosticFlowfield(c) as sv:
        sv.prepare(iterations=n)
        sv.forward_ckpt(n, quantise=quantise)
        return sv.download()


@pytest.mark.gpu
@pytest.mark.parametrize("quantise", [0, 1])
def test_forward_window_equals_oracle(quantise):
    from gral_amd import pf
    from oracle import pyoracle
    c = domain.config1(iterations=NSTEP)
    print(f"[adjoint] forward window config1 quantise={quantise}", flush=True)
    ref = pyoracle.run(c, iterations=NSTEP, project=quantise)
    with pf.PrognosticFlowfield(c) as sv:
        sv.prepare(iterations=NSTEP)
        start = sv.download()
        sv.forward_ckpt(NSTEP, quantise=bool(quantise))
        got = sv.download()
        sv.restore_ckpt()
        back = sv.download()
    for f, r in diff_report(got, ref, ("uk", "vk", "wk", "dpm")).items():
        assert r["n_bits_differ"] == 0, (f, r)
    for f in ("uk", "vk", "wk"):
        assert np.array_equal(back[f], start[f]), f


def _dot(a, b):
    return sum(float(np.dot(x.astype(np.float64).ravel(), y.astype(np.float64).ravel())) for x, y in zip(a, b))


@pytest.mark.gpu
@pytest.mark.parametrize("which", ["config1", "city"])
def test_vjp_against_fp64_reference(which):
    from gral_amd import pf
    from oracle import pyoracle
    if which == "config1":
        c = domain.config1(iterations=NSTEP)
    else:
        bl = domain.city_buildings(40, 40, 2.0, 5, density=0.25, fmin=6.0, fmax=16.0, hmin=4.0, hmax=14.0, margin=3)
        c = domain.build_case("city40", 40, 40, bl, dx=2.0, vertcellsff=12, speed10=4.0, direction=235.0, z0=0.5,
                              iterations=NSTEP)
    print(f"[adjoint] dot-product test {which}", flush=True)
    rng = np.random.default_rng(3)
    lev = np.zeros(c.uk.shape, dtype=bool)
    lev[:, :, :c.kadvmax + 2] = True
    gbar = tuple(np.where(lev, rng.standard_normal(c.uk.shape), 0.0).astype(np.float32) for _ in range(3))
    with pf.PrognosticFlowfield(c) as sv:
        sv.prepare(iterations=NSTEP)
        sv.forward_ckpt(NSTEP)
        g = sv.vjp(*gbar)
    base = pyoracle.run(c, iterations=NSTEP, project=0, fp64=True)
    for seed in (11, 12):
        d = _direction(c, seed)
        pert = pyoracle.run(_shifted(c, d, 1.0), iterations=NSTEP, project=0, fp64=True)
        lhs = _dot(gbar, [pert[f] - base[f] for f in ("uk", "vk", "wk")])
        rhs = _dot(g, d)
        print(f"[adjoint] {which} seed {seed}: <gbar, F(x+d)-F(x)> = {lhs:.9g}, <VJP, d> = {rhs:.9g}", flush=True)
        assert rhs == pytest.approx(lhs, rel=1e-4)


@pytest.mark.gpu
def test_receptor_speed_gradient_central_differences():
    from gral_amd import pf
    c = domain.config1(iterations=NSTEP)
    rec = adjoint.receptors(c)
    assert len(rec) == 64
    print("[adjoint] central differences, 64 receptors at 3 m", flush=True)
    with pf.PrognosticFlowfield(c) as sv:
        sv.prepare(iterations=NSTEP)
        sv.forward_ckpt(NSTEP)
        out = sv.download()
        du, dv, dw = adjoint.mean_receptor_speed_grad(out, rec)
        g = sv.vjp(du, dv, dw)
    h = 1e-3
    for seed in (21, 22, 23):
        d = _direction(c, seed)
        cp, cm = _shifted(c, d, h), _shifted(c, d, -h)
        for s in (cp, cm):                              # the fields the GPU solves: x +- h d in float32
            s.uk, s.vk, s.wk = s.uk.astype(np.float32), s.vk.astype(np.float32), s.wk.astype(np.float32)
        lp = adjoint.mean_receptor_speed(_forward_gpu(cp), rec)
        lm = adjoint.mean_receptor_speed(_forward_gpu(cm), rec)
        fd = (lp - lm) / (2 * h)
        # the direction those rounded fields realise
        d_eff = [(getattr(cp, f).astype(np.float64) - getattr(cm, f)) / (2 * h) for f in ("uk", "vk", "wk")]
        ad = _dot(g, d_eff)
        print(f"[adjoint] seed {seed}: central difference {fd:.7g}, adjoint {ad:.7g}", flush=True)
        assert ad == pytest.approx(fd, rel=1e-3)
    prof = adjoint.inflow_profile_gradient(c, g[0].astype(np.float64), g[1].astype(np.float64))
    assert np.isfinite(prof).all() and np.abs(prof[1:c.kadvmax]).max() > 0
