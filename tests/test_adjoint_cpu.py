"""CPU checks behind the adjoint tests: the forward window the library
differentiates (orc_project without quantisation, in the oracle's FP64
shadow) is linear in the inputs along the directions the GPU tests perturb,
and the receptor functional's gradient matches its finite differences."""
import numpy as np
import pytest

from gral_amd import adjoint, domain
from tests.test_gpu_adjoint import NSTEP, _direction, _shifted


def test_forward_window_is_linear_along_masked_directions():
    from oracle import pyoracle
    c = domain.config1(iterations=NSTEP)
    d = _direction(c, 5)
    base = pyoracle.run(c, iterations=NSTEP, project=0, fp64=True)
    one = pyoracle.run(_shifted(c, d, 1.0), iterations=NSTEP, project=0, fp64=True)
    two = pyoracle.run(_shifted(c, d, 2.0), iterations=NSTEP, project=0, fp64=True)
    assert base["stats"]["dtime"] == one["stats"]["dtime"] == two["stats"]["dtime"]
    for f in ("uk", "vk", "wk"):
        a, b = one[f] - base[f], two[f] - base[f]
        assert np.abs(b - 2 * a).max() <= 1e-9 * max(1.0, np.abs(b).max()), f
        assert np.abs(a).max() > 0.1                  # the window does move the perturbation


def test_receptor_speed_gradient_matches_finite_differences():
    c = domain.config1(iterations=1)
    rec = adjoint.receptors(c)
    assert len(rec) == 64 and len({tuple(r) for r in rec}) == 64
    assert all(c.cutk[i, j] == 0 and c.advdom[i, j] == 1 for i, j, _ in rec)
    rng = np.random.default_rng(1)
    f = {n: getattr(c, n).astype(np.float64) + rng.standard_normal(c.uk.shape) for n in ("uk", "vk", "wk")}
    gu, gv, gw = adjoint.mean_receptor_speed_grad(f, rec)
    d = {n: rng.standard_normal(c.uk.shape) for n in ("uk", "vk", "wk")}
    h = 1e-6
    fp = {n: f[n] + h * d[n] for n in f}
    fm = {n: f[n] - h * d[n] for n in f}
    fd = (adjoint.mean_receptor_speed(fp, rec) - adjoint.mean_receptor_speed(fm, rec)) / (2 * h)
    ad = float((gu * d["uk"]).sum() + (gv * d["vk"]).sum() + (gw * d["wk"]).sum())
    assert ad == pytest.approx(fd, rel=1e-6)
