"""FP64 mass-imbalance reduction on the GPU (gral_pf_mass_imbalance).

Sum over the active cells of |DIV| (DIV formed in FP32 as the divergence pass
does, PrognosticFlowfield.cs:723-758) and of the absolute face fluxes.  Checked
against a closed form (uniform flow through a domain without buildings: no
divergence, and 2 |u| DY DZ per cell), against float64 sums of the same
per-cell terms on the downloaded fields, and for repeatability."""
import numpy as np
import pytest

from gral_amd import domain


def mass_numpy(case, f):
    """Float64 sums of |DIV| and |face flux| over the active cells of fields f."""
    c = case
    uk, vk, wk = f["uk"], f["vk"], f["wk"]
    dy, dx = np.float32(c.dyk), np.float32(c.dxk)
    area = np.float32(dx * dy)
    sd = sf = 0.0
    for i in range(2, c.nii):
        for j in range(2, c.njj):
            if c.advdom[i, j] != 1:
                continue
            kk = int(c.kkart[i, j])
            vi = min(c.nkk - 1, kk + c.vertcellsff)
            k = np.arange(kk + 1, vi + 1)
            dz = c.dzk[k].astype(np.float32)
            u0, u1 = uk[i, j, k], uk[i + 1, j, k]
            v0, v1 = vk[i, j, k], vk[i, j + 1, k]
            w0, w1 = wk[i, j, k], wk[i, j, k + 1]
            t1 = (u0 - u1) * dy * dz
            t2 = (v0 - v1) * dx * dz
            div = np.where(k > kk + 1, t1 + t2 + (w0 - w1) * area, t1 + t2 - w1 * area).astype(np.float32)
            sd += np.abs(div).astype(np.float64).sum()
            fl = (np.abs(u0 * dy * dz).astype(np.float64) + np.abs(u1 * dy * dz) + np.abs(v0 * dx * dz) +
                  np.abs(v1 * dx * dz) + np.abs(w1 * area) + np.where(k > kk + 1, np.abs(w0 * area), 0))
            sf += fl.sum()
    return sd, sf


@pytest.mark.gpu
def test_uniform_flow_closed_form():
    from gral_amd import pf
    c = domain.tiny_case(n=12, vertcellsff=10, building=(6, 6, 6, 6, 0.4), turbulence_model=1, iterations=1)
    c.advdom[1:13, 1:13] = 1                     # one sub-domain over the whole grid
    assert c.kkart.max() == 0                    # the roughness element blocks no level
    c.uk[:] = np.float32(2.5)
    c.vk[:] = 0
    c.wk[:] = 0
    print("[mass] uniform flow", flush=True)
    with pf.PrognosticFlowfield(c) as sv:
        sv.prepare(iterations=1)
        sd, sf = sv.mass_imbalance()
    assert sd == 0.0
    dz = c.dzk[1:c.vertcellsff + 1].astype(np.float32)
    per_col = 2.0 * np.abs(np.float32(2.5) * np.float32(c.dyk) * dz).astype(np.float64).sum()
    assert sf == pytest.approx(per_col * (c.nii - 2) * (c.njj - 2), rel=1e-13)


@pytest.mark.gpu
@pytest.mark.parametrize("model", [1, 2])
def test_against_float64_sums(model):
    from gral_amd import pf
    c = domain.config1(turbulence_model=model, iterations=3)
    print(f"[mass] config1 model={model}", flush=True)
    with pf.PrognosticFlowfield(c) as sv:
        sv.prepare(iterations=3)
        sv.iterate(3)
        a = sv.mass_imbalance()
        b = sv.mass_imbalance()
        st = sv.finish()
        f = sv.download()
    assert a == b                                   # same bits on a repeated call
    assert (st["mass_div"], st["mass_flux"]) == a   # finish reports the same figures
    sd, sf = mass_numpy(c, f)
    assert a[0] == pytest.approx(sd, rel=1e-12)
    assert a[1] == pytest.approx(sf, rel=1e-12)
    assert 0 < a[0] < a[1]
