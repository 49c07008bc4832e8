"""GPU parity of gral_pf_wind_at_particles (Intwind.cs:31-170) against the
oracle, bit for bit, on a solved situation: particles spread over every cell
and over the whole column height (above the solved levels the fields come
from the device copy of the upper levels), heights on level faces included,
both branches (terrain and flat terrain with buildings)."""
import numpy as np
import pytest

from gral_amd import domain
from oracle import intwind_oracle as W

F = np.float32


@pytest.mark.gpu
@pytest.mark.parametrize("topo", [0, 1])
def test_wind_at_particles_matches_oracle(topo):
    from gral_amd import pf
    c = domain.tiny_case(n=12, vertcellsff=10, turbulence_model=1, iterations=3)
    ikoo, jkoo, ahmin = 600000, 5200000, 100.0
    r = np.random.default_rng(40 + topo)
    n = 20000
    ci = r.integers(1, c.nii + 1, n).astype(np.int32)
    cj = r.integers(1, c.njj + 1, n).astype(np.int32)
    x = ikoo + (ci - 1 + r.uniform(0, 1, n)) * c.dxk
    y = jkoo + (cj - 1 + r.uniform(0, 1, n)) * c.dyk
    h = r.uniform(0.0, float(c.hokart[c.nkk]) + 3.0, n).astype(F)
    h[:c.nkk + 2] = c.hokart[:c.nkk + 2]                       # exactly on level faces
    z = (h + F(ahmin)).astype(F) if topo else h
    ahint = r.uniform(0.0, 4.0, n).astype(F)
    print(f"[intwind] topo={topo} {c.nii}x{c.njj}x{c.nkk} particles={n}", flush=True)
    with pf.PrognosticFlowfield(c) as sv:
        sv.solve(3)
        f = sv.download()
        got = sv.wind_at_particles(topo, ikoo, jkoo, ahmin, ci, cj, ahint, x, y, z)
    assert c.kadvmax + 2 < c.nkk                               # levels above the solved ones are read
    want = W.wind_at_particles(f["uk"], f["vk"], f["wk"], c.hokart, c.dzk, c.nkk, c.dxk, c.dyk, ikoo, jkoo, ahmin,
                               topo, ci, cj, ahint, x, y, z)
    for name, a, b in zip(("ux", "uy", "uz"), want[:3], got[:3]):
        bad = np.argwhere(a.view(np.uint32) != b.view(np.uint32))
        assert bad.size == 0, (name, bad[:5].tolist())
    assert np.array_equal(want[3], got[3])
    assert (got[3] > c.kadvmax).any() and (got[3] <= c.kadvmax).any()


@pytest.mark.gpu
def test_wind_at_particles_rejects_cells_outside():
    from gral_amd import pf
    c = domain.tiny_case(n=9, vertcellsff=8, turbulence_model=1, iterations=1)
    with pf.PrognosticFlowfield(c) as sv:
        sv.solve(1)
        with pytest.raises(RuntimeError):
            sv.wind_at_particles(1, 0, 0, 0.0, [c.nii + 1], [1], [0.0], [0.0], [0.0], [1.0])
