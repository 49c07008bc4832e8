"""GPU parity of the mass-conservation post-pass (DiagnosticFlowfield and the
final vertical-velocity step) against the CPU oracle, bit for bit."""
import numpy as np
import pytest

from gral_amd import domain

FIELDS = ("uk", "vk", "wk", "dpm")


def _gpu(case, prognostic_iterations, level, terrain):
    from gral_amd import pf
    with pf.PrognosticFlowfield(case) as s:
        if prognostic_iterations:
            s.prepare(prognostic_iterations)
            s.iterate(prognostic_iterations)
            s.finish()
        s.diagnostic(level, terrain)
        return s.download()


def _oracle(case, prognostic_iterations, level, terrain):
    from oracle import pyoracle
    if prognostic_iterations:
        r = pyoracle.run(case, iterations=prognostic_iterations)
        f = {k: r[k] for k in FIELDS}
    else:
        f = None
    return pyoracle.diagnostic(case, f, flow_field_level=level, terrain=terrain)


def _compare(case, its, level, terrain=False):
    print(f"[diagnostic] {case.name} {case.nii}x{case.njj} prognostic it={its} level={level} terrain={terrain}",
          flush=True)
    got = _gpu(case, its, level, terrain)
    ref = _oracle(case, its, level, terrain)
    for f in FIELDS:
        bad = np.argwhere(got[f].view(np.uint32) != ref[f].view(np.uint32))
        assert bad.size == 0, (f, bad[:5].tolist(), float(np.abs(got[f].astype(np.float64) - ref[f]).max()))


@pytest.mark.gpu
@pytest.mark.parametrize("model", [1, 2])
def test_after_prognostic_tiny(model):
    bl = [domain.Building(3, 4, 3, 5, 6.5), domain.Building(6, 7, 4, 6, 3.2)]
    c = domain.build_case("two", 10, 10, bl, vertcellsff=10, direction=235.0, turbulence_model=model, iterations=3)
    _compare(c, 3, 2)


@pytest.mark.gpu
def test_diagnostic_only_tiny():
    c = domain.tiny_case(n=10, vertcellsff=9, turbulence_model=1, iterations=1)
    _compare(c, 0, 1)


@pytest.mark.gpu
def test_terrain_variant_tiny():
    c = domain.tiny_case(n=10, vertcellsff=9, turbulence_model=1, iterations=2)
    c.dyk = np.float32(2.5)
    _compare(c, 2, 2, terrain=True)


@pytest.mark.gpu
def test_after_prognostic_config1():
    _compare(domain.config1(iterations=4), 4, 2)
