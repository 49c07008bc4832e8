"""The closed-form checks of tests/test_oracle_pieces.py catch planted mistakes.

pf_oracle.c is built with -DORC_MUTANT=n, each n planting one plausible
misreading of the reference in the oracle (a dropped factor, the wrong height,
a swapped upwind side or sign, an off-by-one level, a missing term, a rule
applied regardless of the flow direction).  The check that covers that part
of the oracle must fail on the mutant; on the real oracle it passes."""
import os
import subprocess

import pytest

import tests.test_oracle_pieces as pieces

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(os.path.dirname(HERE), "oracle", "pf_oracle.c")

# mutant -> the check that must catch it (with its parameters)
CASES = {
    1: (pieces.test_eddy_viscosity_of_linear_shear, dict(comp=0, cov=0.0)),
    2: (pieces.test_eddy_viscosity_of_linear_shear, dict(comp=0, cov=0.0)),
    3: (pieces.test_fluxes_conductances_and_power_law, dict(sign=1.0)),
    4: (pieces.test_wall_drag_follows_the_log_law, dict(comp=0)),
    5: (pieces.test_w_equation_fluxes_at_the_w_point, {}),
    6: (pieces.test_tke_production_and_buoyancy, {}),
    7: (pieces.test_divergence_of_a_linear_field, {}),
    8: (pieces.test_subdomain_boundary_copies_inflow_only, {}),
}


@pytest.fixture(scope="module")
def mutant_dir(tmp_path_factory):
    return tmp_path_factory.mktemp("oracle_mutants")


def _build(n, d):
    out = os.path.join(str(d), f"liboracle_mut{n}.so")
    subprocess.run(["gcc", "-O1", "-std=c11", "-fPIC", "-shared", "-ffp-contract=off", "-fno-fast-math",
                    f"-DORC_MUTANT={n}", "-o", out, SRC, "-lm"], check=True)
    return out


@pytest.mark.parametrize("n", sorted(CASES))
def test_check_catches_planted_mistake(n, mutant_dir, monkeypatch):
    fn, kw = CASES[n]
    fn(**kw)                                          # passes on the real oracle
    monkeypatch.setattr(pieces, "LIB_PATH", _build(n, mutant_dir))
    with pytest.raises(AssertionError):
        fn(**kw)
