"""Sub-domain edges inside the grid.  Buildings far apart with a small
sub-domain factor leave separate islands of ADVDOM = 1, so the E/W and N/S
boundary conditions (PrognosticFlowfield.cs:595-714) act at interior edges,
next to cells outside every sub-domain, as well as at the grid's edges, and
the sweeps skip the columns between the islands.  Four iterations (every sweep
direction), bit-exact against the oracle."""
import numpy as np
import pytest

from gral_amd import domain
from tests.helpers import FIELDS, diff_report

F32 = np.float32
# (i0, i1, j0, j1, height in m) on a 60 x 60 grid of 2 m cells
BUILDINGS = [(10, 13, 12, 15, 8.0), (35, 38, 40, 44, 12.0), (48, 50, 8, 10, 6.0), (22, 22, 30, 31, 5.0)]


def _case(factor, direction):
    cut = np.zeros((62, 62), dtype=F32)
    for i0, i1, j0, j1, h in BUILDINGS:
        cut[i0:i1 + 1, j0:j1 + 1] = F32(h)
    return domain.flat_case("islands", 60, 60, cut, vertcellsff=12, speed=3.0, direction=direction, iterations=4,
                            factor=factor)


@pytest.mark.gpu
@pytest.mark.parametrize("factor,direction", [(1, 235.0), (2, 100.0)])
def test_subdomain_islands(factor, direction):
    from gral_amd import pf
    from oracle import pyoracle
    c = _case(factor, direction)
    inner = c.advdom[1:61, 1:61]
    # several islands: cells outside every sub-domain between them, and
    # interior edges in both directions
    assert (inner == 1).any() and (inner != 1).sum() > 0.5 * inner.size
    print(f"[subdomain edges] factor {factor}, direction {direction}: {int((inner == 1).sum())} of {inner.size} "
          f"columns in sub-domains, 4 iterations", flush=True)
    got = pf.run(c, iterations=4)
    ref = pyoracle.run(c, iterations=4)
    rep = diff_report(got, ref, FIELDS)
    for f, r in rep.items():
        assert r["n_bits_differ"] == 0, (f, r)
    for k in ("iterations_run", "topwind", "dtime", "frund", "geschw_max"):
        assert got["stats"][k] == ref["stats"][k], k
