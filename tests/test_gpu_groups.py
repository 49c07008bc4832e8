"""Batches run in stream groups: situations split round-robin into up to 4
groups, each on a stream of its own, with the group count chosen online (one
timed cycle of the four sweep directions per candidate count, then the largest
count within 2 % of the fastest) unless GRAL_PF_GROUPS fixes it.  The grouping
must not change a bit: every situation equals the oracle's plain solve of it,
for the tuned count, for each forced count, and when situations retire early
and the group count's candidates change."""
import os

import pytest

from gral_amd import domain
from tests.test_gpu_batch import _check_against_oracle, _run_batch


def _cities(count, iterations):
    out = []
    for q in range(count):
        bl = domain.city_buildings(20, 20, 2.0, 40 + q, density=0.2, fmin=4.0, fmax=10.0, hmin=3.0, hmax=9.0,
                                   margin=2)
        out.append(domain.build_case(f"g{q}", 20, 20, bl, dx=2.0, vertcellsff=8, speed10=3.0 + 0.3 * q,
                                     direction=45.0 * q, z0=0.5, iterations=iterations))
    return out


@pytest.fixture
def groups_env():
    old = os.environ.get("GRAL_PF_GROUPS")
    yield
    if old is None:
        os.environ.pop("GRAL_PF_GROUPS", None)
    else:
        os.environ["GRAL_PF_GROUPS"] = old


@pytest.mark.gpu
@pytest.mark.parametrize("forced", [None, "1", "2", "3", "4"])
def test_groups_bit_identical(forced, groups_env):
    if forced is None:
        os.environ.pop("GRAL_PF_GROUPS", None)
    else:
        os.environ["GRAL_PF_GROUPS"] = forced
    print(f"[groups] 8 situations, 9 iterations, GRAL_PF_GROUPS={forced}", flush=True)
    cases = _cities(8, 9)
    fields, stats, _ = _run_batch(cases, iterations=[9] * 8)
    _check_against_oracle(cases, fields, stats, iterations=[9] * 8)


@pytest.mark.gpu
def test_groups_with_early_retirement(groups_env):
    """Six situations, of which some stop at their iteration limit after 3
    iterations: the group count is re-chosen for the remaining ones."""
    os.environ.pop("GRAL_PF_GROUPS", None)
    iters = [3, 12, 3, 12, 3, 12]
    cases = _cities(6, 12)
    for c, n in zip(cases, iters):
        c.iter_min = c.iter_max = n
    print("[groups] 6 situations, three retire after 3 iterations", flush=True)
    fields, stats, _ = _run_batch(cases, iterations=iters)
    _check_against_oracle(cases, fields, stats, iterations=iters)
