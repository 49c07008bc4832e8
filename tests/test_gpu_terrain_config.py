"""The terrain configuration at its full grid size (1000 x 1000 cells, dx = 4
m, 30 prognostic layers over 12 Gaussian hills, buildings at 15 %, vegetation, k-eps),
without its adjoint: four situations set up on the GPU and advanced in lock
step.  The checks here are the properties that need no oracle run: the batch equals plain solves of the same
situations bit for bit, every field stays finite, and the strip kernels run.
(One situation is checked against the oracle in tests/test_gpu_configs.py.)"""
import time

import numpy as np
import pytest

from gral_amd import domain

FIELDS = ("uk", "vk", "wk", "dpm", "turb", "tdiss")


@pytest.mark.gpu
def test_terrain_configuration_full_size():
    from gral_amd import pf
    specs = [(270.0, 3.0, 4), (200.0, 5.0, 2), (45.0, 1.5, 6), (120.0, 2.5, 3)]
    t0 = time.time()
    cases = [domain.config5(direction=d, speed=s, stability=st, iterations=3) for d, s, st in specs]
    print(f"[terrain-config] 4 situations {cases[0].nii}x{cases[0].njj}x{cases[0].nkk} set up in "
          f"{time.time() - t0:.1f} s, kadvmax {[c.kadvmax for c in cases]}", flush=True)
    assert all(c.kadvmax + 2 <= 128 for c in cases)             # the strip kernels hold the columns
    solvers = [pf.PrognosticFlowfield(c) for c in cases]
    try:
        for sv in solvers:
            sv.prepare(iterations=3)
        batch = pf.Batch(solvers)
        batch.iterate(3)
        tm = batch.timing()
        batch.close()
        print(f"[terrain-config] batch of 4, 3 iterations: {tm['ms_total'] / 3:.1f} ms per iteration, "
              f"{tm['cell_iterations'] / (tm['ms_total'] / 1e3):.3g} cell-iterations/s", flush=True)
        got = [solvers[q].download() for q in (0, 3)]
    finally:
        for sv in solvers:
            sv.close()
    for q, g in zip((0, 3), got):
        for f in FIELDS:
            assert np.isfinite(g[f]).all(), (q, f)
        with pf.PrognosticFlowfield(cases[q]) as sv:
            sv.prepare(iterations=3)
            sv.iterate(3)
            ref = sv.download()
        for f in FIELDS:
            bad = np.argwhere(g[f].view(np.uint32) != ref[f].view(np.uint32))
            assert bad.size == 0, (q, f, bad[:3].tolist())
        print(f"[terrain-config] situation {q}: batch equals the plain solve", flush=True)
