"""Per-field error of the FP32 solve against an FP64 run of the same algorithm.

TEST INFRASTRUCTURE ONLY (CPU).  Runs the oracle (FP32, the reference's
arithmetic; the GPU solver equals it bit for bit) and its FP64 shadow
(liboracle_pf64.so: every float widened to double) on the same case and
prints, per field, the relative L2 and max-abs difference over the cells the
solver updates.  Usage: python -m oracle.fp64_error [--model 1|2] [--iterations N]
"""
from __future__ import annotations

import argparse
import os
import sys

import numpy as np

if __package__ in (None, ""):
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))

FIELDS = ("uk", "vk", "wk", "dpm", "dpmnew")


def field_errors(case, iterations):
    from oracle import pyoracle
    a = pyoracle.run(case, iterations=iterations)
    b = pyoracle.run(case, iterations=iterations, fp64=True)
    names = FIELDS + (("turb", "tdiss") if case.turbulence_model != 1 else ())
    out = {}
    for f in names:
        x, y = a[f].astype(np.float64), b[f]
        den = float(np.sqrt((y * y).sum()))
        d = x - y
        out[f] = {"rel_l2": float(np.sqrt((d * d).sum()) / den) if den > 0 else float(np.sqrt((d * d).sum())),
                  "max_abs": float(np.abs(d).max())}
    return out


def main():
    from gral_amd import domain
    ap = argparse.ArgumentParser()
    ap.add_argument("--model", type=int, default=1)
    ap.add_argument("--iterations", type=int, default=200)
    a = ap.parse_args()
    c = domain.config1(turbulence_model=a.model, iterations=a.iterations)
    err = field_errors(c, a.iterations)
    print(f"config 1 (100x100, 30 layers, one 20x20x15 m building), model {a.model}, {a.iterations} iterations")
    print("| field | rel. L2 | max abs |")
    print("|---|---|---|")
    for f, e in err.items():
        print(f"| {f} | {e['rel_l2']:.2e} | {e['max_abs']:.2e} |")


if __name__ == "__main__":
    main()
