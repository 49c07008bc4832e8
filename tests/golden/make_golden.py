"""Writes the golden fixtures of the prognostic flow solver (tests/golden/).

The reference ships no golden vectors, known-answer tests or fixtures for
this path, and its C# cannot run here, so the fixtures are the outputs of the
CPU oracle (oracle/pf_oracle.c, itself pinned by the independent numpy
transcription and the closed-form checks in tests/) on seeded synthetic
inputs from gral_amd.domain.  They freeze those outputs so that
  * the GPU tests check the HIP solver against fixed data without running the
    oracle (config 1 at 200 iterations included), and
  * the CPU tests notice any drift of the oracle or of the input generator.

Fixture files:
  tiny_m{model}[_veg].npz  -- full output fields, levels 0..KADVMAX+1, of a
                              few-cell domain (inputs regenerated from the spec)
  golden.json              -- per case: generator spec, SHA-256 of the inputs
                              and of every output field (float32 bit patterns,
                              levels 0..KADVMAX+1), and the solver statistics;
                              for the post-pass cases (key "diagnostic") the
                              digests cover the full column height

Run from the repository root:  python tests/golden/make_golden.py
"""
from __future__ import annotations

import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

FIELDS = ("uk", "vk", "wk", "uks", "vks", "wks", "dpm", "dpmnew")
KE_FIELDS = ("turb", "tdiss")
STATS = ("iteration_loops", "iterations_run", "delta_first", "delta_finish", "topwind", "dtime", "frund",
         "geschw_max", "unreal_u", "unreal_v", "unreal_w", "convergence_problem", "ug", "vg")

# (name, generator, kwargs, iterations, keep full arrays)
CASES = [
    ("tiny_m1", "tiny_case", dict(n=9, vertcellsff=8, turbulence_model=1), 3, True),
    ("tiny_m2", "tiny_case", dict(n=9, vertcellsff=8, turbulence_model=2), 3, True),
    ("tiny_m0", "tiny_case", dict(n=9, vertcellsff=8, turbulence_model=0), 3, True),
    ("tiny_m1_veg", "tiny_case", dict(n=10, vertcellsff=9, turbulence_model=1, veg=True), 4, True),
    ("c1_m1_it1", "config1", dict(turbulence_model=1), 1, False),
    ("c1_m1_it2", "config1", dict(turbulence_model=1), 2, False),
    ("c1_m1_it4", "config1", dict(turbulence_model=1), 4, False),
    ("c1_m1_it200", "config1", dict(turbulence_model=1), 200, False),
    ("c1_m2_it4", "config1", dict(turbulence_model=2), 4, False),
    # the 512-bit mixing-length solvers (simd_width 16)
    ("tiny_m1_w16", "tiny_case", dict(n=9, vertcellsff=16, turbulence_model=1, simd_width=16), 3, True),
    ("c1_m1_w16_it4", "config1", dict(turbulence_model=1, simd_width=16), 4, False),
]


# mass-conservation post-pass after the prognostic iterations (full-height digests)
# (name, generator, kwargs, prognostic iterations, flow_field_level, terrain)
DIAG_CASES = [
    ("c1_m1_it4_diag2", "config1", dict(turbulence_model=1), 4, 2, False),
    ("tiny_m1_diag1", "tiny_case", dict(n=10, vertcellsff=9, turbulence_model=1), 0, 1, False),
]
DIAG_FIELDS = ("uk", "vk", "wk", "dpm")


def make_case(gen, kwargs, iterations):
    from gral_amd import domain
    kw = dict(kwargs)
    width = kw.pop("simd_width", None)
    c = getattr(domain, gen)(iterations=iterations, **kw)
    if width is not None:
        c.simd_width = width
    return c


def crop(a, kadvmax):
    return np.ascontiguousarray(a[:, :, :kadvmax + 2])


def digest(a):
    return hashlib.sha256(np.ascontiguousarray(a).view(np.uint8).tobytes()).hexdigest()


def input_digest(c):
    h = hashlib.sha256()
    for a in (c.uk, c.vk, c.wk, c.advdom, c.kkart, c.cutk, c.dzk, c.hokart, c.z0_wall, c.z0_solver):
        h.update(np.ascontiguousarray(a).view(np.uint8).tobytes())
    return h.hexdigest()


def fields_of(model):
    return FIELDS + (KE_FIELDS if model != 1 else ())


def main():
    from oracle import pyoracle
    meta = {}
    for name, gen, kwargs, its, full in CASES:
        c = make_case(gen, kwargs, its)
        r = pyoracle.run(c, iterations=its, keep_history=False)
        flds = fields_of(c.turbulence_model)
        meta[name] = {
            "generator": gen, "kwargs": kwargs, "iterations": its, "kadvmax": int(c.kadvmax),
            "input_sha256": input_digest(c),
            "field_sha256": {f: digest(crop(r[f], c.kadvmax)) for f in flds},
            "stats": {k: (float(r["stats"][k]) if isinstance(r["stats"][k], (float, np.floating)) else int(r["stats"][k]))
                      for k in STATS},
        }
        if full:
            np.savez_compressed(os.path.join(HERE, f"{name}.npz"), **{f: crop(r[f], c.kadvmax) for f in flds})
        print(name, meta[name]["stats"]["iteration_loops"], flush=True)
    for name, gen, kwargs, its, level, terrain in DIAG_CASES:
        c = make_case(gen, kwargs, max(its, 1))
        f = None
        if its:
            r = pyoracle.run(c, iterations=its, keep_history=False)
            f = {k: r[k] for k in DIAG_FIELDS}
        d = pyoracle.diagnostic(c, f, flow_field_level=level, terrain=terrain)
        meta[name] = {
            "generator": gen, "kwargs": kwargs, "iterations": its, "kadvmax": int(c.kadvmax), "diagnostic": level,
            "terrain": terrain, "input_sha256": input_digest(c),
            "field_sha256": {k: digest(d[k]) for k in DIAG_FIELDS}, "dtime": float(d["dtime"]),
        }
        print(name, flush=True)
    with open(os.path.join(HERE, "golden.json"), "w") as fh:
        json.dump(meta, fh, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
