"""Shared comparison helpers for the parity tests."""
import numpy as np

FIELDS = ("uk", "vk", "wk", "uks", "vks", "wks", "dpm", "dpmnew")
KE_FIELDS = ("turb", "tdiss")


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.sqrt((b * b).sum())
    return float(np.sqrt(((a - b) ** 2).sum()) / den) if den > 0 else float(np.sqrt(((a - b) ** 2).sum()))


def diff_report(got, ref, fields):
    rep = {}
    for f in fields:
        g, r = got[f], ref[f]
        nbad = int((g.view(np.uint32) != r.view(np.uint32)).sum())
        rep[f] = dict(n_bits_differ=nbad, max_abs=float(np.abs(g.astype(np.float64) - r).max()), rel_l2=rel_l2(g, r))
    return rep
