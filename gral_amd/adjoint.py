"""Sensitivities of receptor winds through the library's adjoint.

The C ABI's ``gral_pf_forward_ckpt`` / ``gral_pf_vjp`` pair differentiates
the pressure-correction part of the SIMPLE iteration (divergence, pressure
GS-TDMA sweep, velocity correction: PrognosticFlowfield.cs:723-949) over an
unrolled window of steps, as a linear map of the velocity fields with the
geometry and DTIME of the prepared situation held fixed.  This module builds
the functional the configuration's adjoint asks for (the mean wind speed at
seeded receptor points about 3 m above ground), its gradient with respect to
the fields, and the contraction of a field gradient onto the inflow speed
profile the initial field was built from.
"""
from __future__ import annotations

import numpy as np


def receptors(case, count=64, height=3.0, seed=20264000):
    """``count`` seeded receptor cells (i, j, k) in free sub-domain columns
    (ADVDOM = 1, no building, 3 <= i, j <= n-2) at the level whose centre
    lies nearest ``height`` m above the ground."""
    rng = np.random.default_rng(seed)
    free = np.argwhere((case.advdom[3:case.nii - 1, 3:case.njj - 1] == 1) &
                       (case.cutk[3:case.nii - 1, 3:case.njj - 1] == 0)) + 3
    pick = free[rng.choice(len(free), size=min(count, len(free)), replace=False)]
    out = []
    for i, j in pick:
        kk = int(case.kkart[i, j])
        ground = float(case.ahk[i, j]) if kk > 0 else float(case.hokart[kk])
        zc = case.hokart[:case.nkk] + 0.5 * case.dzk[1:case.nkk + 1] - ground   # level k+1 centre above ground
        k = int(np.argmin(np.abs(zc[kk:] - height))) + kk + 1
        out.append((int(i), int(j), k))
    return np.array(out, dtype=np.int64)


def _centre(f, rec):
    i, j, k = rec[:, 0], rec[:, 1], rec[:, 2]
    u = 0.5 * (f["uk"][i, j, k].astype(np.float64) + f["uk"][i + 1, j, k])
    v = 0.5 * (f["vk"][i, j, k].astype(np.float64) + f["vk"][i, j + 1, k])
    w = 0.5 * (f["wk"][i, j, k].astype(np.float64) + f["wk"][i, j, k + 1])
    return u, v, w


def mean_receptor_speed(fields, rec):
    """Mean over the receptors of the cell-centre wind speed |(u, v, w)|."""
    u, v, w = _centre(fields, rec)
    return float(np.sqrt(u * u + v * v + w * w).mean())


def mean_receptor_speed_grad(fields, rec):
    """dL/d(UK, VK, WK) of ``mean_receptor_speed`` (float64 arrays)."""
    u, v, w = _centre(fields, rec)
    s = np.sqrt(u * u + v * v + w * w)
    s = np.where(s > 0, s, 1.0)
    n = len(rec)
    g = {f: np.zeros(fields[f].shape, dtype=np.float64) for f in ("uk", "vk", "wk")}
    i, j, k = rec[:, 0], rec[:, 1], rec[:, 2]
    for q in range(n):
        cu, cv, cw = 0.5 * u[q] / s[q] / n, 0.5 * v[q] / s[q] / n, 0.5 * w[q] / s[q] / n
        g["uk"][i[q], j[q], k[q]] += cu
        g["uk"][i[q] + 1, j[q], k[q]] += cu
        g["vk"][i[q], j[q], k[q]] += cv
        g["vk"][i[q], j[q] + 1, k[q]] += cv
        g["wk"][i[q], j[q], k[q]] += cw
        g["wk"][i[q], j[q], k[q] + 1] += cw
    return g["uk"], g["vk"], g["wk"]


def inflow_profile_gradient(case, gu, gv):
    """dL/ds(k) for an initial field of the form (UK, VK) = s(k) (cu, cv) m(i, j, k)
    with a fixed direction (cu, cv) and a 0/1 mask m (the power-law set-up on flat
    terrain without near-wall damping): sum over the columns of
    (UK gu + VK gv) / s(k), where s(k) is the largest |(UK, VK)| of level k."""
    uk, vk = case.uk.astype(np.float64), case.vk.astype(np.float64)
    s = np.sqrt(uk * uk + vk * vk).max(axis=(0, 1))
    num = (uk * gu + vk * gv).sum(axis=(0, 1))
    return np.where(s > 0, num / np.where(s > 0, s, 1.0), 0.0)
