"""Inputs for the flat-terrain set-up tests: small random cities, vertical
grids and meteorological profiles, and a vectorised closed form of the
initial field that the serial oracle is checked against."""
import numpy as np

F = np.float32


def vertical_grid(nkk, dz1=2.0, stretch=1.05):
    dzk = np.zeros(nkk + 2, F)
    hok = np.zeros(nkk + 2, F)
    for k in range(1, nkk + 2):
        dzk[k] = F(dz1 * stretch ** (k - 1))
        hok[k] = F(hok[k - 1] + dzk[k])
    return dzk, hok


def city(nii, njj, seed, density=0.08, hmax=14.0):
    r = np.random.default_rng(seed)
    cutk = np.zeros((nii + 2, njj + 2), F)
    n = max(1, int(density * nii * njj / 6))
    for _ in range(n):
        i0, j0 = r.integers(1, nii + 1), r.integers(1, njj + 1)
        w, d = r.integers(1, 4), r.integers(1, 4)
        h = F(r.uniform(0.5, hmax))
        cutk[i0:min(nii, i0 + w) + 1, j0:min(njj, j0 + d) + 1] = np.maximum(cutk[i0:min(nii, i0 + w) + 1,
                                                                                j0:min(njj, j0 + d) + 1], h)
    return cutk


def met(kind, seed=0, nii=None, njj=None, near_wall=0):
    r = np.random.default_rng(seed)
    if kind == "single":
        d = np.deg2rad(r.uniform(0, 360))
        s = r.uniform(0.5, 6.0)
        return {"n_profile": 1, "meas_height": [0.0, 10.0], "obs_u": [0.0, F(s * np.cos(d))],
                "obs_v": [0.0, F(s * np.sin(d))], "ob_length": F(r.choice([-50.0, -500.0, 0.0, 120.0, 1000.0])),
                "ol_column": None, "near_wall": near_wall}
    if kind == "profile3":
        return {"n_profile": 3, "meas_height": [0.0, 6.0, 17.0, 35.0], "obs_u": [0.0, 1.5, 2.25, 3.5],
                "obs_v": [0.0, -0.5, -1.25, -2.0], "ob_length": F(-80.0), "ol_column": None, "near_wall": near_wall}
    if kind == "columns":
        m = met("single", seed, near_wall=near_wall)
        ol = np.zeros((nii + 2, njj + 2), F)
        ol[1:nii + 1, 1:njj + 1] = r.choice(np.array([-30.0, -300.0, 60.0, 900.0, 0.0], F), size=(nii, njj))
        m["ol_column"] = ol
        return m
    raise ValueError(kind)


def closed_form_blocking(nkk, dzk, hok, cutk):
    """KKART, AHK and the blocked-level mask from the definition: a level is
    blocked where its centre is at or below CUTK."""
    vertk = np.zeros(nkk + 2, F)
    vertk[1:nkk + 1] = (hok[0:nkk] + dzk[1:nkk + 1] * F(0.5)).astype(F)
    blocked = np.zeros(cutk.shape + (nkk + 2,), bool)
    blocked[..., 1:nkk + 1] = vertk[None, None, 1:nkk + 1] <= cutk[..., None]
    kk = np.where(blocked.any(-1), (blocked * np.arange(nkk + 2)).max(-1), 0).astype(np.int16)
    ahk = np.where(kk > 0, hok[kk], F(0)).astype(F)
    return vertk, blocked, kk, ahk
