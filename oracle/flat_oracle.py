"""CPU restatement of GRAL's flat-terrain set-up of one situation.

TEST INFRASTRUCTURE ONLY: the checker of gral_pf_subdomains,
gral_pf_reference_point and gral_pf_flat_init.  The product never imports it.

Serial loops in the reference's own order (paths relative to the reference's
src/):
  * subdomains: ReadGeometryAndDomains.cs:321-349;
  * reference_point: MicroscaleTerrainSearchRefPoint.cs:27-103 and the
    fallback of MicroscaleTerrain.cs:936-963;
  * flat_init: MicroscaleFlat.cs:29-345 (zeroing, the column loop over i then
    j with k from NKK down to 1, the ghost copies inside it, KADVMAX, the
    near-wall damping).
MathF.Pow / MathF.Log are the C library's powf / logf, which is what .NET
calls on Linux; float arithmetic is float32 (numpy scalars), the reference's
double locals hold float products and are cast back unchanged.
"""
from __future__ import annotations

import ctypes
import ctypes.util

import numpy as np

F = np.float32
_m = ctypes.CDLL(ctypes.util.find_library("m"))
_m.powf.restype = ctypes.c_float
_m.powf.argtypes = [ctypes.c_float, ctypes.c_float]
_m.logf.restype = ctypes.c_float
_m.logf.argtypes = [ctypes.c_float]
_m.sqrtf.restype = ctypes.c_float
_m.sqrtf.argtypes = [ctypes.c_float]


def powf(x, y):
    return F(_m.powf(float(x), float(y)))


def logf(x):
    return F(_m.logf(float(x)))


def sqrtf(x):
    return F(_m.sqrtf(float(x)))


def net_max(a, b):
    """Math.Max(float, float) of .NET Core."""
    a, b = F(a), F(b)
    if a != b:
        if not np.isnan(a):
            return a if b < a else b
        return a
    return a if np.signbit(b) else b


def subdomains(cutk, nii, njj, dxk, dyk, factor=15, advdom=None):
    adv = np.zeros((nii + 2, njj + 2), np.uint8) if advdom is None else advdom.copy()
    for i in range(1, nii + 1):
        for j in range(1, njj + 1):
            c = F(cutk[i, j])
            if c > 0:
                iplus = int(F(F(factor) * c) / F(dxk))
                jplus = int(F(F(factor) * c) / F(dyk))
                for i1 in range(i - iplus, i + iplus + 1):
                    for j1 in range(j - jplus, j + jplus + 1):
                        if 1 <= i1 <= nii and 1 <= j1 <= njj:
                            adv[i1, j1] = 1
    return adv


def reference_point(adv):
    xmax, ymax = adv.shape
    searchmax = min(xmax, ymax) // 2
    delta, i, j, found = 1, 1, 1, False
    while not found and delta < searchmax:
        i = j = delta
        while j < ymax - delta and not found:
            if adv[i, j] == 1:
                found = True
            else:
                j += 1
        while i < xmax - delta and not found:
            if adv[i, j] == 1:
                found = True
            else:
                i += 1
        while j > delta and not found:
            if adv[i, j] == 1:
                found = True
            else:
                j -= 1
        while i > delta and not found:
            if adv[i, j] == 1:
                found = True
            else:
                i -= 1
        delta += 1
    return (i, j) if found else (1, 1)


def _exponent(ob):
    ob = F(ob)
    if ob <= 0:
        return net_max(F(0.35) - F(0.4) * powf(abs(ob), F(-0.15)), F(0.05))
    return F(0.56) * powf(ob, F(-0.15))


def flat_init(nii, njj, nkk, vcf, dxk, dyk, dzk, hok, cutk, met, uk, vk, wk, advdom=None, ahk=None, kkart=None,
              bui=None):
    """met: dict(n_profile, meas_height, obs_u, obs_v, ob_length, ol_column, near_wall).
    Returns (uk, vk, wk, kkart, ahk, bui, kadvmax) on copies of the inputs."""
    uk, vk, wk = uk.copy(), vk.copy(), wk.copy()
    n2 = (nii + 2, njj + 2)
    ahk = np.zeros(n2, F) if ahk is None else ahk.copy()
    kkart = np.zeros(n2, np.int16) if kkart is None else kkart.copy()
    bui = np.zeros(n2, F) if bui is None else bui.copy()
    H = [F(x) for x in met["meas_height"]]
    U = [F(x) for x in met["obs_u"]]
    V = [F(x) for x in met["obs_v"]]
    inumm = int(met["n_profile"])
    hok = hok.astype(F)
    dzk = dzk.astype(F)
    # MicroscaleFlat.cs:29, :36-50
    kadvmax = 1
    for i in range(1, nii + 2):
        for j in range(1, njj + 2):
            uk[i, j, 1:nkk + 1] = 0
            vk[i, j, 1:nkk + 1] = 0
            wk[i, j, 1:nkk + 1] = 0
    # :55-194 in serial order
    for i in range(1, nii + 1):
        kadv1 = 1
        for j in range(1, njj + 1):
            ahk[i, j] = 0
            kkart[i, j] = 0
            ob = F(met["ob_length"]) if met.get("ol_column") is None else F(met["ol_column"][i, j])
            for k in range(nkk, 0, -1):
                vertk = F(hok[k - 1] + dzk[k] * F(0.5))
                if vertk > cutk[i, j]:
                    if vertk <= H[1]:
                        e = _exponent(ob)
                        dumfac = powf(vertk / H[1], e)
                        ux, uy = F(U[1] * dumfac), F(V[1] * dumfac)
                    elif vertk > H[inumm]:
                        if inumm == 1:
                            e = _exponent(ob)
                            dumfac = powf(vertk / H[inumm], e)
                            ux, uy = F(U[inumm] * dumfac), F(V[inumm] * dumfac)
                        else:
                            ux, uy = U[inumm], V[inumm]
                    else:
                        ipo = 1
                        for ip in range(1, inumm + 1):
                            if vertk > H[ip]:
                                ipo = ip + 1
                        ux = F(U[ipo - 1] + F(F(U[ipo] - U[ipo - 1]) / F(H[ipo] - H[ipo - 1])) * F(vertk - H[ipo - 1]))
                        uy = F(V[ipo - 1] + F(F(V[ipo] - V[ipo - 1]) / F(H[ipo] - H[ipo - 1])) * F(vertk - H[ipo - 1]))
                    uk[i, j, k] = ux
                    vk[i, j, k] = uy
                    if i > 1 and vertk <= cutk[i - 1, j]:
                        uk[i, j, k] = 0
                    if j > 1 and vertk <= cutk[i, j - 1]:
                        vk[i, j, k] = 0
                    uk[nii + 1, j, k] = uk[nii, j, k]
                    vk[i, njj + 1, k] = vk[i, njj, k]
                else:
                    uk[i, j, k] = 0
                    if i < nii:
                        uk[i + 1, j, k] = 0
                    vk[i, j, k] = 0
                    if j < njj:
                        vk[i, j + 1, k] = 0
                    ahk[i, j] = net_max(hok[k], ahk[i, j])
                    bui[i, j] = ahk[i, j]
                    kkart[i, j] = max(k, int(kkart[i, j]))
                    if cutk[i, j] > 0:
                        kadv1 = max(int(kkart[i, j]), kadv1)
        kadvmax = max(kadvmax, kadv1)
    kadvmax = min(nkk, kadvmax + vcf)
    # :201-345
    mode = int(met.get("near_wall", 0))
    igeb = max(int(F(20) / F(dxk)), 1)
    DX, DY = F(dxk), F(dyk)
    if mode:
        for i in range(1 + igeb, nii - igeb + 1):
            for j in range(1 + igeb, njj - igeb + 1):
                if mode == 2 or advdom[i, j] == 0:
                    for k in range(1, nkk):
                        abmind = F(1)
                        if k > kkart[i, j]:
                            vertk = F(hok[k - 1] + dzk[k] * F(0.5))

                            def hit(ig, jg):
                                return cutk[ig, jg] > 1 and vertk <= ahk[ig, jg]

                            searches = (
                                [(ig, j, lambda ig=ig: abs(F((i - ig) * DX))) for ig in range(i - 1, i - igeb - 1, -1)],
                                [(ig, j, lambda ig=ig: abs(F((i - ig) * DX))) for ig in range(i + 1, i + igeb + 1)],
                                [(i, jg, lambda jg=jg: abs(F((j - jg) * DY))) for jg in range(j + 1, j + igeb + 1)],
                                [(i, jg, lambda jg=jg: abs(F((j - jg) * DY))) for jg in range(j - 1, j - igeb - 1, -1)],
                                [(ig, j + ig - i, None) for ig in range(i + 1, i + igeb + 1)],
                                [(ig, j + ig - i, None) for ig in range(i - 1, i - igeb - 1, -1)],
                                [(ig, j + i - ig, None) for ig in range(i + 1, i + igeb + 1)],
                                [(ig, j + i - ig, None) for ig in range(i - 1, i - igeb - 1, -1)],
                            )
                            for cells in searches:
                                entf = F(21)
                                for ig, jg, dist in cells:
                                    if hit(ig, jg):
                                        if dist is not None:
                                            entf = dist()
                                        else:
                                            a = F((i - ig) * DX)
                                            b = F((j - jg) * DY)
                                            entf = sqrtf(F(a * a) + F(b * b))
                                        break
                                if entf <= 20:
                                    abmind = F(abmind * F(F(0.19) * logf(F(F(entf + F(0.5)) * F(10)))))
                            uk[i, j, k] = F(uk[i, j, k] * abmind)
                            vk[i, j, k] = F(vk[i, j, k] * abmind)
    return uk, vk, wk, kkart, ahk, bui, kadvmax
