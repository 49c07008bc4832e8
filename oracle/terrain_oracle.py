"""CPU restatement of GRAL's terrain set-up of one situation from a GRAMM
field (MicroscaleTerrain.cs:57-316, :600-612 and the near-wall damping
:664-832, paths relative to the reference's src/), and of its variant with
the original GRAL topography (:318-598).

TEST INFRASTRUCTURE ONLY: the checker of gral_pf_terrain_init.  The product
never imports it.

Serial loops in the reference's order (columns i ascending, then j, levels k
from NKK down, GRAMM levels ik from NK down).  The reference runs the i loop
in parallel; where a column reads a neighbour's result (KKART of column i-1)
or a ghost copy, the serial order is the one restated.  Float arithmetic is
float32 and the reference's double expressions are float64, operation by
operation.
"""
from __future__ import annotations

import numpy as np

from oracle.flat_oracle import logf, net_max, sqrtf

F = np.float32
D = np.float64


def net_min(a, b):
    """Math.Min(float, float) of .NET Core."""
    a, b = F(a), F(b)
    if a != b:
        if not np.isnan(a):
            return a if a < b else b
        return a
    return a if np.signbit(a) else b


def _p2(x):
    x = F(x)
    return F(x * x)


def _clamp(v, lo, hi):
    return max(lo, min(hi, v))


def weights(xwert, ywert, ddx, ddy):
    """r1..r16 of the 16-point inverse-distance weighting (MicroscaleTerrain.cs:96-111)."""
    xw, yw, X, Y = F(xwert), F(ywert), F(ddx), F(ddy)
    xd, yd = F(X - xw), F(Y - yw)
    pts = [(xw, yw), (xw, yd), (xd, yd), (xd, yw), (F(X + xw), F(Y + yw)), (xw, F(Y + yw)), (xd, F(Y + yw)),
           (F(X + xd), F(Y + yw)), (F(X + xd), yw), (F(X + xd), yd), (F(X + xd), F(Y + yd)), (xd, F(Y + yd)),
           (xw, F(Y + yd)), (F(X + xw), F(Y + yd)), (F(X + xw), yd), (F(X + xw), yw)]
    return [F(F(1) / _p2(F(F(_p2(a) + _p2(b)) + F(0.1)))) for a, b in pts]


# AHE neighbours of r1..r16, as (di, dj) from (indi, indj)
NEIGH = [(0, 0), (0, 1), (1, 1), (1, 0), (-1, -1), (0, -1), (1, -1), (2, -1), (2, 0), (2, 1), (2, 2), (1, 2), (0, 2),
         (-1, 2), (-1, 1), (-1, 0)]


def terrain_init(nii, njj, nkk, vcf, dxk, dyk, dzk, hok, cutk, advdom, gr, uk, vk, wk, kkart, ahk, bui, ahmin,
                 near_wall=0):
    """gr: dict(nx, ny, nz, ddx, ddy, delta_i, delta_j, ahe, uwin, vwin, wwin, zsp, ah, zko1, zko2).
    Returns (uk, vk, wk, kkart, ahk, bui, ahmin_new, kadvmax, (vec_x, vec_y, vec_n))."""
    uk, vk, wk = uk.copy(), vk.copy(), wk.copy()
    kkart, ahk, bui = kkart.copy(), ahk.copy(), bui.copy()
    NX, NY, NK = gr["nx"], gr["ny"], gr["nz"]
    DDX1, DDY1 = F(gr["ddx"]), F(gr["ddy"])
    AHE, UW, VW, WW, ZSP, AH = gr["ahe"], gr["uwin"], gr["vwin"], gr["wwin"], gr["zsp"], gr["ah"]
    AHMIN = F(ahmin)
    DX, DY = F(dxk), F(dyk)
    hok, dzk = hok.astype(F), dzk.astype(F)
    vec_x = vec_y = 0.0
    vec_n = 0
    kadvmax = 1
    zlim = D(10) + (D(gr["zko2"]) - D(gr["zko1"])) * D(F(0.5))
    for i in range(1, nii + 1):
        # the 10 m wind is summed per row i (thread-local in the reference's
        # parallel loop over i, :67) and the row sums are added in i order (:311-313)
        lx = ly = 0.0
        ln = 0
        for j in range(1, njj + 1):
            ahk[i, j] = AHMIN
            kkart[i, j] = 0
            indi = int((D(F(gr["delta_i"]) + F(DX * F(i))) - D(DX) * 0.5) / D(DDX1)) + 1
            indj = int((D(F(gr["delta_j"]) + F(DY * F(j))) - D(DY) * 0.5) / D(DDY1)) + 1
            xwert = F(F(F(F(DX * F(i)) - F(DX * F(0.5))) + F(gr["delta_i"])) - F(F(indi - 1) * DDX1))
            ywert = F(F(F(F(DY * F(j)) - F(DY * F(0.5))) + F(gr["delta_j"])) - F(F(indj - 1) * DDY1))
            ahe_l = AHE[_clamp(indi, 1, NX), _clamp(indj, 1, NY)]
            interior = 2 <= indi <= NX - 1 and 2 <= indj <= NY - 1
            ch = np.zeros(NK + 2, F)
            chp = np.zeros(NK + 2, F)
            if interior:
                r = weights(xwert, ywert, DDX1, DDY1)
                gew = F(0)
                for q in range(16):
                    gew = F(gew + r[q])
                for ik in range(NK, 0, -1):
                    ikp = min(ik + 1, NK)
                    s = F(0)
                    sp = F(0)
                    for q, (di, dj) in enumerate(NEIGH):
                        s = F(s + F(AHE[indi + di, indj + dj, ik] * r[q])) if q else F(AHE[indi + di, indj + dj, ik] * r[q])
                        sp = F(sp + F(AHE[indi + di, indj + dj, ikp] * r[q])) if q else F(AHE[indi + di, indj + dj, ikp] * r[q])
                    ch[ik] = F(F(s / gew) - AHMIN)
                    chp[ik] = F(F(sp / gew) - AHMIN)
            counted = False
            for k in range(nkk, 0, -1):
                vertk = D(hok[k - 1]) + D(dzk[k]) * 0.5 + D(AHMIN)
                indk = 0
                indk_p = 0
                c_h = c_hp = dz = F(0)
                for ik in range(NK, 0, -1):
                    ikp = min(ik + 1, NK)
                    if not interior:
                        c_h = F(ahe_l[ik] - AHMIN)
                        c_hp = F(ahe_l[ikp] - AHMIN)
                    else:
                        c_h, c_hp = ch[ik], chp[ik]
                    if vertk - D(AHMIN) > D(F(c_h + cutk[i, j])):
                        indk = ik
                        indk_p = min(indk + 1, NK)
                        dz = F(F(vertk - D(AHMIN)) - F(c_h + cutk[i, j]))
                        break
                if indk > 0:
                    indi = _clamp(indi, 1, NX)
                    indj = _clamp(indj, 1, NY)
                    ixm = _clamp(indi - 1, 1, NX) if xwert < F(DDX1 * F(0.5)) else _clamp(indi + 1, 1, NX)
                    iym = _clamp(indj - 1, 1, NY) if ywert < F(DDY1 * F(0.5)) else _clamp(indj + 1, 1, NY)
                    fx = F(xwert - F(DDX1 * F(0.5)))
                    fy = F(ywert - F(DDY1 * F(0.5)))
                    den = net_max(F(c_hp - c_h), F(0.1))

                    def interp(W):
                        def hx(jj, kk_):
                            a = W[indi, jj, kk_]
                            return F(a + F(F(F(F(W[ixm, jj, kk_] - a) / DDX1) * F(ixm - indi)) * fx))
                        x1, x2 = hx(indj, indk), hx(iym, indk)
                        x3, x4 = hx(indj, indk_p), hx(iym, indk_p)
                        x5 = F(x1 + F(F(F(x3 - x1) / den) * dz))
                        x6 = F(x2 + F(F(F(x4 - x2) / den) * dz))
                        return F(x5 + F(F(F(F(x6 - x5) / DDY1) * F(iym - indj)) * fy))

                    uk[i, j, k] = interp(UW)
                    vk[i, j, k] = interp(VW)
                    wk[i, j, k] = interp(WW)
                    if D(F(ZSP[indi, indj, indk] - AH[indi, indj])) < zlim and not counted:
                        counted = True
                        lx += D(uk[i, j, k])
                        ly += D(vk[i, j, k])
                        ln += 1
                    if i > 1 and k <= kkart[i - 1, j]:
                        uk[i, j, k] = 0
                    if j > 1 and k <= kkart[i, j - 1]:
                        vk[i, j, k] = 0
                    uk[nii + 1, j, k] = uk[nii, j, k]
                    vk[i, njj + 1, k] = vk[i, njj, k]
                else:
                    uk[i, j, k] = 0
                    vk[i, j, k] = 0
                    wk[i, j, k] = 0
                    if i < nii:
                        uk[i + 1, j, k] = 0
                    if j < njj:
                        vk[i, j + 1, k] = 0
                    if k < nkk:
                        wk[i, j, k + 1] = 0
                    ahk[i, j] = net_max(F(hok[k] + AHMIN), ahk[i, j])
                    if cutk[i, j] > 0:
                        bui[i, j] = net_max(F(F(ahk[i, j] - AHMIN) - c_h), bui[i, j])
                    kkart[i, j] = max(k, int(kkart[i, j]))
            if advdom[i, j] > 0:
                kadvmax = max(int(kkart[i, j]), kadvmax)
        vec_x += lx
        vec_y += ly
        vec_n += ln
    return _epilogue(nii, njj, nkk, vcf, dxk, dyk, dzk, hok, cutk, advdom, uk, vk, wk, kkart, ahk, bui, kadvmax,
                     (vec_x, vec_y, vec_n), near_wall)


def terrain_init_topo(nii, njj, nkk, vcf, dxk, dyk, dzk, hok, cutk, advdom, gr, uk, vk, wk, kkart, ahk, bui, ahmin,
                      ahko, ahko_min, near_wall=0):
    """The variant with the original GRAL topography (GRAL_topofile.txt),
    MicroscaleTerrain.cs:318-598: a GRAL level is blocked when its centre
    (float, above AHKOriMin) is at or below AHKOri + CUTK of the column; the
    free levels take the GRAMM wind at the highest GRAMM level whose height
    (absolute, interpolated, in the interior; relative to the incoming AHMIN
    at the GRAMM border, as the reference has it) plus CUTK lies below the
    centre, or at GRAMM level 1 with the level below it (index 0) when none
    does.  The ghost u of row NII+1 (v of column NJJ+1) is copied while
    column i = 1 (j = 1) is set, in the serial order restated here.  The
    10 m wind counts the highest free level below AHKOri + 10 m + DZK/2, its
    u summed in double and v in float per row i.  gr as for terrain_init
    (zsp, ah and zko are not read); ahko [nii+2][njj+2]."""
    uk, vk, wk = uk.copy(), vk.copy(), wk.copy()
    kkart, ahk, bui = kkart.copy(), ahk.copy(), bui.copy()
    NX, NY, NK = gr["nx"], gr["ny"], gr["nz"]
    DDX1, DDY1 = F(gr["ddx"]), F(gr["ddy"])
    AHE, UW, VW, WW = gr["ahe"], gr["uwin"], gr["vwin"], gr["wwin"]
    AHMIN, AOM = F(ahmin), F(ahko_min)
    DX, DY = F(dxk), F(dyk)
    hok, dzk = hok.astype(F), dzk.astype(F)
    dI, dJ = F(gr["delta_i"]), F(gr["delta_j"])
    vec_x = vec_y = 0.0
    vec_n = 0
    kadvmax = 1
    for i in range(1, nii + 1):
        lx, ly, ln = 0.0, F(0), 0
        for j in range(1, njj + 1):
            ahk[i, j] = AOM
            kkart[i, j] = 0
            indi = int(F(F(F(dI + F(DX * F(i))) - F(DX * F(0.5))) / DDX1)) + 1
            indj = int(F(F(F(dJ + F(DY * F(j))) - F(DY * F(0.5))) / DDY1)) + 1
            xwert = F(F(F(F(DX * F(i)) - F(DX * F(0.5))) + dI) - F(F(indi - 1) * DDX1))
            ywert = F(F(F(F(DY * F(j)) - F(DY * F(0.5))) + dJ) - F(F(indj - 1) * DDY1))
            ahe_l = AHE[_clamp(indi, 1, NX), _clamp(indj, 1, NY)]
            interior = 2 <= indi <= NX - 1 and 2 <= indj <= NY - 1
            ch = np.zeros(NK + 2, F)
            chp = np.zeros(NK + 2, F)
            if interior:
                r = weights(xwert, ywert, DDX1, DDY1)
                gew = F(0)
                for q in range(16):
                    gew = F(gew + r[q])
                for ik in range(NK, 0, -1):
                    ikp = min(ik + 1, NK)
                    sm = F(0)
                    sp = F(0)
                    for q, (di, dj) in enumerate(NEIGH):
                        sm = F(sm + F(AHE[indi + di, indj + dj, ik] * r[q])) if q else F(AHE[indi + di, indj + dj, ik] * r[q])
                        sp = F(sp + F(AHE[indi + di, indj + dj, ikp] * r[q])) if q else F(AHE[indi + di, indj + dj, ikp] * r[q])
                    ch[ik] = F(sm / gew)
                    chp[ik] = F(sp / gew)
            counted = False
            ground = F(ahko[i, j] + cutk[i, j])
            for k in range(nkk, 0, -1):
                vertk = F(F(hok[k - 1] + F(dzk[k] * F(0.5))) + AOM)
                indk, indz, indk_p = 1, 1, 0
                c_h = c_hp = dz = F(0)
                for ik in range(NK, 0, -1):
                    ikp = min(ik + 1, NK)
                    if not interior:
                        c_h = F(ahe_l[ik] - AHMIN)
                        c_hp = F(ahe_l[ikp] - AHMIN)
                    else:
                        c_h, c_hp = ch[ik], chp[ik]
                    if vertk <= ground:
                        indz = 0
                        indk = ik
                        break
                    if vertk > F(c_h + cutk[i, j]):
                        indk = ik
                        indk_p = min(indk + 1, NK)
                        dz = F(vertk - F(c_h + cutk[i, j]))
                        break
                if indz > 0:
                    indi = _clamp(indi, 1, NX)
                    indj = _clamp(indj, 1, NY)
                    ixm = _clamp(indi - 1, 1, NX) if xwert < F(DDX1 * F(0.5)) else _clamp(indi + 1, 1, NX)
                    iym = _clamp(indj - 1, 1, NY) if ywert < F(DDY1 * F(0.5)) else _clamp(indj + 1, 1, NY)
                    fx = F(xwert - F(DDX1 * F(0.5)))
                    fy = F(ywert - F(DDY1 * F(0.5)))
                    den = net_max(F(c_hp - c_h), F(0.1))

                    def interp(W):
                        def hx(jj, kk_):
                            a = W[indi, jj, kk_]
                            return F(a + F(F(F(F(W[ixm, jj, kk_] - a) / DDX1) * F(ixm - indi)) * fx))
                        x1, x2 = hx(indj, indk), hx(iym, indk)
                        x3, x4 = hx(indj, indk_p), hx(iym, indk_p)
                        x5 = F(x1 + F(F(F(x3 - x1) / den) * dz))
                        x6 = F(x2 + F(F(F(x4 - x2) / den) * dz))
                        return F(x5 + F(F(F(F(x6 - x5) / DDY1) * F(iym - indj)) * fy))

                    uk[i, j, k] = interp(UW)
                    vk[i, j, k] = interp(VW)
                    wk[i, j, k] = interp(WW)
                    if i > 1 and k <= kkart[i - 1, j]:
                        uk[i, j, k] = 0
                    if j > 1 and k <= kkart[i, j - 1]:
                        vk[i, j, k] = 0
                    if i == 1:
                        uk[nii + 1, j, k] = uk[nii, j, k]
                    if j == 1:
                        vk[i, njj + 1, k] = vk[i, njj, k]
                    if vertk < F(F(ahko[i, j] + F(10)) + F(F(0.5) * dzk[k])) and not counted:
                        counted = True
                        lx += D(uk[i, j, k])
                        ly = F(ly + vk[i, j, k])
                        ln += 1
                else:
                    uk[i, j, k] = 0
                    vk[i, j, k] = 0
                    wk[i, j, k] = 0
                    if i < nii:
                        uk[i + 1, j, k] = 0
                    if j < njj:
                        vk[i, j + 1, k] = 0
                    if k < nkk:
                        wk[i, j, k + 1] = 0
                    ahk[i, j] = net_max(F(hok[k] + AOM), ahk[i, j])
                    kkart[i, j] = max(k, int(kkart[i, j]))
                    if advdom[i, j] > 0:
                        kadvmax = max(int(kkart[i, j]), kadvmax)
                    if cutk[i, j] > 0 and vertk >= ahko[i, j] and k > 1:
                        bui[i, j] = net_max(F(ahk[i, j] - F(hok[k - 1] + AOM)), bui[i, j])
        vec_x += lx
        vec_y += D(ly)
        vec_n += ln
    return _epilogue(nii, njj, nkk, vcf, dxk, dyk, dzk, hok, cutk, advdom, uk, vk, wk, kkart, ahk, bui, kadvmax,
                     (vec_x, vec_y, vec_n), near_wall)


def _epilogue(nii, njj, nkk, vcf, dxk, dyk, dzk, hok, cutk, advdom, uk, vk, wk, kkart, ahk, bui, kadvmax, vec,
              near_wall):
    """KADVMAX, the new AHMIN and the near-wall damping, common to both
    terrain variants (MicroscaleTerrain.cs:600-612, :664-832)."""
    DX, DY = F(dxk), F(dyk)
    kadvmax = min(nkk, kadvmax + vcf)
    ahmin_new = F(100000)
    for i in range(1, nii + 1):
        for j in range(1, njj + 1):
            ahmin_new = net_min(ahmin_new, ahk[i, j])
    # near-wall damping (MicroscaleTerrain.cs:664-832): u, v and w, levels above AHMIN
    if near_wall:
        igeb = max(int(F(20) / DX), 1)
        for i in range(1 + igeb, nii - igeb + 1):
            for j in range(1 + igeb, njj - igeb + 1):
                if near_wall == 2 or advdom[i, j] == 0:
                    for k in range(1, nkk):
                        if k > kkart[i, j]:
                            vertk = F(F(hok[k - 1] + F(dzk[k] * F(0.5))) + ahmin_new)
                            ab = F(1)
                            dirs = [[(i - g, j, F(abs(F(F(g) * DX)))) for g in range(1, igeb + 1)],
                                    [(i + g, j, F(abs(F(F(-g) * DX)))) for g in range(1, igeb + 1)],
                                    [(i, j + g, F(abs(F(F(-g) * DY)))) for g in range(1, igeb + 1)],
                                    [(i, j - g, F(abs(F(F(g) * DY)))) for g in range(1, igeb + 1)],
                                    [(i + g, j + g, None) for g in range(1, igeb + 1)],
                                    [(i - g, j - g, None) for g in range(1, igeb + 1)],
                                    [(i + g, j - g, None) for g in range(1, igeb + 1)],
                                    [(i - g, j + g, None) for g in range(1, igeb + 1)]]
                            for cells in dirs:
                                entf = F(21)
                                for ig, jg, dist in cells:
                                    if cutk[ig, jg] > 1 and vertk <= ahk[ig, jg]:
                                        if dist is None:
                                            a = F(F(i - ig) * DX)
                                            b = F(F(j - jg) * DY)
                                            dist = sqrtf(F(F(a * a) + F(b * b)))
                                        entf = dist
                                        break
                                if entf <= 20:
                                    ab = F(ab * F(F(0.19) * logf(F(F(entf + F(0.5)) * F(10)))))
                            uk[i, j, k] = F(uk[i, j, k] * ab)
                            vk[i, j, k] = F(vk[i, j, k] * ab)
                            wk[i, j, k] = F(wk[i, j, k] * ab)
    return uk, vk, wk, kkart, ahk, bui, ahmin_new, kadvmax, vec
