"""Second, independent transcription of PrognosticFlowfield.Calculate in numpy.

TEST INFRASTRUCTURE: it pins the C oracle (oracle/pf_oracle.c) by agreeing
with it bit for bit on tiny domains.  It is written separately from the C
oracle, directly from the reference's C# (paths relative to the reference's
src/), in a different shape:
  * the Vector<float> solvers (U/V/W-prognostic-microscale_1.cs) are executed
    block by block on width-W numpy float32 vectors, as the JIT does;
  * the scalar solvers (_0/_2, TKE) loop over k with numpy float32 scalars;
  * FP64 sub-expressions are evaluated with Python floats;
  * the fused multiply-adds of the 512-bit solvers
    (U/V/W-prognostic-microscale_1_Vec512.cs, simd_width 16) are evaluated by
    fma32() below: exact product and error-free sum in float64, then one
    rounding to float32 (no C library fmaf, which the C oracle uses).
NumPy 2 treats Python scalars as weak, so every float32 literal is written as
F(...) and every C# double promotion as float(...).
Only for small cases: pure Python loops over columns.
"""
from __future__ import annotations

import math

import numpy as np

F = np.float32


def vmax(a, b):                      # Vector.Max / FloatMax: maxps semantics
    return np.where(a > b, a, b).astype(F) if isinstance(a, np.ndarray) or isinstance(b, np.ndarray) else (a if a > b else b)


def vmin(a, b):
    return np.where(a < b, a, b).astype(F) if isinstance(a, np.ndarray) or isinstance(b, np.ndarray) else (a if a < b else b)


def net_max(a, b):                   # Math.Max(float/double) of .NET Core
    if a != b:
        if not math.isnan(a):
            return a if b < a else b
        return a
    return a if math.copysign(1.0, b) < 0 else b


def net_min(a, b):
    if a != b:
        if not math.isnan(a):
            return a if a < b else b
        return a
    return a if math.copysign(1.0, a) < 0 else b


def fma32(a, b, c):
    """a*b+c of float32 operands with a single rounding to float32.

    The product of two float32 values is exact in float64; the sum s = p + c is
    rounded once to float64 and its exact error e is recovered by TwoSum, so
    p + c == s + e exactly.  Rounding s to float32 can differ from rounding
    s + e only when s is a tie between two float32 neighbours (rounding is
    monotonic, so s never crosses one); then e decides the side."""
    a64, b64, c64 = (np.asarray(v, dtype=np.float64) for v in (a, b, c))
    p = a64 * b64
    s = p + c64
    bb = s - p
    e = (p - (s - bb)) + (c64 - bb)
    r = s.astype(F)
    r64 = r.astype(np.float64)
    other = np.nextafter(r, np.where(s > r64, F(np.inf), F(-np.inf)).astype(F))
    tie = (s != r64) & (s == (r64 + other.astype(np.float64)) * 0.5)
    fix = tie & (e != 0) & (np.sign(e) == np.sign(s - r64))
    out = np.where(fix, other, r).astype(F)
    return out if out.ndim else F(out)


def conv_to_int(x: float) -> int:    # ProgramFunctions.cs:792-804 (double overload)
    return int(x - 0.5) if x < 0 else int(x + 0.5)


class Solver:
    def __init__(self, case, iterations):
        c = case
        self.c = c
        self.NII, self.NJJ, self.NKK = c.nii, c.njj, c.nkk
        self.KADVMAX, self.VCF = c.kadvmax, c.vertcellsff
        self.DXK, self.DYK = F(c.dxk), F(c.dyk)
        self.DZK = c.dzk.astype(F)
        self.HOK = c.hokart.astype(F)
        self.ADV = c.advdom.astype(np.int64)
        self.KKART = c.kkart.astype(np.int64)
        self.CUTK = c.cutk.astype(F)
        self.COV = c.cov.astype(F)
        self.VEG = None if c.veg is None else c.veg.astype(F)
        self.VEGP = None if c.veg_present is None else c.veg_present.astype(bool)
        self.model = c.turbulence_model
        self.W = int(getattr(c, "simd_width", 8))
        self.relax, self.relaxp = F(c.relax), F(c.relaxp)
        self.bz0 = F(c.building_z0)
        self.Cor = F(c.coriolis)
        self.iters = iterations
        self.UK = c.uk.astype(F).copy()
        self.VK = c.vk.astype(F).copy()
        self.WK = c.wk.astype(F).copy()
        shp = self.UK.shape
        for n in ("UKS", "VKS", "WKS", "DPM", "DPMNEW", "DIV", "TURB", "TDISS"):
            setattr(self, n, np.zeros(shp, dtype=F))

    # ------------------------------------------------------------------ prologue
    def prologue(self):
        """PrognosticFlowfield.cs:145-579."""
        c, NII, NJJ, NKK = self.c, self.NII, self.NJJ, self.NKK
        UK, VK, WK = self.UK, self.VK, self.WK
        rx, ry = c.refp
        tw = F(math.sqrt(float(F(UK[rx, ry, NKK - 1] * UK[rx, ry, NKK - 1]) + F(VK[rx, ry, NKK - 1] * VK[rx, ry, NKK - 1]))))
        tw = F(net_max(tw, F(0.01)))
        self.topwind = tw
        sp = np.sqrt(F(UK[rx, ry, self.KADVMAX] * UK[rx, ry, self.KADVMAX] + VK[rx, ry, self.KADVMAX] * VK[rx, ry, self.KADVMAX]))
        dt = F(F(F(0.5) * F(net_min(self.DXK, self.DZK[1]))) / F(net_max(F(0.01), F(sp))))
        self.DTIME = F(net_min(dt, F(5)))
        self.AREA = F(self.DXK * self.DYK)
        self.AP0 = np.zeros(self.KADVMAX + 2, dtype=F)
        for k in range(1, self.KADVMAX + 1):
            self.AP0[k] = F(F(self.AREA * self.DZK[k]) / self.DTIME)
        if self.model != 1:
            self.TURB[1:NII + 1, 1:NJJ + 1, 1:NKK + 2] = F(0.001)
            self.TDISS[1:NII + 1, 1:NJJ + 1, 1:NKK + 2] = F(0.0000001)
        self.Frund = float(F(F(10000 if self.model == 2 else 8000) / tw))
        adv = self.ADV
        g = 0.0
        for i in range(1, NII + 1):
            for j in range(1, NJJ + 1):
                ip, jp = min(i + 1, NII + 1), min(j + 1, NJJ + 1)
                if not (adv[i, j] or adv[i - 1, j] or adv[ip, j] or adv[i, j - 1] or adv[i - 1, j - 1] or adv[ip, j - 1]
                        or adv[i, jp] or adv[i - 1, jp] or adv[ip, jp]):
                    continue
                ks = slice(1, NKK)
                self.UKS[i, j, ks] = F(0.5) * (UK[i, j, ks] + UK[i + 1, j, ks])
                self.VKS[i, j, ks] = F(0.5) * (VK[i, j, ks] + VK[i, j + 1, ks])
                self.WKS[i, j, ks] = F(0.5) * (WK[i, j, 1:NKK] + WK[i, j, 2:NKK + 1])
                s2 = self.UKS[i, j, ks] * self.UKS[i, j, ks] + self.VKS[i, j, ks] * self.VKS[i, j, ks]
                for v in s2:
                    g = net_max(g, math.sqrt(float(v)))
        self.GESCHW_MAX = g
        self.VI = np.minimum(NKK - 1, self.KKART + self.VCF)
        self.HT = np.zeros_like(self.CUTK)
        self.HO = np.zeros_like(self.CUTK)
        for i in range(1, NII + 1):
            for j in range(1, NJJ + 1):
                k = int(self.KKART[i, j]) + 1
                ks = 1 if self.CUTK[i, j] != 0 else k
                if not (ks <= k <= self.VI[i, j]):
                    continue
                rough = F(c.z0_wall[i, j])
                D = self.DZK
                if self.CUTK[i, j] < 1:
                    if float(rough) >= float(D[k]) * 0.1:
                        self.HT[i, j] = F(0.4) / F(math.log(float(F(F(F(D[k] + F(D[k + 1] * F(0.5))) / D[k]) * F(0.1)))))
                    else:
                        self.HT[i, j] = F(0.4) / F(math.log(float(F(F(D[k] * F(0.5)) / rough))))
                else:
                    self.HO[i, j] = F(0.4) / F(math.log(float(F(F(D[k] * F(0.5)) / self.bz0))))
        self.UG = self.UKS[rx, ry, NKK - 1]
        self.VG = self.VKS[rx, ry, NKK - 1]
        for i in range(1, NII + 1):
            for j in range(1, NJJ + 1):
                ob, ust, rough = F(c.ob_init[i, j]), F(c.ust_init[i, j]), F(c.z0_init[i, j])
                for k in range(1, NKK):
                    if i == 1:
                        UK[NII + 1, j, k] = UK[NII, j, k]
                    if j == 1:
                        VK[i, NJJ + 1, k] = VK[i, NJJ, k]
                    if self.model == 1:
                        continue
                    wh = net_max(math.sqrt(float(F(UK[i, j, k] * UK[i, j, k] + VK[i, j, k] * VK[i, j, k]))), 0.01)
                    u0 = F(wh * (0.2 * math.pow(wh, -0.9) + 0.32 * float(rough) + 0.18))
                    u0 = F(F(net_max(float(u0), 0.3)) * F(c.std_dev_v))
                    v0 = F(c.v0_1)
                    if ob < 0:
                        pw = np.power(F(F(c.bdlayheight) / F(-ob)), F(0.67))
                        varw = F(F(ust * ust) * F(F(F(1.15) + F(F(0.1) * F(pw))) ** 2)) * F(F(c.std_dev_w) * F(c.std_dev_w))
                    else:
                        t = F(F(ust * F(c.std_dev_w)) * F(1.25))
                        varw = F(t * t)
                    self.TURB[i, j, k] = F(0.5 * float(F(F(u0 * u0) + F(v0 * v0)) + varw))
        self.VISHMIN = F(0.01) if self.model == 2 else F(0.0)
        self.VISVMIN = self.VISHMIN
        self.UFX = F(0.4 / math.log(float(self.DXK) * 0.5 / float(self.bz0)))
        for i in range(1, NII + 1):
            for j in range(1, NJJ + 1):
                if adv[i, j] != 1:
                    continue
                vi, kk = int(self.VI[i, j]), int(self.KKART[i, j])
                for k in range(1, vi + 1):
                    if kk >= k:
                        UK[i, j, k] = 0
                        if i == NII:
                            UK[i + 1, j, k] = 0
                        VK[i, j, k] = 0
                        if j == NJJ:
                            VK[i, j + 1, k] = 0
                        WK[i, j, k] = 0
                        if k == vi:
                            WK[i, j, k + 1] = 0
                        if self.model != 1:
                            self.TURB[i, j, k] = 0
                            self.TDISS[i, j, k] = 0
                        self.UKS[i, j, k] = 0
                        self.VKS[i, j, k] = 0
                        self.WKS[i, j, k] = 0

    # ------------------------------------------------------------ SIMPLE pieces
    def boundary(self):
        """PrognosticFlowfield.cs:595-714."""
        NII, NJJ, adv, UK, VK, WK = self.NII, self.NJJ, self.ADV, self.UK, self.VK, self.WK
        for j in range(2, NJJ + 1):
            for i in range(3, NII):
                if adv[i, j] != 1:
                    continue
                n = int(self.VI[i, j]) + 1
                if adv[i - 1, j] < 1 or i == 3:
                    u = UK[i - 2, j, 1:n].copy()
                    UK[i - 1, j, 1:n] = np.where(u >= 0, u, UK[i - 1, j, 1:n])
                    VK[i - 2, j, 1:n] = VK[i - 1, j, 1:n]
                    WK[i - 2, j, 1:n] = WK[i - 1, j, 1:n]
                if adv[i + 1, j] < 1 or i == NII - 1:
                    u = UK[i + 2, j, 1:n].copy()
                    UK[i + 1, j, 1:n] = np.where(u <= 0, u, UK[i + 1, j, 1:n])
                    VK[i + 1, j, 1:n] = VK[i, j, 1:n]
                    WK[i + 1, j, 1:n] = WK[i, j, 1:n]
        for i in range(2, NII + 1):
            for j in range(3, NJJ):
                if adv[i, j] != 1:
                    continue
                n = int(self.VI[i, j]) + 1
                if adv[i, j - 1] < 1 or j == 3:
                    v = VK[i, j - 2, 1:n].copy()
                    VK[i, j - 1, 1:n] = np.where(v >= 0, v, VK[i, j - 1, 1:n])
                    UK[i, j - 2, 1:n] = UK[i, j - 1, 1:n]
                    WK[i, j - 2, 1:n] = WK[i, j - 1, 1:n]
                if adv[i, j + 1] < 1 or j == NJJ - 1:
                    v = VK[i, j + 2, 1:n].copy()
                    VK[i, j + 1, 1:n] = np.where(v <= 0, v, VK[i, j + 1, 1:n])
                    UK[i, j + 1, 1:n] = UK[i, j, 1:n]
                    WK[i, j + 1, 1:n] = WK[i, j, 1:n]
        for i in range(2, NII):
            for j in range(2, NJJ):
                if adv[i, j] == 1:
                    k = int(self.VI[i, j])
                    WK[i, j, k + 1] = WK[i, j, k]
                    UK[i, j, k + 1] = UK[i, j, k]
                    VK[i, j, k + 1] = VK[i, j, k]

    def divergence(self):
        """PrognosticFlowfield.cs:723-758."""
        for i in range(2, self.NII):
            for j in range(2, self.NJJ):
                if self.ADV[i, j] != 1:
                    continue
                kk, vi = int(self.KKART[i, j]), int(self.VI[i, j])
                ks = np.arange(kk + 1, vi + 1)
                dz = self.DZK[ks]
                t1 = (self.UK[i, j, ks] - self.UK[i + 1, j, ks]) * self.DYK * dz
                t2 = (self.VK[i, j, ks] - self.VK[i, j + 1, ks]) * self.DXK * dz
                t3 = (self.WK[i, j, ks] - self.WK[i, j, ks + 1]) * self.AREA
                t3b = self.WK[i, j, ks + 1] * self.AREA
                self.DIV[i, j, ks] = np.where(ks > kk + 1, (t1 + t2) + t3, (t1 + t2) - t3b)
                self.DPM[i, j, ks] = 0

    def pressure(self, IS, JS):
        """PrognosticFlowfield.cs:804-901 (deterministic: rows in order)."""
        NII, NJJ, D = self.NII, self.NJJ, self.DZK
        dmax = F(0)
        for j1 in range(2, NJJ):
            PIMP = np.zeros(self.NKK + 2, dtype=F)
            QIMP = np.zeros(self.NKK + 2, dtype=F)
            intern = F(0)
            j = NJJ - j1 + 1 if JS == -1 else j1
            for i1 in range(2, NII):
                i = NII - i1 + 1 if IS == -1 else i1
                if self.ADV[i, j] != 1:
                    continue
                kk, vi = int(self.KKART[i, j]), int(self.VI[i, j])
                DPM = self.DPM
                for k in range(vi, kk, -1):
                    DIM = F(F(self.DIV[i, j, k] / self.DTIME) + F(F(F((DPM[i - 1, j, k] + DPM[i + 1, j, k]) / self.DXK) * self.DYK) * D[k])) \
                        + F(F(F((DPM[i, j - 1, k] - DPM[i, j + 1, k]) / self.DYK) * self.DXK) * D[k])
                    APP = F(2) * F(F(F(self.DYK * D[k]) / self.DXK) + F(F(self.DXK * D[k]) / self.DYK) + F(self.AREA / D[k]))
                    ABP = self.AREA / F(F(0.5) * (D[k - 1] + D[k])) if k > 1 else self.AREA / D[k]
                    ATP = self.AREA / F(F(0.5) * (D[k + 1] + D[k]))
                    TERMP = F(1) / F(APP - F(ATP * PIMP[k + 1]))
                    PIMP[k] = ABP * TERMP
                    QIMP[k] = F(F(ATP * QIMP[k + 1]) + DIM) * TERMP
                for k in range(kk + 1, vi + 1):
                    temp = F(F(PIMP[k] * DPM[i, j, k - 1]) + QIMP[k])
                    q = conv_to_int(float(temp) * self.Frund) / self.Frund
                    intern = F(net_max(float(intern), abs(float(DPM[i, j, k]) - q)))
                    DPM[i, j, k] = temp
                    self.DPMNEW[i, j, k] = F(self.DPMNEW[i, j, k] + F(self.relaxp * temp))
            if intern > dmax:
                dmax = intern
        return dmax

    def correction(self):
        """PrognosticFlowfield.cs:904-949."""
        fr, dt = self.Frund, self.DTIME
        for i in range(2, self.NII):
            for j in range(2, self.NJJ):
                if not (self.ADV[i, j] == 1 or self.ADV[i - 1, j] == 1 or self.ADV[i, j - 1] == 1):
                    continue
                kk, vi = int(self.KKART[i, j]), int(self.VI[i, j])
                kim, kjm = int(self.KKART[i - 1, j]), int(self.KKART[i, j - 1])
                for k in range(1, vi + 1):
                    d0 = self.DPM[i, j, k]
                    if i > 2 and kk < k and kim < k:
                        e = F(self.UK[i, j, k] + F(F(F(self.DPM[i - 1, j, k] - d0) / self.DXK) * dt))
                        self.UK[i, j, k] = F(conv_to_int(float(e) * fr) / fr)
                    if j > 2 and kk < k and kjm < k:
                        e = F(self.VK[i, j, k] + F(F(F(self.DPM[i, j - 1, k] - d0) / self.DYK) * dt))
                        self.VK[i, j, k] = F(conv_to_int(float(e) * fr) / fr)
                    if kk + 1 < k:
                        e = F(self.WK[i, j, k] + F(F(F(F(self.DPM[i, j, k - 1] - d0) / F(self.DZK[k] + self.DZK[k - 1])) * F(2)) * dt))
                        self.WK[i, j, k] = F(conv_to_int(float(e) * fr) / fr)
                    else:
                        self.WK[i, j, k] = 0

    def averages(self):
        """PrognosticFlowfield.cs:958-1039."""
        NII, NJJ = self.NII, self.NJJ
        for i in range(2, NII):
            for j in range(2, NJJ):
                if not (self.ADV[i, j] == 1 or self.ADV[i - 1, j] == 1 or self.ADV[i, j - 1] == 1):
                    continue
                kk, vi = int(self.KKART[i, j]), int(self.VI[i, j])
                ks0 = 1 if self.CUTK[i, j] != 0 else kk + 1
                ks = np.arange(ks0, vi + 1)
                m1 = ks > kk
                m2 = ks > kk + 1
                self.UKS[i, j, ks[m1]] = F(0.5) * (self.UK[i, j, ks[m1]] + self.UK[i + 1, j, ks[m1]])
                self.VKS[i, j, ks[m1]] = F(0.5) * (self.VK[i, j, ks[m1]] + self.VK[i, j + 1, ks[m1]])
                self.WKS[i, j, ks[m2]] = F(0.5) * (self.WK[i, j, ks[m2]] + self.WK[i, j, ks[m2] + 1])
                if j == 2:
                    self.UKS[i, 1, ks] = self.UKS[i, j, ks]
                if j == NJJ - 1 and self.ADV[i, NJJ] == 1:
                    self.UKS[i, NJJ, ks] = self.UKS[i, j, ks]
        for j in range(2, NJJ):
            if self.ADV[1, j] == 1:
                self.VKS[1, j, 0:self.NKK + 1] = self.VKS[2, j, 0:self.NKK + 1]
            if self.ADV[NII - 1, j] == 1:
                self.VKS[NII, j, 0:self.NKK + 1] = self.VKS[NII - 1, j, 0:self.NKK + 1]

    # ------------------------------------------------- mixing-length solvers
    def _ml_block(self, comp, i, j, k, UG):
        """One Vector<float> block (lanes k..k+W-1) of U/V/W_PrognosticMicroscaleV1."""
        W, kk = self.W, int(self.KKART[i, j])
        KKP1 = kk + 1
        ks = np.arange(k, k + W)
        D = self.DZK[ks]
        DXKDZK, DYKDZK = self.DXK * D, self.DYK * D
        AREA = self.AREA
        UK, VK, WK, UKS, VKS, WKS = self.UK, self.VK, self.WK, self.UKS, self.VKS, self.WKS
        if W == 16:                          # U-prognostic-microscale_1_Vec512.cs:218
            zs = self.HOK[ks] - fma32(D, F(0.5), np.full(W, self.HOK[kk]))
        else:
            zs = (self.HOK[ks] - self.HOK[kk]) - D * F(0.5)
        VIS = np.zeros(W, dtype=F)
        mask_on = KKP1 >= k and KKP1 < k + W
        Mask = (ks > KKP1).astype(F)
        if k + W > KKP1:
            mixL = vmin(F(0.071) * zs, np.full(W, F(100)))
            if self.COV[i, j] > 0:
                mixL = mixL * (F(1) - F(0.99) * self.COV[i, j])
            it = (UK[i, j, ks + 1] - UK[i, j, ks - 1]) / (F(2) * D)
            dudz = it * it
            it = (VK[i, j, ks + 1] - VK[i, j, ks - 1]) / (F(2) * D)
            dvdz = it * it
            v = vmin(mixL * mixL * np.sqrt(F(0.5) * (dudz + dvdz)), np.full(W, F(15)))
            VIS = v * Mask if mask_on else v
        mv = vmax(VIS, np.full(W, self.VISHMIN))
        DE = mv * DYKDZK / self.DXK
        DW = DE.copy()
        DS = mv * DXKDZK / self.DYK
        DN = DS.copy()
        DT = mv * AREA / D
        DB = np.zeros(W, dtype=F)
        if mask_on:
            DB = DT * Mask
        elif k > KKP1:
            DB = DT.copy()
        z = np.zeros(W, dtype=F)
        if self.ADV[i, j - 1] < 1 or j == 2:
            DS = z.copy()
        if self.ADV[i, j + 1] < 1 or j == self.NJJ - 1:
            DN = z.copy()
        if self.ADV[i - 1, j] < 1 or i == 2:
            DW = z.copy()
        if self.ADV[i + 1, j] < 1 or i == self.NII - 1:
            DE = z.copy()
        q = F(0.25)
        if comp == "u":
            FE = UKS[i, j, ks] * DYKDZK
            FW = UKS[i - 1, j, ks] * DYKDZK
            FS = q * (VKS[i - 1, j, ks] + VKS[i, j, ks] + VKS[i - 1, j - 1, ks] + VKS[i, j - 1, ks]) * DXKDZK
            FN = q * (VKS[i - 1, j, ks] + VKS[i, j, ks] + VKS[i - 1, j + 1, ks] + VKS[i, j + 1, ks]) * DXKDZK
            FT = q * (WKS[i - 1, j, ks] + WKS[i, j, ks] + WKS[i - 1, j, ks + 1] + WKS[i, j, ks + 1]) * AREA
            fb = q * (WKS[i - 1, j, ks] + WKS[i, j, ks] + WKS[i - 1, j, ks - 1] + WKS[i, j, ks - 1]) * AREA
        elif comp == "v":
            FE = q * (UKS[i, j, ks] + UKS[i + 1, j, ks] + UKS[i, j - 1, ks] + UKS[i + 1, j - 1, ks]) * DYKDZK
            FW = q * (UKS[i, j, ks] + UKS[i - 1, j, ks] + UKS[i, j - 1, ks] + UKS[i - 1, j - 1, ks]) * DYKDZK
            FS = VKS[i, j - 1, ks] * DXKDZK
            FN = VKS[i, j, ks] * DXKDZK
            FT = q * (WKS[i, j - 1, ks] + WKS[i, j, ks] + WKS[i, j - 1, ks + 1] + WKS[i, j, ks + 1]) * AREA
            fb = q * (WKS[i, j - 1, ks] + WKS[i, j, ks] + WKS[i, j - 1, ks - 1] + WKS[i, j, ks - 1]) * AREA
        else:
            FE = q * (UKS[i, j, ks] + UKS[i + 1, j, ks] + UKS[i, j, ks - 1] + UKS[i + 1, j, ks - 1]) * DYKDZK
            FW = q * (UKS[i, j, ks] + UKS[i - 1, j, ks] + UKS[i, j, ks - 1] + UKS[i - 1, j, ks - 1]) * DYKDZK
            FS = q * (VKS[i, j, ks] + VKS[i, j - 1, ks] + VKS[i, j, ks - 1] + VKS[i, j - 1, ks - 1]) * DXKDZK
            FN = q * (VKS[i, j, ks] + VKS[i, j + 1, ks] + VKS[i, j, ks - 1] + VKS[i, j + 1, ks - 1]) * DXKDZK
            FT = WKS[i, j, ks] * AREA
            fb = WKS[i, j, ks - 1] * AREA
        FB = np.zeros(W, dtype=F)
        if mask_on:
            FB = fb * Mask
        elif k > KKP1:
            FB = fb
        lo = np.full(W, F(0.0001))
        DE, DB, DW, DS, DN, DT = vmax(DE, lo), vmax(DB, lo), vmax(DW, lo), vmax(DS, lo), vmax(DN, lo), vmax(DT, lo)
        PE, PB, PW = np.abs(FE / DE), np.abs(FB / DB), np.abs(FW / DW)
        PS, PN, PT = np.abs(FS / DS), np.abs(FN / DN), np.abs(FT / DT)

        def pl(Dv, Pv, Fv):
            t = F(1) - F(0.1) * Pv
            if W == 16:                      # U-prognostic-microscale_1_Vec512.cs:355-377
                return fma32(Dv, vmax(z, t * ((t * t) * (t * t))), vmax(Fv, z))
            return Dv * vmax(z, t * t * t * t * t) + vmax(Fv, z)
        BIM = pl(DT, PT, -FT)
        CIM = pl(DB, PB, FB)
        AE1 = pl(DE, PE, -FE)
        AW1 = pl(DW, PW, FW)
        AS1 = pl(DS, PS, FS)
        AN1 = pl(DN, PN, -FN)
        AP0 = self.AP0[ks]
        AIM = BIM + CIM + AW1 + AS1 + AE1 + AN1 + AP0
        vegp = self.VEGP is not None and self.VEGP[i, j]
        if comp == "u":
            X = UK
            ddp, ddm = self.DPMNEW[i - 1, j, ks] - self.DPMNEW[i, j, ks], DYKDZK
        elif comp == "v":
            X = VK
            ddp, ddm = self.DPMNEW[i, j - 1, ks] - self.DPMNEW[i, j, ks], DXKDZK
        else:
            X = WK
            ddp, ddm = self.DPMNEW[i, j, ks - 1] - self.DPMNEW[i, j, ks], AREA
        DDP = ddp * ddm
        x = X[i, j, ks]
        # AP0_V * Vect_05 * (a + b) is (AP0*0.5)*(a+b)
        if comp == "u":
            ssum = UKS[i - 1, j, ks] + UKS[i, j, ks]
        elif comp == "v":
            ssum = VKS[i, j - 1, ks] + VKS[i, j, ks]
        else:
            ssum = WKS[i, j, ks - 1] + WKS[i, j, ks]
        # the 512-bit solvers fuse the neighbour products pairwise, except V
        # without vegetation (V-prognostic-microscale_1_Vec512.cs:414-419)
        if W == 16 and (comp != "v" or vegp):
            nb = fma32(AW1, X[i - 1, j, ks], AS1 * X[i, j - 1, ks]) + fma32(AE1, X[i + 1, j, ks], AN1 * X[i, j + 1, ks])
        else:
            nb = AW1 * X[i - 1, j, ks] + AS1 * X[i, j - 1, ks] + AE1 * X[i + 1, j, ks] + AN1 * X[i, j + 1, ks]
        DIM = nb + AP0 * F(0.5) * ssum
        if comp in ("u", "v"):
            G = UG
            if vegp:
                in2 = F(0.5) * (VKS[i - 1, j, ks] + VKS[i, j, ks])
                in1 = F(0.5) * (UKS[i - 1, j, ks] + UKS[i, j, ks])
                wh = vmax(np.sqrt(in1 * in1 + in2 * in2), np.full(W, F(0.01)))
                DIM = DIM + DDP + (self.Cor * (G - x) - self.VEG[i, j, ks - 1] * x * wh) * AREA * D
            elif W == 16 and comp == "u":    # U-prognostic-microscale_1_Vec512.cs:419-421
                DIM = DIM + fma32(ddp, ddm, self.Cor * (G - x) * AREA * D)
            else:
                DIM = DIM + DDP + self.Cor * (G - x) * AREA * D
        else:
            DIM = DIM + DDP
            if vegp:
                in2 = F(0.5) * (VKS[i, j - 1, ks] + VKS[i, j, ks])
                in1 = F(0.5) * (UKS[i - 1, j, ks] + UKS[i, j, ks])
                wh = vmax(np.sqrt(in1 * in1 + in2 * in2), np.full(W, F(0.01)))
                DIM = DIM - self.VEG[i, j, ks - 1] * x * wh * AREA * D
        if comp in ("u", "v") and k <= KKP1 < k + W:
            ii = KKP1 - k
            m2 = np.zeros(W, dtype=F)
            m2[ii] = 1
            in1 = F(0.5) * (UKS[i - 1, j, ks] + UKS[i, j, ks])
            in2 = F(0.5) * (VKS[i - 1, j, ks] + VKS[i, j, ks])
            wh = vmax(np.sqrt(in1 * in1 + in2 * in2), np.full(W, F(0.1)))
            if self.CUTK[i, j] < 1:
                us = F(self.HT[i, j] * wh[ii])
                if F(self.c.z0_solver[i, j]) >= F(D[ii] * F(0.1)):
                    ind = k + ii + 1
                    a_ = F(F(0.5) * F(UKS[i - 1, j, ind] + UKS[i, j, ind]))
                    b_ = F(F(0.5) * F(VKS[i - 1, j, ind] + VKS[i, j, ind]))
                    us = F(self.HT[i, j] * F(np.sqrt(F(F(a_ * a_) + F(b_ * b_)))))
            else:
                us = F(self.HO[i, j] * wh[ii])
            DIM = DIM - x / wh * F(us * us) * AREA * m2
        if k > KKP1:
            h = F(0.5)
            if comp == "u":
                t1 = (UK[i + 1, j, ks] - x) * DYKDZK
                t2 = (x - UK[i - 1, j, ks]) * DYKDZK
                t3 = (h * (VKS[i + 1, j + 1, ks] + VKS[i, j + 1, ks]) - h * (VKS[i, j + 1, ks] + VKS[i - 1, j + 1, ks])) * DXKDZK
                t4 = (h * (VKS[i + 1, j - 1, ks] + VKS[i, j - 1, ks]) - h * (VKS[i, j - 1, ks] + VKS[i - 1, j - 1, ks])) * DXKDZK
                t5 = (h * (WK[i + 1, j, ks] + WKS[i, j, ks]) - h * (WKS[i, j, ks] + WKS[i - 1, j, ks])) * AREA
                t6 = (h * (WK[i + 1, j, ks - 1] + WKS[i, j, ks - 1]) - h * (WKS[i, j, ks - 1] + WKS[i - 1, j, ks - 1])) * AREA
                div = self.DXK
            elif comp == "v":
                t3 = (VK[i, j + 1, ks] - x) * DXKDZK
                t4 = (x - VK[i, j - 1, ks]) * DXKDZK
                t1 = (h * (UKS[i + 1, j + 1, ks] + UKS[i + 1, j, ks]) - h * (UKS[i + 1, j, ks] + UKS[i + 1, j - 1, ks])) * DXKDZK
                t2 = (h * (UKS[i - 1, j + 1, ks] + UKS[i - 1, j, ks]) - h * (UKS[i - 1, j, ks] + UKS[i - 1, j - 1, ks])) * DXKDZK
                t5 = (h * (WK[i, j + 1, ks] + WKS[i, j, ks]) - h * (WKS[i, j, ks] + WKS[i, j - 1, ks])) * AREA
                t6 = (h * (WK[i, j + 1, ks - 1] + WKS[i, j, ks - 1]) - h * (WKS[i, j, ks - 1] + WKS[i, j - 1, ks])) * AREA
                div = self.DXK
            else:
                t1 = (h * (UKS[i, j, ks + 1] + UKS[i, j, ks]) - h * (UKS[i, j, ks] + UKS[i, j, ks - 1])) * DYKDZK
                t2 = (h * (UKS[i - 1, j, ks + 1] + UKS[i - 1, j, ks]) - h * (UKS[i - 1, j, ks] + UKS[i - 1, j, ks - 1])) * DYKDZK
                t3 = (h * (VKS[i, j, ks + 1] + VKS[i, j, ks]) - h * (VKS[i, j, ks] + VKS[i, j, ks - 1])) * DXKDZK
                t4 = (h * (VKS[i, j - 1, ks + 1] + VKS[i, j - 1, ks]) - h * (VKS[i, j - 1, ks] + VKS[i, j - 1, ks - 1])) * DXKDZK
                t5 = (WK[i, j, ks + 1] - x) * AREA
                t6 = (x - WK[i, j, ks - 1]) * AREA
                div = D
            add = VIS / div * (t1 - t2 + t3 - t4 + t5 - t6)
            DIM = DIM + (add * Mask if mask_on else add)
        return AIM.astype(F), BIM.astype(F), CIM.astype(F), DIM.astype(F)

    def _ml_column(self, comp, i, j, G):
        W, kk = self.W, int(self.KKART[i, j])
        KKP1 = kk + 1
        KSTART = 1 if self.CUTK[i, j] != 0 else kk + 1
        KEND = int(self.VI[i, j]) + 1
        P = np.zeros(self.KADVMAX + 2, dtype=F)
        Q = np.zeros(self.KADVMAX + 2, dtype=F)
        k = KSTART
        while k < KEND:
            k_real, end = k, False
            if k > KEND - W:
                k, end = KEND - W, True
            A, B, C, Dd = self._ml_block(comp, i, j, k, G)
            s0 = 0
            if end:
                s0 = k_real - k
                k_real -= s0
            for l in range(s0, W):
                idx = k_real + l
                if idx < KEND:
                    if idx > KKP1:
                        T = F(F(1) / F(A[l] - F(C[l] * P[idx - 1])))
                        P[idx] = F(B[l] * T)
                        Q[idx] = F(F(Dd[l] + F(C[l] * Q[idx - 1])) * T)
                    else:
                        T = F(F(1) / A[l])
                        P[idx] = F(B[l] * T)
                        Q[idx] = F(Dd[l] * T)
            if end:
                k = KEND + 1
            k += W
        return P, Q, KSTART

    def ml_sweep(self, comp, IS, JS):
        NII, NJJ = self.NII, self.NJJ
        i1r = range(3, NII) if comp == "u" else range(2, NII)
        j1r = range(3, NJJ) if comp == "v" else range(2, NJJ)
        X = {"u": self.UK, "v": self.VK, "w": self.WK}[comp]
        G = {"u": self.UG, "v": self.VG, "w": F(0)}[comp]
        for i1 in i1r:
            i = NII - i1 + 1 if IS == -1 else i1
            for j1 in j1r:
                j = NJJ - j1 + 1 if JS == -1 else j1
                if self.ADV[i, j] != 1:
                    continue
                P, Q, KSTART = self._ml_column(comp, i, j, G)
                kk = int(self.KKART[i, j])
                kkm = max(kk, int(self.KKART[i - 1, j])) if comp == "u" else max(kk, int(self.KKART[i, j - 1])) if comp == "v" else -1
                for k in range(int(self.VI[i, j]), KSTART - 1, -1):
                    if kkm < k:
                        X[i, j, k] = F(X[i, j, k] + F(self.relax * F(F(F(P[k] * X[i, j, k + 1]) + Q[k]) - X[i, j, k])))

    # -------------------------------------------------- k-eps / no-diffusion
    def ke_sweep(self, comp, IS, JS, diffusion):
        NII, NJJ = self.NII, self.NJJ
        i1r = range(3, NII) if comp == "u" else range(2, NII)
        j1r = range(3, NJJ) if comp == "v" else range(2, NJJ)
        UK, VK, WK, UKS, VKS, WKS = self.UK, self.VK, self.WK, self.UKS, self.VKS, self.WKS
        X = {"u": UK, "v": VK, "w": WK}[comp]
        h, q = F(0.5), F(0.25)
        for i1 in i1r:
            for j1 in j1r:
                j = NJJ - j1 + 1 if JS == -1 else j1
                i = NII - i1 + 1 if IS == -1 else i1
                if self.ADV[i, j] != 1:
                    continue
                kk, vi = int(self.KKART[i, j]), int(self.VI[i, j])
                ks0 = 1 if self.CUTK[i, j] != 0 else kk + 1
                P = np.zeros(self.KADVMAX + 2, dtype=F)
                Q = np.zeros(self.KADVMAX + 2, dtype=F)
                for k in range(ks0, vi + 1):
                    dz = self.DZK[k]
                    dxz, dyz = F(self.DXK * dz), F(self.DYK * dz)
                    VIS = F(0)
                    DE = DW = DS = DN = DT = DB = F(0)
                    if diffusion:
                        zs = F(F(self.HOK[k] - self.HOK[kk]) - F(dz * h))
                        VIS = F(F(F(math.sqrt(float(self.TURB[i, j, k]))) * zs) * F(0.09))
                        m = vmax(VIS, self.VISHMIN)
                        DE = F(F(m * dyz) / self.DXK); DW = DE
                        DS = F(F(m * dxz) / self.DYK); DN = DS
                        DT = F(F(m * self.AREA) / dz)
                        DB = DT if k > kk + 1 else F(0)
                        if self.ADV[i, j - 1] < 1 or j == 2: DS = F(0)
                        if self.ADV[i, j + 1] < 1 or j == NJJ - 1: DN = F(0)
                        if self.ADV[i - 1, j] < 1 or i == 2: DW = F(0)
                        if self.ADV[i + 1, j] < 1 or i == NII - 1: DE = F(0)
                    if comp == "u":
                        FE, FW = F(UKS[i, j, k] * dyz), F(UKS[i - 1, j, k] * dyz)
                        FS = F(F(q * F(F(F(VKS[i - 1, j, k] + VKS[i, j, k]) + VKS[i - 1, j - 1, k]) + VKS[i, j - 1, k])) * dxz)
                        FN = F(F(q * F(F(F(VKS[i - 1, j, k] + VKS[i, j, k]) + VKS[i - 1, j + 1, k]) + VKS[i, j + 1, k])) * dxz)
                        FT = F(F(q * F(F(F(WKS[i - 1, j, k] + WKS[i, j, k]) + WKS[i - 1, j, k + 1]) + WKS[i, j, k + 1])) * self.AREA)
                        FB = F(F(q * F(F(F(WKS[i - 1, j, k] + WKS[i, j, k]) + WKS[i - 1, j, k - 1]) + WKS[i, j, k - 1])) * self.AREA) if k > kk + 1 else F(0)
                    elif comp == "v":
                        FE = F(F(q * F(F(F(UKS[i, j, k] + UKS[i + 1, j, k]) + UKS[i, j - 1, k]) + UKS[i + 1, j - 1, k])) * dyz)
                        FW = F(F(q * F(F(F(UKS[i, j, k] + UKS[i - 1, j, k]) + UKS[i, j - 1, k]) + UKS[i - 1, j - 1, k])) * dyz)
                        FS, FN = F(VKS[i, j - 1, k] * dxz), F(VKS[i, j, k] * dxz)
                        FT = F(F(q * F(F(F(WKS[i, j - 1, k] + WKS[i, j, k]) + WKS[i, j - 1, k + 1]) + WKS[i, j, k + 1])) * self.AREA)
                        FB = F(F(q * F(F(F(WKS[i, j - 1, k] + WKS[i, j, k]) + WKS[i, j - 1, k - 1]) + WKS[i, j, k - 1])) * self.AREA) if k > kk + 1 else F(0)
                    else:
                        FE = F(F(q * F(F(F(UKS[i, j, k] + UKS[i + 1, j, k]) + UKS[i, j, k - 1]) + UKS[i + 1, j, k - 1])) * dyz)
                        FW = F(F(q * F(F(F(UKS[i, j, k] + UKS[i - 1, j, k]) + UKS[i, j, k - 1]) + UKS[i - 1, j, k - 1])) * dyz)
                        FS = F(F(q * F(F(F(VKS[i, j, k] + VKS[i, j - 1, k]) + VKS[i, j, k - 1]) + VKS[i, j - 1, k - 1])) * dxz)
                        FN = F(F(q * F(F(F(VKS[i, j, k] + VKS[i, j + 1, k]) + VKS[i, j, k - 1]) + VKS[i, j + 1, k - 1])) * dxz)
                        FT = F(WKS[i, j, k] * self.AREA)
                        FB = F(WKS[i, j, k - 1] * self.AREA) if k > kk + 1 else F(0)
                    if diffusion:
                        lo = F(0.0001)
                        DE, DB, DW, DS, DN, DT = (vmax(v, lo) for v in (DE, DB, DW, DS, DN, DT))

                        def pl(Dv, Pv):
                            t = 1.0 - 0.1 * float(Pv)
                            return F(Dv * vmax(F(0), F(t * t * t * t * t)))
                        BIM = F(pl(DT, abs(F(FT / DT))) + vmax(F(-FT), F(0)))
                        CIM = F(pl(DB, abs(F(FB / DB))) + vmax(FB, F(0)))
                        AE1 = F(pl(DE, abs(F(FE / DE))) + vmax(F(-FE), F(0)))
                        AW1 = F(pl(DW, abs(F(FW / DW))) + vmax(FW, F(0)))
                        AS1 = F(pl(DS, abs(F(FS / DS))) + vmax(FS, F(0)))
                        AN1 = F(pl(DN, abs(F(FN / DN))) + vmax(F(-FN), F(0)))
                    else:
                        BIM, CIM, AE1 = vmax(F(-FT), F(0)), vmax(FB, F(0)), vmax(F(-FE), F(0))
                        AW1, AS1, AN1 = vmax(FW, F(0)), vmax(FS, F(0)), vmax(F(-FN), F(0))
                    ap0 = self.AP0[k]
                    AIM = F(F(F(F(F(F(BIM + CIM) + AW1) + AS1) + AE1) + AN1) + ap0)
                    part = F(F(F(F(AW1 * X[i - 1, j, k]) + F(AS1 * X[i, j - 1, k])) + F(AE1 * X[i + 1, j, k])) + F(AN1 * X[i, j + 1, k]))
                    x = X[i, j, k]
                    if comp == "w":
                        t5 = float(ap0) * 0.5 * float(F(WKS[i, j, k - 1] + WKS[i, j, k]))
                        Dd = F(float(part) + t5 + float(F(F(self.DPMNEW[i, j, k - 1] - self.DPMNEW[i, j, k]) * self.AREA)))
                    else:
                        if comp == "u":
                            t5 = float(ap0) * 0.5 * float(F(UKS[i - 1, j, k] + UKS[i, j, k]))
                            t6 = F(F(self.DPMNEW[i - 1, j, k] - self.DPMNEW[i, j, k]) * dyz)
                            G = self.UG
                        else:
                            t5 = float(ap0) * 0.5 * float(F(VKS[i, j - 1, k] + VKS[i, j, k]))
                            t6 = F(F(self.DPMNEW[i, j - 1, k] - self.DPMNEW[i, j, k]) * dxz)
                            G = self.VG
                        t7 = F(F(F(self.Cor * F(G - x)) * self.AREA) * dz)
                        Dd = F(float(part) + t5 + float(t6) + float(t7))
                        if k == kk + 1:
                            a = 0.5 * float(F(UKS[i - 1, j, k] + UKS[i, j, k]))
                            b = 0.5 * float(F(VKS[i - 1, j, k] + VKS[i, j, k]))
                            wh = vmax(F(math.sqrt(a * a + b * b)), F(0.01))
                            if self.CUTK[i, j] < 1:
                                us = F(self.HT[i, j] * wh)
                                if float(self.c.z0_solver[i, j]) >= float(dz) * 0.1:
                                    a1 = 0.5 * float(F(UKS[i - 1, j, k + 1] + UKS[i, j, k + 1]))
                                    b1 = 0.5 * float(F(VKS[i - 1, j, k + 1] + VKS[i, j, k + 1]))
                                    us = F(self.HT[i, j] * F(math.sqrt(a1 * a1 + b1 * b1)))
                            else:
                                us = F(self.HO[i, j] * wh)
                            Dd = F(Dd - F(F(F(x / wh) * F(us * us)) * self.AREA))
                    if diffusion and k > kk + 1:
                        if comp == "u":
                            e1 = F(F(X[i + 1, j, k] - x) * dyz); e2 = F(F(x - X[i - 1, j, k]) * dyz)
                            e3 = F(F(F(h * F(VKS[i + 1, j + 1, k] + VKS[i, j + 1, k])) - F(h * F(VKS[i, j + 1, k] + VKS[i - 1, j + 1, k]))) * dxz)
                            e4 = F(F(F(h * F(VKS[i + 1, j - 1, k] + VKS[i, j - 1, k])) - F(h * F(VKS[i, j - 1, k] + VKS[i - 1, j - 1, k]))) * dxz)
                            e5 = F(F(F(h * F(WK[i + 1, j, k] + WKS[i, j, k])) - F(h * F(WKS[i, j, k] + WKS[i - 1, j, k]))) * self.AREA)
                            e6 = F(F(F(h * F(WK[i + 1, j, k - 1] + WKS[i, j, k - 1])) - F(h * F(WKS[i, j, k - 1] + WKS[i - 1, j, k - 1]))) * self.AREA)
                            dv = self.DXK
                        elif comp == "v":
                            e3 = F(F(X[i, j + 1, k] - x) * dxz); e4 = F(F(x - X[i, j - 1, k]) * dxz)
                            e1 = F(F(F(h * F(UKS[i + 1, j + 1, k] + UKS[i + 1, j, k])) - F(h * F(UKS[i + 1, j, k] + UKS[i + 1, j - 1, k]))) * dxz)
                            e2 = F(F(F(h * F(UKS[i - 1, j + 1, k] + UKS[i - 1, j, k])) - F(h * F(UKS[i - 1, j, k] + UKS[i - 1, j - 1, k]))) * dxz)
                            e5 = F(F(F(h * F(WK[i, j + 1, k] + WKS[i, j, k])) - F(h * F(WKS[i, j, k] + WKS[i, j - 1, k]))) * self.AREA)
                            e6 = F(F(F(h * F(WK[i, j + 1, k - 1] + WKS[i, j, k - 1])) - F(h * F(WKS[i, j, k - 1] + WKS[i, j - 1, k - 1]))) * self.AREA)
                            dv = self.DXK
                        else:
                            e1 = F(F(F(h * F(UKS[i, j, k + 1] + UKS[i, j, k])) - F(h * F(UKS[i, j, k] + UKS[i, j, k - 1]))) * dyz)
                            e2 = F(F(F(h * F(UKS[i - 1, j, k + 1] + UKS[i - 1, j, k])) - F(h * F(UKS[i - 1, j, k] + UKS[i - 1, j, k - 1]))) * dyz)
                            e3 = F(F(F(h * F(VKS[i, j, k + 1] + VKS[i, j, k])) - F(h * F(VKS[i, j, k] + VKS[i, j, k - 1]))) * dxz)
                            e4 = F(F(F(h * F(VKS[i, j - 1, k + 1] + VKS[i, j - 1, k])) - F(h * F(VKS[i, j - 1, k] + VKS[i, j - 1, k - 1]))) * dxz)
                            e5 = F(F(WK[i, j, k + 1] - x) * self.AREA)
                            e6 = F(F(x - WK[i, j, k - 1]) * self.AREA)
                            dv = dz
                        Dd = F(Dd + F(F(VIS / dv) * F(F(F(F(F(e1 - e2) + e3) - e4) + e5) - e6)))
                    if k > kk + 1:
                        den = F(AIM - F(CIM * P[k - 1]))
                        P[k] = F(BIM / den)
                        Q[k] = F(F(Dd + F(CIM * Q[k - 1])) / den)
                    else:
                        P[k] = F(BIM / AIM)
                        Q[k] = F(Dd / AIM)
                kkn = int(self.KKART[i, j - 1]) if comp == "v" else int(self.KKART[i - 1, j])
                for k in range(vi, ks0 - 1, -1):
                    if comp == "w" or (kk < k and kkn < k):
                        X[i, j, k] = F(X[i, j, k] + F(self.relax * F(F(F(P[k] * X[i, j, k + 1]) + Q[k]) - X[i, j, k])))

    def tke_sweep(self, IS, JS):
        """TKE-prognostic-microscale.cs:25-352."""
        NII, NJJ = self.NII, self.NJJ
        UK, VK, WK, TU, TD = self.UK, self.VK, self.WK, self.TURB, self.TDISS
        h = F(0.5)
        rx, ry = F(1) / self.DXK, F(1) / self.DYK
        for i1 in range(2, NII):
            for j1 in range(2, NJJ):
                j = NJJ - j1 + 1 if JS == -1 else j1
                i = NII - i1 + 1 if IS == -1 else i1
                if self.ADV[i, j] != 1:
                    continue
                kk, vi = int(self.KKART[i, j]), int(self.VI[i, j])
                ks0 = 1 if self.CUTK[i, j] != 0 else kk + 1
                z0 = F(self.c.z0_solver[i, j])
                n = self.KADVMAX + 2
                P1, Q1, P2, Q2 = (np.zeros(n, dtype=F) for _ in range(4))
                for k in range(ks0, vi + 1):
                    dz = self.DZK[k]
                    dxz, dyz, rz = F(self.DXK * dz), F(self.DYK * dz), F(F(1) / dz)
                    zs = F(F(self.HOK[k] - self.HOK[kk]) - F(dz * h))
                    VIS = F(F(F(math.sqrt(float(TU[i, j, k]))) * zs) * F(0.09))
                    VIS = F(net_min(VIS, F(15)))
                    m = F(net_max(VIS, self.VISHMIN))
                    DE = F(F(m * dyz) * rx); DW = DE
                    DS = F(F(m * dxz) * ry); DN = DS
                    DT = F(F(m * self.AREA) * rz); DB = F(0)
                    DTE = F(F(F(net_max(F(VIS / F(1)), self.VISVMIN)) * self.AREA) * rz); DBE = F(0)
                    if k > kk + 1:
                        DB, DBE = DT, DTE
                    if self.ADV[i, j - 1] < 1 or j == 2: DS = F(0)
                    if self.ADV[i, j + 1] < 1 or j == NJJ - 1: DN = F(0)
                    if self.ADV[i - 1, j] < 1 or i == 2: DW = F(0)
                    if self.ADV[i + 1, j] < 1 or i == NII - 1: DE = F(0)
                    FE, FW = F(UK[i + 1, j, k] * dyz), F(UK[i, j, k] * dyz)
                    FS, FN = F(VK[i, j, k] * dxz), F(VK[i, j + 1, k] * dxz)
                    FT = F(WK[i, j, k + 1] * self.AREA)
                    FB = F(WK[i, j, k] * self.AREA) if k > kk + 1 else F(0)
                    lo = F(0.0001)
                    DE, DB, DW, DS, DN, DT, DBE, DTE = (F(net_max(v, lo)) for v in (DE, DB, DW, DS, DN, DT, DBE, DTE))

                    def pl(Dv, Pv):
                        t = F(F(1) - F(F(0.1) * Pv))
                        return F(Dv * F(net_max(F(0), F(F(F(F(t * t) * t) * t) * t))))
                    PE, PB, PW = abs(F(FE / DE)), abs(F(FB / DB)), abs(F(FW / DW))
                    PS, PN, PT = abs(F(FS / DS)), abs(F(FN / DN)), abs(F(FT / DT))
                    PBE, PTE = abs(F(FN / DBE)), abs(F(FT / DTE))
                    BIM = F(pl(DT, PT) + F(net_max(F(-FT), F(0))))
                    CIM = F(pl(DB, PB) + F(net_max(FB, F(0))))
                    AE1 = F(pl(DE, PE) + F(net_max(F(-FE), F(0))))
                    AW1 = F(pl(DW, PW) + F(net_max(FW, F(0))))
                    AS1 = F(pl(DS, PS) + F(net_max(FS, F(0))))
                    AN1 = F(pl(DN, PN) + F(net_max(F(-FN), F(0))))
                    BIME = F(pl(DT, PTE) + F(net_max(F(-FT), F(0))))
                    CIME = F(pl(DB, PBE) + F(net_max(FB, F(0))))
                    ap0 = self.AP0[k]
                    AIM = F(F(F(F(F(F(BIM + CIM) + AW1) + AS1) + AE1) + AN1) + ap0)
                    VX1 = F(h * F(VK[i + 1, j, k] + VK[i + 1, j + 1, k])); VX2 = F(h * F(VK[i - 1, j, k] + VK[i - 1, j + 1, k]))
                    WX1 = F(h * F(WK[i + 1, j, k] + WK[i + 1, j, k + 1])); WX2 = F(h * F(WK[i - 1, j, k] + WK[i - 1, j, k + 1]))
                    UY1 = F(h * F(UK[i, j + 1, k] + UK[i + 1, j + 1, k])); UY2 = F(h * F(UK[i, j - 1, k] + UK[i + 1, j - 1, k]))
                    WY1 = F(h * F(WK[i, j + 1, k] + WK[i, j + 1, k + 1])); WY2 = F(h * F(WK[i, j - 1, k] + WK[i, j - 1, k + 1]))
                    DUDX = F(F(UK[i + 1, j, k] - UK[i, j, k]) * rx)
                    DVDX = F(F(F(VX1 - VX2) * rx) * h)
                    DWDX = F(F(F(WX1 - WX2) * rx) * h)
                    DUDY = F(F(F(UY1 - UY2) * ry) * h)
                    DVDY = F(F(VK[i, j + 1, k] - VK[i, j, k]) * ry)
                    DWDY = F(F(F(WY1 - WY2) * ry) * h)
                    if k > kk + 1:
                        DUDZ = F(F(F(UK[i, j, k + 1] - UK[i, j, k - 1]) * rz) * h)
                        DVDZ = F(F(F(VK[i, j, k + 1] - VK[i, j, k - 1]) * rz) * h)
                        DWDZ = F(F(F(WK[i, j, k + 1] - WK[i, j, k - 1]) * rz) * h)
                    else:
                        DUDZ = F(F(UK[i, j, k + 1] * rz) * h)
                        DVDZ = F(F(VK[i, j, k + 1] * rz) * h)
                        DWDZ = F(F(WK[i, j, k + 1] * rz) * h)
                    s = F(F(2) * F(DUDX * DUDX))
                    s = F(s + F(F(2) * F(DVDY * DVDY)))
                    s = F(s + F(F(2) * F(DWDZ * DWDZ)))
                    s = F(s + F(F(DUDY + DVDX) * F(DUDY + DVDX)))
                    s = F(s + F(F(DUDZ + DWDX) * F(DUDZ + DWDX)))
                    s = F(s + F(F(DVDZ + DWDY) * F(DVDZ + DWDY)))
                    s = F(s - TD[i, j, k])
                    PRO = F(F(F(VIS * self.AREA) * dz) * s)
                    PROB = F((((((float(F(-F(self.c.potdtdz))) * 9.81) / 283.0) * float(VIS)) * 1.35) * float(self.AREA)) * float(dz))
                    DT1 = F(F(F(F(F(F(F(AW1 * TU[i - 1, j, k]) + F(AS1 * TU[i, j - 1, k])) + F(AE1 * TU[i + 1, j, k])) + F(AN1 * TU[i, j + 1, k]))
                                + F(ap0 * TU[i, j, k])) + PRO) + PROB)
                    DT2 = F(F(F(F(F(AW1 * TD[i - 1, j, k]) + F(AS1 * TD[i, j - 1, k])) + F(AE1 * TD[i + 1, j, k])) + F(AN1 * TD[i, j + 1, k]))
                            + F(ap0 * TD[i, j, k]))
                    DT2 = F(DT2 + F(F(TD[i, j, k] / TU[i, j, k]) * F(F(F(1.44) * PRO) - F(F(1.92) * TD[i, j, k]))))
                    if k > kk + 1:
                        den1 = F(AIM - F(CIM * P1[k - 1])); den2 = F(AIM - F(CIME * P2[k - 1]))
                        P1[k] = F(BIM / den1); Q1[k] = F(F(DT1 + F(CIM * Q1[k - 1])) / den1)
                        P2[k] = F(BIME / den2); Q2[k] = F(F(DT2 + F(CIME * Q2[k - 1])) / den2)
                    else:
                        P1[k] = F(BIM / AIM); Q1[k] = F(DT1 / AIM); P2[k] = F(BIME / AIM); Q2[k] = F(DT2 / AIM)
                for k in range(vi, ks0 - 1, -1):
                    if not (kk < k):
                        continue
                    dz, dz1 = float(self.DZK[k]), float(self.DZK[k + 1])
                    c = self.CUTK
                    if k == kk + 1 and c[i, j] == 0:
                        if float(z0) >= dz * 0.5:
                            hz = dz + dz1 * 0.5
                            a = 0.5 * float(F(UK[i, j, k + 1] + UK[i + 1, j, k + 1])); b = 0.5 * float(F(VK[i, j, k + 1] + VK[i, j + 1, k + 1]))
                            us = F(0.4 / math.log(hz / float(z0)) * math.sqrt(a * a + b * b))
                            TU[i, j, k] = F(3.33 * float(F(us * us)))
                            TD[i, j, k] = F(float(F(F(us * us) * us)) / (hz * 0.4))
                        else:
                            a = 0.5 * float(F(UK[i, j, k + 1] + UK[i + 1, j, k])); b = 0.5 * float(F(VK[i, j, k] + VK[i, j + 1, k]))
                            us = F(0.4 / math.log(dz * 0.5 / float(z0)) * math.sqrt(a * a + b * b))
                            TU[i, j, k] = F(3.33 * float(F(us * us)))
                            TD[i, j, k] = F(float(F(F(us * us) * us)) / (dz * 0.5 * 0.4))
                    elif k == kk + 1 and c[i, j] > 0:
                        a = 0.5 * float(F(UK[i, j, k] + UK[i + 1, j, k])); b = 0.5 * float(F(VK[i, j, k] + VK[i, j + 1, k]))
                        us = F(self.HO[i, j] * F(math.sqrt(a * a + b * b)))
                        TU[i, j, k] = F(3.33 * float(F(us * us)))
                        TD[i, j, k] = F(float(F(F(us * us) * us)) / (dz * 0.5 * 0.4))
                    elif (k == self.KKART[i + 1, j] + 1 and c[i + 1, j] > 0) or (k == self.KKART[i - 1, j] + 1 and c[i - 1, j] > 0):
                        a = 0.5 * float(F(WK[i, j, k] + WK[i, j, k + 1])); b = 0.5 * float(F(VK[i, j, k] + VK[i, j + 1, k]))
                        us = F(float(self.UFX) * math.sqrt(a * a + b * b))
                        TU[i, j, k] = F(3.33 * float(F(us * us)))
                        TD[i, j, k] = F(float(F(F(us * us) * us)) / (float(self.bz0) * 0.4))
                    elif (k == self.KKART[i, j + 1] + 1 and c[i, j + 1] > 0) or (k == self.KKART[i, j - 1] + 1 and c[i, j - 1] > 0):
                        a = 0.5 * float(F(WK[i, j, k] + WK[i, j, k + 1])); b = 0.5 * float(F(UK[i, j, k] + UK[i + 1, j, k]))
                        us = F(float(self.UFX) * math.sqrt(a * a + b * b))
                        TU[i, j, k] = F(3.33 * float(F(us * us)))
                        TD[i, j, k] = F(float(F(F(us * us) * us)) / (float(self.bz0) * 0.4))
                    else:
                        TU[i, j, k] = F(F(P1[k] * TU[i, j, k + 1]) + Q1[k])
                        TD[i, j, k] = F(F(P2[k] * TD[i, j, k + 1]) + Q2[k])
                    TU[i, j, k] = F(net_max(float(TU[i, j, k]), 0.001))
                    TD[i, j, k] = F(net_max(float(TD[i, j, k]), 0.0000001))

    # -------------------------------------------------------------- driver
    def run(self):
        self.prologue()
        hist = []
        for it in range(1, self.iters + 1):
            self.boundary()
            self.divergence()
            IS, JS = {0: (-1, 1), 1: (-1, -1), 2: (1, -1), 3: (1, 1)}[it % 4]
            hist.append(self.pressure(IS, JS))
            self.correction()
            self.averages()
            if self.model == 1:
                self.ml_sweep("u", IS, JS)
                self.ml_sweep("v", -IS, -JS)
                self.ml_sweep("w", IS, JS)
            else:
                d = self.model == 2
                self.ke_sweep("u", IS, JS, d)
                self.ke_sweep("v", IS, JS, d)
                self.ke_sweep("w", IS, JS, d)
                if d:
                    self.tke_sweep(IS, JS)
        out = {n.lower(): getattr(self, n) for n in ("UK", "VK", "WK", "UKS", "VKS", "WKS", "DPM", "DPMNEW", "TURB", "TDISS")}
        out["delta_umax_hist"] = np.array(hist, dtype=F)
        return out


def run(case, iterations):
    return Solver(case, iterations).run()


# ----------------------------------------------------------------------------
# DiagnosticFlowfield.Calculate (DiagnosticFlowfield.cs:24-436) and the final
# vertical-velocity mass adjustment of MicroscaleFlat.cs:388-462.  Stencil
# passes are whole-array float32 numpy operations (elementwise IEEE single
# precision, same operation order as the C#); the Gauss-Seidel pressure sweep
# loops over columns in Python.
# ----------------------------------------------------------------------------
def diagnostic(case, fields, level, iterations=0, terrain=False):
    c = case
    NII, NJJ, NKK = c.nii, c.njj, c.nkk
    DXK, DYK = F(c.dxk), F(c.dyk)
    DZK = c.dzk.astype(F)
    KK = c.kkart.astype(np.int64)
    UK = fields["uk"].astype(F).copy()
    VK = fields["vk"].astype(F).copy()
    WK = fields["wk"].astype(F).copy()
    DPM = fields["dpm"].astype(F).copy()
    DIV = np.zeros_like(UK)
    u, v = UK[1, 1, c.kadvmax], VK[1, 1, c.kadvmax]
    s2 = F(F(u * u) + F(v * v))
    DTIME = F(0.5 * float(net_min(DXK, DYK)) / math.sqrt(float(s2)) if s2 > 0 else math.inf)
    DTIME = F(net_min(float(DTIME), 5.0))
    n_it = iterations if iterations > 0 else (10 if level == 2 else 5)
    DPM[1:NII + 1, 1:NJJ + 1, 1:NKK + 2] = 0
    I = slice(2, NII)
    J = slice(2, NJJ)
    ks = np.arange(NKK + 2)
    above = ks[None, None, :] > KK[:, :, None]                       # KKART < k
    for it in range(1, n_it + 1):
        # divergence over every interior column, k = 1..NKK-1
        dz = DZK[None, None, 1:NKK]
        t1 = F(UK[I, J, 1:NKK] - UK[3:NII + 1, J, 1:NKK]) * DYK * dz
        t2 = F(VK[I, J, 1:NKK] - VK[I, 3:NJJ + 1, 1:NKK]) * DXK * dz
        top = WK[I, J, 2:NKK + 1] * DXK * DYK
        inner = t1 + t2 + F(WK[I, J, 1:NKK] - WK[I, J, 2:NKK + 1]) * DXK * DYK
        bottom = t1 + t2 - top
        kk = KK[I, J][:, :, None]
        k3 = ks[None, None, 1:NKK]
        new = np.where(k3 > kk + 1, inner, bottom).astype(F)
        DIV[I, J, 1:NKK] = np.where(k3 > kk, new, DIV[I, J, 1:NKK])
        # pressure sweep
        IS, JS = {0: (1, 1), 1: (1, -1), 2: (-1, -1), 3: (-1, 1)}[it % 4]
        PIMP = np.zeros(NKK + 2, dtype=F)
        QIMP = np.zeros(NKK + 2, dtype=F)
        for j1 in range(2, NJJ):
            j = NJJ - j1 + 1 if JS == -1 else j1
            for i1 in range(2, NII):
                i = NII - i1 + 1 if IS == -1 else i1
                kb = int(KK[i, j])
                for k in range(NKK - 1, kb, -1):
                    dzk = DZK[k]
                    DIM = F(F(F(DIV[i, j, k] / DTIME) + F(F(F(DPM[i - 1, j, k] + DPM[i + 1, j, k]) / DXK) * DYK) * dzk)
                            + F(F(F(F(DPM[i, j - 1, k] - DPM[i, j + 1, k]) / DYK) * DXK) * dzk))
                    APP = F(F(2) * F(F(F(F(DYK * dzk) / DXK) + F(F(DXK * dzk) / DYK)) + F(F(DXK * DYK) / dzk)))
                    if k > kb + 1:
                        ABP = F(F(DXK * DYK) / F(F(0.5) * F(DZK[k - 1] + dzk)))
                    else:
                        ABP = F(F(DXK * DYK) / dzk)
                    ATP = F(F(DXK * DYK) / F(F(0.5) * F(DZK[k + 1] + dzk)))
                    TERMP = F(F(1) / F(APP - F(ATP * PIMP[k + 1])))
                    PIMP[k] = F(ABP * TERMP)
                    QIMP[k] = F(F(F(ATP * QIMP[k + 1]) + DIM) * TERMP)
                DPM[i, j, 0] = 0
                for k in range(1, NKK):
                    DPM[i, j, k] = F(F(PIMP[k] * DPM[i, j, k - 1]) + QIMP[k]) if kb < k else F(0)
        # velocity correction, whole interior at once
        d0 = DPM[I, J, 1:NKK]
        DDPX = F(DPM[1:NII - 1, J, 1:NKK] - d0)
        DDPY = F(DPM[I, 1:NJJ - 1, 1:NKK] - d0)
        DDPZ = np.where(k3 > kk + 1, F(DPM[I, J, 0:NKK - 1] - d0), F(0)).astype(F)
        okU = (k3 > kk) & (k3 > KK[1:NII - 1, J][:, :, None])
        okV = (k3 > kk) & (k3 > KK[I, 1:NJJ - 1][:, :, None])
        UK[I, J, 1:NKK] = np.where(okU, F(UK[I, J, 1:NKK] + F(F(DDPX / DXK) * DTIME)), F(0))
        VK[I, J, 1:NKK] = np.where(okV, F(VK[I, J, 1:NKK] + F(F(DDPY / DYK) * DTIME)), F(0))
        dzs = F(DZK[None, None, 1:NKK] + DZK[None, None, 0:NKK - 1])
        WK[I, J, 1:NKK] = np.where(k3 > kk + 1, F(WK[I, J, 1:NKK] + F(F(F(DDPZ * F(2)) / dzs) * DTIME)), F(0))
    # vertical velocity from the mass balance of each cell (serial in k per column)
    for i in range(2, NII):
        for j in range(2, NJJ):
            if not terrain and not (level <= 1 or c.advdom[i, j] == 0):
                continue
            DSN = DYK if terrain else DXK
            kb = int(KK[i, j])
            for k in range(1, NKK):
                if k <= kb:
                    continue
                fwo1 = F(0) if k <= KK[i - 1, j] else F(F(DZK[k] * DYK) * UK[i, j, k])
                fwo2 = F(0) if k <= KK[i + 1, j] else F(F(DZK[k] * DYK) * UK[i + 1, j, k])
                fsn1 = F(0) if k <= KK[i, j - 1] else F(F(DZK[k] * DSN) * VK[i, j, k])
                fsn2 = F(0) if k <= KK[i, j + 1] else F(F(DZK[k] * DSN) * VK[i, j + 1, k])
                fbt1 = F(0) if (k <= kb + 1 or k == 1) else F(F(DXK * DYK) * WK[i, j, k])
                WK[i, j, k + 1] = F(F(F(F(F(fwo1 - fwo2) + fsn1) - fsn2) + fbt1) / F(DXK * DYK))
    return {"uk": UK, "vk": VK, "wk": WK, "dpm": DPM, "div": DIV, "dtime": float(DTIME)}
