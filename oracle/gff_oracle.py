"""CPU restatement of GRAL's flow-field file payloads (*.gff, *.tke).

TEST INFRASTRUCTURE ONLY: the checker of the GPU packer behind
gral_pf_gff_payload / gral_pf_gff_write.  The product never imports it.

Follows GFFWrite.cs (paths relative to the reference's src/):
  * kind 0: WriteCompressedGFFFiles, header -1, column-major records of
    (u, v, w) per level (:167-225);
  * kind 1: WriteStrongCompressedGFFFiles, header -2 and -NKK, level-major rows
    of u, then v, then w over j (:256-360);
  * kind 2: WriteTerrainCompressedGFFFiles, header -3 and -NKK, KKART, then per
    column u, v, w from max(KKART, 1) up, rounded (:388-510);
  * kind 3: the *.tke file of the k-eps model, header -1 (:79-133).
Value conversions (all on FP32 value*100):
  * kind 0: Math.Max(Int16.MinValue, Math.Min(Int16.MaxValue, v)) in float,
    then (Int16): truncation toward zero;
  * kind 1: v > 32767 -> 32767, v < -32768 -> -32768, else (Int16)v;
  * kind 2: MathF.Round(v) (half to even), then clamped as kind 1;
  * kind 3: Convert.ToInt16 of the clamped value: round half to even.
A NaN converts to 0 here; what .NET gives for (Int16)NaN is unspecified by the
language ("parity unpinned" for NaN inputs, which the solver does not produce).
The reader (read_payload) restates GFFRead.cs:181-321 for round-trip checks.
"""
from __future__ import annotations

import struct

import numpy as np

F = np.float32
HEADER = struct.Struct("<iiiffifi")      # NKK, NJJ, NII, dir, speed, stability class, DXK, format flag


def _trunc16(v):
    v = np.asarray(v, dtype=F)
    c = np.minimum(np.maximum(v, F(-32768.0)), F(32767.0))          # NaN stays NaN
    out = np.trunc(np.nan_to_num(c, nan=0.0)).astype(np.int16)
    return out


def _round16(v):
    v = np.rint(np.asarray(v, dtype=F))                             # MathF.Round: half to even
    c = np.minimum(np.maximum(v, F(-32768.0)), F(32767.0))
    return np.nan_to_num(c, nan=0.0).astype(np.int16)


def header(kind, nii, njj, nkk, dxk, wind_dir, wind_vel, stab_class):
    flag = {0: -1, 1: -2, 2: -3, 3: -1}[kind]
    n = -nkk if kind in (1, 2) else nkk
    return HEADER.pack(int(n), int(njj), int(nii), float(F(wind_dir)), float(F(wind_vel)), int(stab_class),
                       float(F(dxk)), flag)


def payload(kind, uk, vk, wk, kkart, nii, njj, nkk, dxk, wind_dir=270.0, wind_vel=3.0, stab_class=4, turb=None):
    """Bytes inside the zip entry of a *.gff (kind 0, 1, 2) or *.tke (kind 3) file.
    Fields are [nii+2][njj+2][nkk+2] float32, kkart [nii+2][njj+2]."""
    h = header(kind, nii, njj, nkk, dxk, wind_dir, wind_vel, stab_class)
    I = slice(1, nii + 2)
    J = slice(1, njj + 2)
    if kind == 3:
        t = _round16(np.asarray(turb, dtype=F)[1:nii + 1, 1:njj + 1, 1:nkk + 1] * F(100))
        return h + t.astype("<i2").tobytes()
    u = np.asarray(uk, dtype=F)[I, J, 0:nkk + 1] * F(100)             # UK[i][j][k-1], k = 1..NKK+1
    v = np.asarray(vk, dtype=F)[I, J, 0:nkk + 1] * F(100)
    w = np.asarray(wk, dtype=F)[I, J, 1:nkk + 2] * F(100)             # WK[i][j][k]
    if kind == 0:
        rec = np.stack([_trunc16(u), _trunc16(v), _trunc16(w)], axis=-1)        # [i][j][k][comp]
        return h + rec.astype("<i2").tobytes()
    if kind == 1:
        rows = np.stack([_trunc16(u), _trunc16(v), _trunc16(w)], axis=0)       # [comp][i][j][k]
        rows = np.transpose(rows, (3, 1, 0, 2))                                  # [k][i][comp][j]
        return h + np.ascontiguousarray(rows).astype("<i2").tobytes()
    # kind 2
    kk = np.asarray(kkart)
    out = [h, kk[1:nii + 1, 1:njj + 1].astype("<i2").tobytes()]
    ru, rv, rw = _round16(u), _round16(v), _round16(w)
    for a in range(nii + 1):
        for b in range(njj + 1):
            k0 = max(int(kk[a + 1, b + 1]), 1)                           # levels k0..NKK+1 -> index k0-1..NKK
            out.append(ru[a, b, k0 - 1:].astype("<i2").tobytes())
            out.append(rv[a, b, k0 - 1:].astype("<i2").tobytes())
            out.append(rw[a, b, k0 - 1:].astype("<i2").tobytes())
    return b"".join(out)


def read_payload(data, kkart=None):
    """GFFRead.cs:181-321 restated: header values and the fields it fills
    (u, v at k-1, w at k for k = 1..NKK+1; value*0.01f)."""
    nkk, njj, nii, wdir, wvel, ak, dxk, flag = HEADER.unpack_from(data, 0)
    nkk = abs(nkk)
    body = np.frombuffer(data, dtype="<i2", offset=HEADER.size)
    shp = (nii + 2, njj + 2, nkk + 2)
    uk, vk, wk = np.zeros(shp, F), np.zeros(shp, F), np.zeros(shp, F)
    s = F(0.01)
    if flag == -1:
        r = body.reshape(nii + 1, njj + 1, nkk + 1, 3).astype(F)
        uk[1:, 1:, 0:nkk + 1] = r[..., 0] * s
        vk[1:, 1:, 0:nkk + 1] = r[..., 1] * s
        wk[1:, 1:, 1:nkk + 2] = r[..., 2] * s
    elif flag == -2:
        r = body.reshape(nkk + 1, nii + 1, 3, njj + 1).astype(F)
        uk[1:, 1:, 0:nkk + 1] = np.transpose(r[:, :, 0, :], (1, 2, 0)) * s
        vk[1:, 1:, 0:nkk + 1] = np.transpose(r[:, :, 1, :], (1, 2, 0)) * s
        wk[1:, 1:, 1:nkk + 2] = np.transpose(r[:, :, 2, :], (1, 2, 0)) * s
    else:
        kk = body[:nii * njj].reshape(nii, njj)
        if kkart is not None:
            assert np.array_equal(kk, np.asarray(kkart)[1:nii + 1, 1:njj + 1])
        p = nii * njj
        for a in range(nii + 1):
            for b in range(njj + 1):
                k0 = max(int(kkart[a + 1, b + 1]) if kkart is not None else 1, 1)
                n = nkk + 2 - k0
                uk[a + 1, b + 1, k0 - 1:nkk + 1] = body[p:p + n].astype(F) * s
                vk[a + 1, b + 1, k0 - 1:nkk + 1] = body[p + n:p + 2 * n].astype(F) * s
                wk[a + 1, b + 1, k0:nkk + 2] = body[p + 2 * n:p + 3 * n].astype(F) * s
                p += 3 * n
        assert p == body.size
    return {"nkk": nkk, "njj": njj, "nii": nii, "dir": wdir, "speed": wvel, "stab": ak, "dxk": dxk, "flag": flag,
            "uk": uk, "vk": vk, "wk": wk}
