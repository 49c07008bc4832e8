"""Flow-field files (GFFWrite.cs / GFFRead.cs), CPU side: the payload oracle
pinned by hand-computed values, the header layout, the size formulas and the
reader restated from GFFRead.cs; the zip container through the C ABI
(host-only code, no GPU call), read back with Python's zipfile.

The reference ships no *.gff or *.tke files, so there are no golden payloads:
the oracle is pinned by known answers and by the writer/reader round trip
(GFFWrite.cs and GFFRead.cs are independent code in the reference, so a
transposed layout or a shifted level fails the round trip)."""
import os
import struct
import zipfile

import numpy as np
import pytest

from oracle import gff_oracle as G

F = np.float32


def _fields(nii, njj, nkk, seed, scale=3.0):
    r = np.random.default_rng(seed)
    shp = (nii + 2, njj + 2, nkk + 2)
    return [(r.standard_normal(shp) * scale).astype(F) for _ in range(4)]


def _kkart(nii, njj, nkk, seed):
    r = np.random.default_rng(seed)
    kk = np.zeros((nii + 2, njj + 2), np.int16)
    kk[1:nii + 2, 1:njj + 2] = r.integers(0, max(2, nkk // 2), size=(nii + 1, njj + 1))
    return kk


def test_conversion_known_answers():
    # value*100 is formed in FP32: 0.125*100 = 12.5 exactly, 0.375*100 = 37.5
    v = np.array([0.125, 0.375, -0.125, -0.375, 1.999, -1.999, 400.0, -400.0, 0.0, -0.004, 327.67, -327.68],
                 dtype=F) * F(100)
    assert G._trunc16(v).tolist() == [12, 37, -12, -37, 199, -199, 32767, -32768, 0, 0, 32767, -32768]
    # MathF.Round / Convert.ToInt16: half to even
    assert G._round16(v).tolist() == [12, 38, -12, -38, 200, -200, 32767, -32768, 0, 0, 32767, -32768]
    assert G._round16(np.array([2.5, 3.5, -2.5, -0.5, 0.49999997], F)).tolist() == [2, 4, -2, 0, 0]
    assert G._trunc16(np.array([np.nan], F)).tolist() == [0]


def test_header_layout():
    h = G.header(1, nii=7, njj=5, nkk=12, dxk=2.0, wind_dir=235.5, wind_vel=3.25, stab_class=4)
    assert h == struct.pack("<3i2fifi", -12, 5, 7, 235.5, 3.25, 4, 2.0, -2)
    assert G.header(0, 7, 5, 12, 2.0, 1, 2, 3)[:4] == struct.pack("<i", 12)
    assert G.header(0, 7, 5, 12, 2.0, 1, 2, 3)[-4:] == struct.pack("<i", -1)
    assert G.header(2, 7, 5, 12, 2.0, 1, 2, 3)[:4] == struct.pack("<i", -12)
    assert G.header(2, 7, 5, 12, 2.0, 1, 2, 3)[-4:] == struct.pack("<i", -3)
    assert G.header(3, 7, 5, 12, 2.0, 1, 2, 3)[-4:] == struct.pack("<i", -1)
    assert len(h) == 32


def test_payload_sizes():
    nii, njj, nkk = 6, 4, 9
    u, v, w, t = _fields(nii, njj, nkk, 1)
    kk = _kkart(nii, njj, nkk, 2)
    col = (nii + 1) * (njj + 1)
    assert len(G.payload(0, u, v, w, kk, nii, njj, nkk, 2.0)) == 32 + col * (nkk + 1) * 6
    assert len(G.payload(1, u, v, w, kk, nii, njj, nkk, 2.0)) == 32 + col * (nkk + 1) * 6
    k0 = np.maximum(kk[1:, 1:].astype(int), 1)
    assert len(G.payload(2, u, v, w, kk, nii, njj, nkk, 2.0)) == 32 + nii * njj * 2 + int((6 * (nkk + 2 - k0)).sum())
    assert len(G.payload(3, u, v, w, kk, nii, njj, nkk, 2.0, turb=t)) == 32 + nii * njj * nkk * 2


def test_first_record_by_hand():
    # kind 0 starts with (UK[1][1][0], VK[1][1][0], WK[1][1][1]); kind 1 with the
    # row UK[1][j][0], j = 1..NJJ+1
    nii, njj, nkk = 3, 2, 4
    u, v, w, _ = _fields(nii, njj, nkk, 3)
    b = G.payload(0, u, v, w, None, nii, njj, nkk, 1.0)
    r = struct.unpack_from("<3h", b, 32)
    assert r == (int(np.trunc(u[1, 1, 0] * F(100))), int(np.trunc(v[1, 1, 0] * F(100))), int(np.trunc(w[1, 1, 1] * F(100))))
    b = G.payload(1, u, v, w, None, nii, njj, nkk, 1.0)
    row = struct.unpack_from(f"<{njj + 1}h", b, 32)
    assert list(row) == [int(np.trunc(u[1, j, 0] * F(100))) for j in range(1, njj + 2)]
    tke = G.payload(3, u, v, w, None, nii, njj, nkk, 1.0, turb=np.abs(u))
    first = struct.unpack_from("<2h", tke, 32)
    assert list(first) == [int(np.rint(abs(u[1, 1, k]) * F(100))) for k in (1, 2)]


@pytest.mark.parametrize("kind", [0, 1, 2])
def test_round_trip_through_reader(kind):
    nii, njj, nkk = 7, 5, 11
    u, v, w, _ = _fields(nii, njj, nkk, 10 + kind)
    u[2, 3, 4] = F(500.0)                     # clamps
    v[4, 1, 2] = F(-500.0)
    kk = _kkart(nii, njj, nkk, 20 + kind)
    data = G.payload(kind, u, v, w, kk, nii, njj, nkk, 2.0, 123.0, 4.5, 6)
    rd = G.read_payload(data, kk if kind == 2 else None)
    assert (rd["nkk"], rd["njj"], rd["nii"], rd["stab"], rd["dxk"]) == (nkk, njj, nii, 6, 2.0)
    assert (rd["dir"], rd["speed"]) == (123.0, 4.5)
    conv = G._round16 if kind == 2 else G._trunc16
    s = F(0.01)
    for name, a, lev in (("uk", u, slice(0, nkk + 1)), ("vk", v, slice(0, nkk + 1)), ("wk", w, slice(1, nkk + 2))):
        want = np.zeros_like(a)
        want[1:, 1:, lev] = conv(a[1:, 1:, lev] * F(100)).astype(F) * s
        if kind == 2:
            # the reader zeroes the levels below max(KKART, 1) (GFFRead.cs:285-290)
            for i in range(1, nii + 2):
                for j in range(1, njj + 2):
                    k0 = max(int(kk[i, j]), 1)
                    lo = (0, k0 - 1) if name != "wk" else (1, k0)
                    want[i, j, lo[0]:lo[1]] = 0
        assert np.array_equal(rd[name], want), name


def test_mode1_is_a_permutation_of_mode0():
    nii, njj, nkk = 5, 6, 7
    u, v, w, _ = _fields(nii, njj, nkk, 30)
    a = np.frombuffer(G.payload(0, u, v, w, None, nii, njj, nkk, 1.0)[32:], "<i2")
    b = np.frombuffer(G.payload(1, u, v, w, None, nii, njj, nkk, 1.0)[32:], "<i2")
    assert not np.array_equal(a, b)
    assert np.array_equal(np.sort(a), np.sort(b))


def test_a_shifted_w_level_fails_the_round_trip():
    # a plausible writer mistake: WK at level k-1 like UK and VK
    nii, njj, nkk = 4, 4, 6
    u, v, w, _ = _fields(nii, njj, nkk, 40)
    good = G.payload(0, u, v, w, None, nii, njj, nkk, 1.0)
    bad = bytearray(good)
    rec = np.frombuffer(good[32:], "<i2").reshape(nii + 1, njj + 1, nkk + 1, 3).copy()
    rec[..., 2] = G._trunc16(w[1:, 1:, 0:nkk + 1] * F(100))
    bad[32:] = rec.astype("<i2").tobytes()
    rd = G.read_payload(bytes(bad))
    assert not np.array_equal(rd["wk"][1:, 1:, 1:nkk + 2], G._trunc16(w[1:, 1:, 1:nkk + 2] * F(100)).astype(F) * F(0.01))


# ---- zip container through the C ABI (host code only) ----------------------

def _zip(tmp_path, name, data):
    from gral_amd import pf
    p = os.path.join(tmp_path, name)
    pf.zip_write(p, name, data)
    with zipfile.ZipFile(p) as z:
        assert z.testzip() is None            # CRC-32 of the entry checks
        infos = z.infolist()
        assert [i.filename for i in infos] == [name]
        assert infos[0].compress_type == zipfile.ZIP_DEFLATED
        return z.read(name), infos[0]


@pytest.mark.parametrize("n", [0, 1, 1000, (4 << 20) - 1, 4 << 20, (9 << 20) + 12345])
def test_zip_round_trip_sizes(tmp_path, n):
    r = np.random.default_rng(n)
    # compressible (int16 of a smooth field) so chunk boundaries land inside deflate runs
    data = (np.cumsum(r.integers(-3, 4, size=(n + 1) // 2)).astype("<i2").tobytes())[:n]
    got, info = _zip(str(tmp_path), "00001.gff", data)
    assert got == data
    assert info.file_size == n
    if n > 100000:
        assert info.compress_size < n


def test_zip_incompressible(tmp_path):
    data = np.random.default_rng(5).integers(0, 256, size=(5 << 20) + 7, dtype=np.uint8).tobytes()
    got, _ = _zip(str(tmp_path), "x.tke", data)
    assert got == data


def test_zip_of_an_oracle_payload(tmp_path):
    nii, njj, nkk = 9, 8, 10
    u, v, w, _ = _fields(nii, njj, nkk, 50)
    kk = _kkart(nii, njj, nkk, 51)
    data = G.payload(2, u, v, w, kk, nii, njj, nkk, 2.0)
    got, _ = _zip(str(tmp_path), "00002.gff", data)
    assert got == data
    rd = G.read_payload(got, kk)
    assert rd["flag"] == -3


def test_zip_bad_arguments(tmp_path):
    import ctypes as C
    from gral_amd import pf
    lib = pf.load()
    assert lib.gral_pf_zip_write(None, b"a", b"x", 1) != 0
    assert lib.gral_pf_zip_write(os.fsencode(os.path.join(str(tmp_path), "a.zip")), b"", b"x", 1) != 0
    # an unwritable path leaves nothing behind
    bad = os.path.join(str(tmp_path), "no_such_dir", "a.gff")
    assert lib.gral_pf_zip_write(os.fsencode(bad), b"a.gff", b"x", 1) != 0
    assert not os.path.exists(bad)
    assert lib.gral_pf_zip_write(os.fsencode(os.path.join(str(tmp_path), "b.zip")), b"b", C.c_void_p(None), 0) == 0
