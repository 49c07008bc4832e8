"""ctypes binding of the CPU oracle (pf_oracle.c).

TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker.  The product (gral_amd) never
imports this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle_pf.so")


def _structs(ft):
    """The oracle's structures with ``ft`` as the floating-point type: c_float
    for liboracle_pf.so, c_double for the FP64 shadow liboracle_pf64.so."""
    FPt = C.POINTER(ft)

    class OrcParams(C.Structure):
        _fields_ = [("nii", C.c_int), ("njj", C.c_int), ("nkk", C.c_int), ("kadvmax", C.c_int),
                    ("vertcellsff", C.c_int), ("dxk", ft), ("dyk", ft),
                    ("turbulence_model", C.c_int), ("simd_width", C.c_int), ("iter_min", C.c_int),
                    ("iter_max", C.c_int), ("relax", ft), ("relaxp", ft),
                    ("building_z0", ft), ("coriolis", ft), ("potdtdz", ft),
                    ("std_dev_v", ft), ("std_dev_w", ft), ("bdlayheight", ft),
                    ("v0_1", ft), ("refp_i", C.c_int), ("refp_j", C.c_int)]

    class OrcInputs(C.Structure):
        _fields_ = [("dzk", FPt), ("hokart", FPt), ("advdom", U8P), ("kkart", I16P), ("cutk", FPt),
                    ("cov", FPt), ("ahk", FPt), ("z0_wall", FPt), ("z0_solver", FPt), ("ob_init", FPt),
                    ("ust_init", FPt), ("z0_init", FPt), ("veg", FPt), ("veg_present", U8P),
                    ("istatistics", C.c_int), ("n_profile", C.c_int), ("meas_height", FPt), ("u0", FPt),
                    ("v0", FPt)]

    class OrcFields(C.Structure):
        _fields_ = [("uk", FPt), ("vk", FPt), ("wk", FPt), ("uks", FPt), ("vks", FPt), ("wks", FPt),
                    ("dpm", FPt), ("dpmnew", FPt), ("div", FPt), ("turb", FPt), ("tdiss", FPt)]

    class OrcStats(C.Structure):
        _fields_ = [("iteration_loops", C.c_int), ("iterations_run", C.c_int),
                    ("delta_first", ft), ("delta_finish", ft), ("topwind", ft),
                    ("dtime", ft), ("ug", ft), ("vg", ft),
                    ("ustern_factor_x", ft), ("frund", C.c_double), ("geschw_max", C.c_double),
                    ("unreal_u", C.c_int), ("unreal_v", C.c_int), ("unreal_w", C.c_int),
                    ("convergence_problem", C.c_int), ("delta_umax_hist", FPt)]

    class OrcCapture(C.Structure):
        _fields_ = [("comp", C.c_int), ("i", C.c_int), ("j", C.c_int), ("iteration", C.c_int),
                    ("a", FPt), ("b", FPt), ("c", FPt), ("d", FPt), ("x_old", FPt), ("x_new", FPt),
                    ("captured", C.c_int), ("piece", FPt)]

    class OrcDiagParams(C.Structure):
        _fields_ = [("nii", C.c_int), ("njj", C.c_int), ("nkk", C.c_int), ("kadvmax", C.c_int), ("dxk", ft),
                    ("dyk", ft), ("flow_field_level", C.c_int), ("iterations", C.c_int), ("terrain", C.c_int)]

    return OrcParams, OrcInputs, OrcFields, OrcStats, OrcCapture, OrcDiagParams


FP = C.POINTER(C.c_float)
U8P = C.POINTER(C.c_uint8)
I16P = C.POINTER(C.c_int16)
OrcParams, OrcInputs, OrcFields, OrcStats, OrcCapture, OrcDiagParams = _structs(C.c_float)
_structs_f32 = (OrcParams, OrcInputs, OrcFields, OrcStats)
# the per-level pieces the checking hook records (enum of pf_oracle.h, in order)
PIECES = ("vis", "de", "dw", "ds", "dn", "dt", "db", "fe", "fw", "fs", "fn", "ft", "fb",
          "ae", "aw", "as", "an", "at", "ab", "press", "cor", "wall", "cross", "vegt", "prod", "buoy")
_S64 = None


_lib = None


_other = {}


def load(build_if_missing: bool = True, path: str = None):
    """The oracle library; ``path`` loads another build of pf_oracle.c (the
    mutant builds of the oracle's own tests) with the same interface."""
    global _lib
    if path is not None:
        if path not in _other:
            _other[path] = _configure(C.CDLL(path))
        return _other[path]
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH) and build_if_missing:
        subprocess.run(["make", "-s", "-C", HERE], check=True)
    _lib = _configure(C.CDLL(LIB_PATH))
    return _lib


def _configure(_lib):
    _lib.orc_calculate.argtypes = [C.POINTER(OrcParams), C.POINTER(OrcInputs), C.POINTER(OrcFields),
                                   C.POINTER(OrcStats)]
    _lib.orc_calculate.restype = C.c_int
    _lib.orc_calculate_capture.argtypes = [C.POINTER(OrcParams), C.POINTER(OrcInputs), C.POINTER(OrcFields),
                                           C.POINTER(OrcStats), C.POINTER(OrcCapture)]
    _lib.orc_calculate_capture.restype = C.c_int
    _lib.orc_calculate_threads.argtypes = [C.POINTER(OrcParams), C.POINTER(OrcInputs), C.POINTER(OrcFields),
                                           C.POINTER(OrcStats), C.c_int]
    _lib.orc_calculate_threads.restype = C.c_int
    for n in ("orc_powlaw_vector", "orc_powlaw_scalar"):
        getattr(_lib, n).argtypes = [C.c_float, C.c_float]
        getattr(_lib, n).restype = C.c_float
    _lib.orc_conv_to_int.argtypes = [C.c_double]
    _lib.orc_conv_to_int.restype = C.c_int
    _lib.orc_diagnostic.argtypes = [C.POINTER(OrcDiagParams), FP, U8P, I16P, FP, FP, FP, FP, FP, FP]
    _lib.orc_diagnostic.restype = C.c_int
    return _lib


SIMD_PATH = os.path.join(HERE, "liboracle_pf_simd.so")


def simd_available() -> bool:
    """The AVX2 timing build exists and this CPU runs AVX2 code."""
    try:
        flags = open("/proc/cpuinfo").read()
    except OSError:
        return False
    return os.path.exists(SIMD_PATH) and " avx2" in flags and " fma" in flags


def load64(build_if_missing: bool = True):
    """The FP64 shadow (liboracle_pf64.so): the same restatement with every
    float widened to double; for error statements only, never a checker."""
    global _lib64, _S64
    if _lib64 is not None:
        return _lib64
    path = os.path.join(HERE, "liboracle_pf64.so")
    if not os.path.exists(path) and build_if_missing:
        subprocess.run(["make", "-s", "-C", HERE], check=True)
    _S64 = _structs(C.c_double)
    P, I, F, S = _S64[:4]
    lib = C.CDLL(path)
    lib.orc_calculate.argtypes = [C.POINTER(P), C.POINTER(I), C.POINTER(F), C.POINTER(S)]
    lib.orc_calculate.restype = C.c_int
    _lib64 = lib
    return lib


_lib64 = None


def _p(a, t=FP):
    if a is None:
        return C.cast(None, t)
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(t)


def run(case, iterations=None, simd_width=None, keep_history=True, capture=None, threads=1, fp64=False,
        lib_path=None, project=None):
    """Run PrognosticFlowfield.Calculate on ``case`` (a gral_amd.domain.Case)
    with the deterministic CPU oracle.  Returns a dict of output fields and stats.
    The input case is not modified.  threads > 1 selects the timing mode: the
    reference's default multi-threaded schedule, not deterministic (CPU
    baseline only, never a checker).  fp64 runs the FP64 shadow instead: the
    same algorithm in double precision (fields come back as float64).
    project = 0 / 1 runs orc_project instead: the pressure-correction part of
    each iteration only, without (0) or with (1) the quantisation."""
    if fp64:
        lib = load64()
        OrcParams, OrcInputs, OrcFields, OrcStats = _S64[:4]
        dt, FPt = np.float64, C.POINTER(C.c_double)
        assert capture is None and threads == 1
    else:
        lib = load(path=lib_path)
        OrcParams, OrcInputs, OrcFields, OrcStats = _structs_f32
        dt, FPt = np.float32, FP
    c = case
    nit = c.iter_max if iterations is None else iterations
    prm = OrcParams(c.nii, c.njj, c.nkk, c.kadvmax, c.vertcellsff, c.dxk, c.dyk, c.turbulence_model,
                    c.simd_width if simd_width is None else simd_width,
                    c.iter_min if iterations is None else iterations, nit, c.relax, c.relaxp,
                    c.building_z0, c.coriolis, c.potdtdz, c.std_dev_v, c.std_dev_w, c.bdlayheight,
                    c.v0_1, int(c.refp[0]), int(c.refp[1]))
    f32 = lambda a: np.ascontiguousarray(a, dtype=dt)
    pf_ = lambda a: _p(a, FPt)
    keep = dict(dzk=f32(c.dzk), hokart=f32(c.hokart), advdom=np.ascontiguousarray(c.advdom, dtype=np.uint8),
                kkart=np.ascontiguousarray(c.kkart, dtype=np.int16), cutk=f32(c.cutk), cov=f32(c.cov),
                ahk=f32(c.ahk), z0_wall=f32(c.z0_wall), z0_solver=f32(c.z0_solver), ob_init=f32(c.ob_init),
                ust_init=f32(c.ust_init), z0_init=f32(c.z0_init),
                veg=None if c.veg is None else f32(c.veg),
                veg_present=None if c.veg_present is None else np.ascontiguousarray(c.veg_present, dtype=np.uint8))
    prof = c.istatistics in (0, 3)
    for n in ("meas_height", "u0", "v0"):
        keep[n] = f32(getattr(c, n)) if prof else None
    inp = OrcInputs(pf_(keep["dzk"]), pf_(keep["hokart"]), _p(keep["advdom"], U8P), _p(keep["kkart"], I16P),
                    pf_(keep["cutk"]), pf_(keep["cov"]), pf_(keep["ahk"]), pf_(keep["z0_wall"]),
                    pf_(keep["z0_solver"]), pf_(keep["ob_init"]), pf_(keep["ust_init"]), pf_(keep["z0_init"]),
                    pf_(keep["veg"]), _p(keep["veg_present"], U8P), int(c.istatistics),
                    len(keep["meas_height"]) - 1 if prof else 0, pf_(keep["meas_height"]), pf_(keep["u0"]),
                    pf_(keep["v0"]))
    shp = c.uk.shape
    out = {n: np.zeros(shp, dtype=dt) for n in ("uks", "vks", "wks", "dpm", "dpmnew", "div", "turb", "tdiss")}
    out["uk"] = f32(c.uk).copy()
    out["vk"] = f32(c.vk).copy()
    out["wk"] = f32(c.wk).copy()
    fld = OrcFields(*[pf_(out[n]) for n in ("uk", "vk", "wk", "uks", "vks", "wks", "dpm", "dpmnew", "div", "turb", "tdiss")])
    hist = np.zeros(max(nit, 1), dtype=dt)
    st = OrcStats()
    st.delta_umax_hist = pf_(hist) if keep_history else C.cast(None, FPt)
    cap = None
    if capture is not None:
        comp, ci, cj, cit = capture
        nk = c.nkk + 2
        capbuf = {n: np.zeros(nk, dtype=np.float32) for n in ("a", "b", "c", "d", "x_old", "x_new")}
        pieces = np.zeros((len(PIECES), nk), dtype=np.float32)
        cap = OrcCapture(comp, ci, cj, cit, *[_p(capbuf[n]) for n in ("a", "b", "c", "d", "x_old", "x_new")], 0,
                         _p(pieces))
        rc = lib.orc_calculate_capture(C.byref(prm), C.byref(inp), C.byref(fld), C.byref(st), C.byref(cap))
    elif project is not None:
        lib.orc_project.argtypes = [C.POINTER(OrcParams), C.POINTER(OrcInputs), C.POINTER(OrcFields),
                                    C.POINTER(OrcStats), C.c_int]
        lib.orc_project.restype = C.c_int
        rc = lib.orc_project(C.byref(prm), C.byref(inp), C.byref(fld), C.byref(st), int(project))
    elif threads > 1:
        rc = lib.orc_calculate_threads(C.byref(prm), C.byref(inp), C.byref(fld), C.byref(st), int(threads))
    else:
        rc = lib.orc_calculate(C.byref(prm), C.byref(inp), C.byref(fld), C.byref(st))
    if rc != 0:
        raise RuntimeError(f"orc_calculate failed: {rc}")
    if cap is not None:
        capbuf["captured"] = bool(cap.captured)
        capbuf["pieces"] = {n: pieces[q] for q, n in enumerate(PIECES)}
        out["capture"] = capbuf
    stats = {n: getattr(st, n) for n, _ in OrcStats._fields_ if n != "delta_umax_hist"}
    stats["delta_umax_hist"] = hist[: st.iterations_run].copy()
    out["stats"] = stats
    return out


def diagnostic(case, fields=None, flow_field_level=2, iterations=0, terrain=False):
    """DiagnosticFlowfield.Calculate() and the final vertical-velocity mass
    adjustment of MicroscaleFlat on ``fields`` (dict with uk, vk, wk, dpm; by
    default the case's initial wind field and zero DPM).  Returns new arrays
    uk, vk, wk, dpm, div and the time step used."""
    lib = load()
    c = case
    f32 = lambda a: np.ascontiguousarray(a, dtype=np.float32).copy()
    src = fields if fields is not None else {"uk": c.uk, "vk": c.vk, "wk": c.wk, "dpm": np.zeros_like(c.uk)}
    out = {n: f32(src[n]) for n in ("uk", "vk", "wk", "dpm")}
    out["div"] = np.zeros_like(out["uk"])
    dzk = np.ascontiguousarray(c.dzk, dtype=np.float32)
    adv = np.ascontiguousarray(c.advdom, dtype=np.uint8)
    kk = np.ascontiguousarray(c.kkart, dtype=np.int16)
    prm = OrcDiagParams(c.nii, c.njj, c.nkk, c.kadvmax, c.dxk, c.dyk, flow_field_level, iterations, int(bool(terrain)))
    dt = C.c_float(0)
    rc = lib.orc_diagnostic(C.byref(prm), _p(dzk), _p(adv, U8P), _p(kk, I16P), _p(out["uk"]), _p(out["vk"]),
                            _p(out["wk"]), _p(out["dpm"]), _p(out["div"]), C.cast(C.pointer(dt), FP))
    if rc != 0:
        raise RuntimeError(f"orc_diagnostic failed: {rc}")
    out["dtime"] = dt.value
    return out
