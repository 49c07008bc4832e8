"""ctypes binding of libgral_pf.so (the C ABI declared in include/gral_pf.h).

``PrognosticFlowfield`` mirrors the reference entry point
``PrognosticFlowfield.Calculate()`` (src/PrognosticFlowfield.cs:32): it takes
one situation (a ``gral_amd.domain.Case``: the state GRAL's Program.* statics
hold before the call), runs the SIMPLE loop on the GPU and returns the
updated fields.  The library must be built (``gral_amd.build``); there is no
CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libgral_pf.so")

FP = C.POINTER(C.c_float)
U8P = C.POINTER(C.c_uint8)
I16P = C.POINTER(C.c_int16)

EXPORTS = ["gral_pf_create", "gral_pf_set_masks", "gral_pf_set_roughness", "gral_pf_upload_fields",
           "gral_pf_set_met_profile", "gral_pf_prepare", "gral_pf_iterate", "gral_pf_mass_imbalance",
           "gral_pf_mass_imbalance_noalloc", "gral_pf_forward_ckpt", "gral_pf_restore_ckpt", "gral_pf_vjp",
           "gral_pf_momentum_sweep", "gral_pf_momentum_vjp", "gral_pf_set_velocity", "gral_pf_finish", "gral_pf_solve", "gral_pf_download_fields",
           "gral_pf_diagnostic",
           "gral_pf_last_timing", "gral_pf_profile_iterations", "gral_pf_batch_create", "gral_pf_batch_iterate",
           "gral_pf_batch_solve", "gral_pf_batch_profile_iterations", "gral_pf_batch_last_timing",
           "gral_pf_batch_add", "gral_pf_batch_step", "gral_pf_batch_size",
           "gral_pf_gff_payload", "gral_pf_gff_write", "gral_pf_zip_write", "gral_pf_wind_at_particles",
           "gral_pf_subdomains", "gral_pf_reference_point", "gral_pf_flat_init", "gral_pf_terrain_init",
           "gral_pf_terrain_init_topo", "gral_pf_setup_last_error",
           "gral_pf_last_error", "gral_pf_destroy", "gral_pf_batch_last_error", "gral_pf_batch_destroy"]
FAMILIES = ("stencil", "pressure", "u", "v", "w", "tke", "coef")


class Grid(C.Structure):
    _fields_ = [("nii", C.c_int32), ("njj", C.c_int32), ("nkk", C.c_int32), ("kadvmax", C.c_int32),
                ("vertcellsff", C.c_int32), ("dxk", C.c_float), ("dyk", C.c_float), ("dzk", FP), ("hokart", FP)]


class Params(C.Structure):
    _fields_ = [("turbulence_model", C.c_int32), ("simd_width", C.c_int32), ("iter_min", C.c_int32),
                ("iter_max", C.c_int32), ("relax", C.c_float), ("relaxp", C.c_float), ("building_z0", C.c_float),
                ("coriolis", C.c_float), ("potdtdz", C.c_float), ("std_dev_v", C.c_float), ("std_dev_w", C.c_float),
                ("bdlayheight", C.c_float), ("v0_1", C.c_float), ("refp_i", C.c_int32), ("refp_j", C.c_int32)]


class Stats(C.Structure):
    _fields_ = [("iteration_loops", C.c_int32), ("iterations_run", C.c_int32), ("delta_first", C.c_float),
                ("delta_finish", C.c_float), ("convergence_problem", C.c_int32), ("unreal_u", C.c_int32),
                ("unreal_v", C.c_int32), ("unreal_w", C.c_int32), ("topwind", C.c_float), ("dtime", C.c_float),
                ("ug", C.c_float), ("vg", C.c_float), ("frund", C.c_double), ("geschw_max", C.c_double),
                ("last_delta", C.c_double), ("mass_div", C.c_double), ("mass_flux", C.c_double)]


class Timing(C.Structure):
    _fields_ = [("ms_total", C.c_double), ("launches", C.c_int64), ("active_cells", C.c_int64),
                ("cell_iterations", C.c_int64), ("situations", C.c_int32), ("groups", C.c_int32)]


class FamilyTime(C.Structure):
    _fields_ = [("ms", C.c_double), ("launches", C.c_int64), ("cells", C.c_int64)]


_lib = None


class GffMeta(C.Structure):
    _fields_ = [("wind_dir", C.c_float), ("wind_vel", C.c_float), ("stab_class", C.c_int32)]


class FlatMet(C.Structure):
    _fields_ = [("n_profile", C.c_int32), ("meas_height", C.POINTER(C.c_float)), ("obs_u", C.POINTER(C.c_float)),
                ("obs_v", C.POINTER(C.c_float)), ("ob_length", C.c_float), ("ol_column", C.POINTER(C.c_float)),
                ("near_wall", C.c_int32)]


class Gramm(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32), ("ddx", C.c_float), ("ddy", C.c_float),
                ("delta_i", C.c_int32), ("delta_j", C.c_int32), ("ahe", C.POINTER(C.c_float)),
                ("uwin", C.POINTER(C.c_float)), ("vwin", C.POINTER(C.c_float)), ("wwin", C.POINTER(C.c_float)),
                ("zsp", C.POINTER(C.c_float)), ("ah", C.POINTER(C.c_float)), ("zko1", C.c_double),
                ("zko2", C.c_double)]


# flow-field file kinds: GRAL_FlowFields.txt writing modes 0, 1, 2 and the *.tke file
GFF_KINDS = {"mode0": 0, "mode1": 1, "mode2": 2, "tke": 3}


def load(path: str = LIB_PATH):
    """Load the HIP library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"{path} not found: build it with `python -m gral_amd.build` (no CPU fallback exists)")
    lib = C.CDLL(path)
    H = C.c_void_p
    lib.gral_pf_create.argtypes = [C.POINTER(Grid), C.POINTER(H)]
    lib.gral_pf_set_masks.argtypes = [H, U8P, I16P, FP, FP, FP, FP, U8P]
    lib.gral_pf_set_roughness.argtypes = [H, FP, FP, FP, FP, FP]
    lib.gral_pf_upload_fields.argtypes = [H, FP, FP, FP]
    lib.gral_pf_set_met_profile.argtypes = [H, C.c_int32, C.c_int32, FP, FP, FP]
    lib.gral_pf_mass_imbalance.argtypes = [H, C.POINTER(C.c_double), C.POINTER(C.c_double)]
    lib.gral_pf_mass_imbalance_noalloc.argtypes = [H, C.POINTER(C.c_double), C.POINTER(C.c_double)]
    lib.gral_pf_forward_ckpt.argtypes = [H, C.c_int32, C.c_int32]
    lib.gral_pf_restore_ckpt.argtypes = [H]
    lib.gral_pf_vjp.argtypes = [H, FP, FP, FP, FP, FP, FP]
    lib.gral_pf_momentum_sweep.argtypes = [H, C.c_int32]
    lib.gral_pf_momentum_vjp.argtypes = [H, FP, FP]
    lib.gral_pf_set_velocity.argtypes = [H, FP, FP, FP]
    lib.gral_pf_prepare.argtypes = [H, C.POINTER(Params)]
    lib.gral_pf_iterate.argtypes = [H, C.c_int32, C.c_int32, C.POINTER(Stats)]
    lib.gral_pf_finish.argtypes = [H, C.POINTER(Stats)]
    lib.gral_pf_solve.argtypes = [H, C.POINTER(Params), C.POINTER(Stats)]
    lib.gral_pf_download_fields.argtypes = [H] + [FP] * 10
    lib.gral_pf_diagnostic.argtypes = [H, C.c_int32, C.c_int32]
    lib.gral_pf_last_timing.argtypes = [H, C.POINTER(Timing)]
    lib.gral_pf_profile_iterations.argtypes = [H, C.c_int32, C.POINTER(FamilyTime), C.c_int32]
    lib.gral_pf_batch_create.argtypes = [C.POINTER(H), C.c_int32, C.POINTER(H)]
    lib.gral_pf_batch_iterate.argtypes = [H, C.c_int32, C.c_int32, C.POINTER(Stats)]
    lib.gral_pf_batch_solve.argtypes = [H, C.POINTER(Stats)]
    lib.gral_pf_batch_profile_iterations.argtypes = [H, C.c_int32, C.POINTER(FamilyTime), C.c_int32]
    lib.gral_pf_batch_last_timing.argtypes = [H, C.POINTER(Timing)]
    lib.gral_pf_batch_add.argtypes = [H, H]
    lib.gral_pf_batch_step.argtypes = [H, C.POINTER(C.c_int32), C.POINTER(H), C.POINTER(Stats), C.c_int32]
    lib.gral_pf_batch_size.argtypes = [H, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]
    lib.gral_pf_gff_payload.argtypes = [H, C.c_int32, C.POINTER(GffMeta), C.c_void_p, C.c_size_t,
                                        C.POINTER(C.c_size_t)]
    lib.gral_pf_gff_write.argtypes = [H, C.c_int32, C.POINTER(GffMeta), C.c_char_p]
    lib.gral_pf_zip_write.argtypes = [C.c_char_p, C.c_char_p, C.c_void_p, C.c_size_t]
    lib.gral_pf_subdomains.argtypes = [C.c_int32, C.c_int32, C.c_float, C.c_float, C.c_int32, FP, U8P]
    lib.gral_pf_reference_point.argtypes = [C.c_int32, C.c_int32, U8P, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]
    lib.gral_pf_flat_init.argtypes = [C.POINTER(Grid), C.POINTER(FlatMet), FP, U8P, FP, FP, FP, I16P, FP, FP,
                                      C.POINTER(C.c_int32)]
    lib.gral_pf_terrain_init.argtypes = [C.POINTER(Grid), C.POINTER(Gramm), FP, U8P, C.c_int32, FP, FP, FP, I16P, FP,
                                         FP, FP, C.POINTER(C.c_int32), C.POINTER(C.c_double)]
    lib.gral_pf_terrain_init_topo.argtypes = [C.POINTER(Grid), C.POINTER(Gramm), FP, C.c_float, FP, U8P, C.c_int32,
                                              FP, FP, FP, I16P, FP, FP, FP, C.POINTER(C.c_int32),
                                              C.POINTER(C.c_double)]
    lib.gral_pf_setup_last_error.argtypes = []
    lib.gral_pf_setup_last_error.restype = C.c_char_p
    I32P, F64P = C.POINTER(C.c_int32), C.POINTER(C.c_double)
    lib.gral_pf_wind_at_particles.argtypes = [H, C.c_int32, C.c_int32, C.c_int32, C.c_float, C.c_int32, I32P, I32P, FP,
                                              F64P, F64P, FP, FP, FP, FP, I32P]
    for n in ("gral_pf_last_error", "gral_pf_batch_last_error"):
        getattr(lib, n).argtypes = [H]
        getattr(lib, n).restype = C.c_char_p
    for n in ("gral_pf_destroy", "gral_pf_batch_destroy"):
        getattr(lib, n).argtypes = [H]
        getattr(lib, n).restype = None
    for n in EXPORTS[:-4]:
        if n != "gral_pf_setup_last_error":
            getattr(lib, n).restype = C.c_int
    _lib = lib
    return lib


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _ptr(a, t=FP):
    if a is None:
        return C.cast(None, t)
    return a.ctypes.data_as(t)


class PrognosticFlowfield:
    """One situation on the GPU: create -> set_masks/roughness/fields -> prepare
    -> iterate -> finish -> download."""

    def __init__(self, case):
        self.lib = load()
        self.case = case
        c = case
        self._keep = {"dzk": _f32(c.dzk), "hokart": _f32(c.hokart)}
        g = Grid(c.nii, c.njj, c.nkk, c.kadvmax, c.vertcellsff, c.dxk, c.dyk, _ptr(self._keep["dzk"]),
                 _ptr(self._keep["hokart"]))
        h = C.c_void_p()
        rc = self.lib.gral_pf_create(C.byref(g), C.byref(h))
        if rc != 0:
            raise RuntimeError(f"gral_pf_create failed ({rc})")
        self.h = h
        k = self._keep
        k.update(advdom=np.ascontiguousarray(c.advdom, dtype=np.uint8), kkart=np.ascontiguousarray(c.kkart, dtype=np.int16),
                 cutk=_f32(c.cutk), cov=_f32(c.cov), ahk=_f32(c.ahk))
        veg = None if c.veg is None else _f32(c.veg)
        vegp = None if c.veg_present is None else np.ascontiguousarray(c.veg_present, dtype=np.uint8)
        k.update(veg=veg, vegp=vegp)
        self._check(self.lib.gral_pf_set_masks(h, _ptr(k["advdom"], U8P), _ptr(k["kkart"], I16P), _ptr(k["cutk"]),
                                               _ptr(k["cov"]), _ptr(k["ahk"]), _ptr(veg), _ptr(vegp, U8P)))
        k.update(z0w=_f32(c.z0_wall), z0s=_f32(c.z0_solver), obi=_f32(c.ob_init), usti=_f32(c.ust_init), z0i=_f32(c.z0_init))
        self._check(self.lib.gral_pf_set_roughness(h, _ptr(k["z0w"]), _ptr(k["z0s"]), _ptr(k["obi"]), _ptr(k["usti"]),
                                                   _ptr(k["z0i"])))
        if c.istatistics in (0, 3):
            k.update(mh=_f32(c.meas_height), u0=_f32(c.u0), v0=_f32(c.v0))
            self._check(self.lib.gral_pf_set_met_profile(h, int(c.istatistics), len(k["mh"]) - 1, _ptr(k["mh"]),
                                                         _ptr(k["u0"]), _ptr(k["v0"])))
        elif c.istatistics != 4:
            self._check(self.lib.gral_pf_set_met_profile(h, int(c.istatistics), 0, None, None, None))
        uk, vk, wk = _f32(c.uk), _f32(c.vk), _f32(c.wk)
        self._check(self.lib.gral_pf_upload_fields(h, _ptr(uk), _ptr(vk), _ptr(wk)))

    def _check(self, rc):
        if rc != 0:
            msg = self.lib.gral_pf_last_error(self.h)
            raise RuntimeError(f"gral_pf call failed ({rc}): {msg.decode() if msg else ''}")

    def params(self, iterations=None):
        c = self.case
        it_min = c.iter_min if iterations is None else iterations
        it_max = c.iter_max if iterations is None else iterations
        return Params(c.turbulence_model, c.simd_width, it_min, it_max, c.relax, c.relaxp, c.building_z0, c.coriolis,
                      c.potdtdz, c.std_dev_v, c.std_dev_w, c.bdlayheight, c.v0_1, int(c.refp[0]), int(c.refp[1]))

    def prepare(self, iterations=None):
        self._params = self.params(iterations)
        self._check(self.lib.gral_pf_prepare(self.h, C.byref(self._params)))

    def iterate(self, n, check_convergence=False):
        st = Stats()
        self._check(self.lib.gral_pf_iterate(self.h, int(n), int(bool(check_convergence)), C.byref(st)))
        return _stats(st)

    def mass_imbalance(self):
        """(sum |DIV|, sum |face flux|) over the active cells of the current
        device fields, reduced in FP64 on the GPU."""
        a, b = C.c_double(0), C.c_double(0)
        self._check(self.lib.gral_pf_mass_imbalance(self.h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def forward_ckpt(self, n, quantise=False):
        """n pressure-correction steps (divergence, pressure sweep, velocity
        correction) on the device fields, keeping a checkpoint of the start."""
        self._check(self.lib.gral_pf_forward_ckpt(self.h, int(n), int(bool(quantise))))

    def restore_ckpt(self):
        self._check(self.lib.gral_pf_restore_ckpt(self.h))

    def set_velocity(self, uk, vk, wk):
        """Replace UK, VK, WK of the prepared situation (full-depth arrays),
        keeping its derived fields and iteration count."""
        uk, vk, wk = _f32(uk), _f32(vk), _f32(wk)
        self._check(self.lib.gral_pf_set_velocity(self.h, _ptr(uk), _ptr(vk), _ptr(wk)))

    def momentum_sweep(self, comp):
        """One mixing-length momentum sweep of component comp (0 U, 1 V, 2 W) in
        the directions of the next iteration, keeping the fields before it."""
        self._check(self.lib.gral_pf_momentum_sweep(self.h, int(comp)))

    def momentum_vjp(self, dx):
        """dL/dX before the last momentum_sweep from dL/dX after it (X the swept
        component; full-depth array), the other fields held fixed."""
        dx = _f32(dx)
        gx = np.zeros_like(dx)
        self._check(self.lib.gral_pf_momentum_vjp(self.h, _ptr(dx), _ptr(gx)))
        return gx

    def vjp(self, du, dv, dw):
        """dL/d(UK, VK, WK) at the start of the last forward_ckpt window from
        dL/d(UK, VK, WK) at its end (full-depth arrays)."""
        du, dv, dw = _f32(du), _f32(dv), _f32(dw)
        gu, gv, gw = np.zeros_like(du), np.zeros_like(dv), np.zeros_like(dw)
        self._check(self.lib.gral_pf_vjp(self.h, _ptr(du), _ptr(dv), _ptr(dw), _ptr(gu), _ptr(gv), _ptr(gw)))
        return gu, gv, gw

    def finish(self):
        st = Stats()
        self._check(self.lib.gral_pf_finish(self.h, C.byref(st)))
        return _stats(st)

    def solve(self, iterations=None):
        """The whole Calculate(): prologue, SIMPLE loop with the reference's
        convergence test, epilogue checks."""
        self._params = self.params(iterations)
        st = Stats()
        self._check(self.lib.gral_pf_solve(self.h, C.byref(self._params), C.byref(st)))
        return _stats(st)

    def timing(self):
        t = Timing()
        self._check(self.lib.gral_pf_last_timing(self.h, C.byref(t)))
        return {n: getattr(t, n) for n, _ in Timing._fields_}

    def profile_families(self, n):
        """Run n more iterations with every launch bracketed by HIP events;
        returns {family: {ms, launches, cells}} for the families that ran."""
        arr = (FamilyTime * len(FAMILIES))()
        self._check(self.lib.gral_pf_profile_iterations(self.h, int(n), arr, len(FAMILIES)))
        return {f: {"ms": arr[q].ms, "launches": arr[q].launches, "cells": arr[q].cells}
                for q, f in enumerate(FAMILIES) if arr[q].launches > 0}

    def diagnostic(self, flow_field_level=2, terrain=False):
        """The mass-conservation post-pass (DiagnosticFlowfield + final W step)."""
        self._check(self.lib.gral_pf_diagnostic(self.h, int(flow_field_level), int(bool(terrain))))

    def gff_payload_into(self, kind, out, wind_dir=270.0, wind_vel=3.0, stab_class=4):
        """Packs the payload of a flow-field file into out (a contiguous uint8
        numpy array of at least gff_size(kind) bytes); returns its length."""
        meta = GffMeta(float(wind_dir), float(wind_vel), int(stab_class))
        n = C.c_size_t(0)
        assert out.dtype == np.uint8 and out.flags["C_CONTIGUOUS"]
        self._check(self.lib.gral_pf_gff_payload(self.h, int(kind), C.byref(meta), out.ctypes.data_as(C.c_void_p),
                                                 out.nbytes, C.byref(n)))
        return n.value

    def gff_size(self, kind):
        meta = GffMeta(0.0, 0.0, 0)
        n = C.c_size_t(0)
        self._check(self.lib.gral_pf_gff_payload(self.h, int(kind), C.byref(meta), None, 0, C.byref(n)))
        return n.value

    def gff_payload(self, kind, wind_dir=270.0, wind_vel=3.0, stab_class=4):
        """The bytes GFFWrite.cs puts into a flow-field file's zip entry (kind 0,
        1, 2: writing modes of GRAL_FlowFields.txt; 3: the *.tke file),
        packed on the GPU."""
        buf = np.empty(self.gff_size(kind), dtype=np.uint8)
        n = self.gff_payload_into(kind, buf, wind_dir, wind_vel, stab_class)
        return buf[:n].tobytes()

    def write_gff(self, path, kind, wind_dir=270.0, wind_vel=3.0, stab_class=4):
        """GFFWrite.Write() for one file: payload packed on the GPU, written as
        a zip archive with one entry named like the file."""
        meta = GffMeta(float(wind_dir), float(wind_vel), int(stab_class))
        self._check(self.lib.gral_pf_gff_write(self.h, int(kind), C.byref(meta), os.fsencode(path)))

    def wind_at_particles(self, topo, ikooagral, jkooagral, ahmin, cell_i, cell_j, ahint, x, y, z):
        """The particle model's wind at particle positions on the current fields
        (Intwind.cs:31-170) through gral_pf_wind_at_particles; returns (ux, uy, uz, index_k)."""
        ci, cj = np.ascontiguousarray(cell_i, np.int32), np.ascontiguousarray(cell_j, np.int32)
        ah, zz = np.ascontiguousarray(ahint, np.float32), np.ascontiguousarray(z, np.float32)
        xx, yy = np.ascontiguousarray(x, np.float64), np.ascontiguousarray(y, np.float64)
        n = len(ci)
        ux, uy, uz = np.zeros(n, np.float32), np.zeros(n, np.float32), np.zeros(n, np.float32)
        ck = np.zeros(n, np.int32)
        I32P, F64P = C.POINTER(C.c_int32), C.POINTER(C.c_double)
        self._check(self.lib.gral_pf_wind_at_particles(
            self.h, int(topo), int(ikooagral), int(jkooagral), float(ahmin), n, _ptr(ci, I32P), _ptr(cj, I32P),
            _ptr(ah), _ptr(xx, F64P), _ptr(yy, F64P), _ptr(zz), _ptr(ux), _ptr(uy), _ptr(uz), _ptr(ck, I32P)))
        return ux, uy, uz, ck

    def download(self):
        shp = self.case.uk.shape
        names = ("uk", "vk", "wk", "uks", "vks", "wks", "dpm", "dpmnew", "turb", "tdiss")
        out = {n: np.zeros(shp, dtype=np.float32) for n in names}
        self._check(self.lib.gral_pf_download_fields(self.h, *[_ptr(out[n]) for n in names]))
        return out

    def close(self):
        if getattr(self, "h", None):
            self.lib.gral_pf_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Several prepared situations advanced in lock step by the same launches
    (gral_pf_batch_*).  Each situation keeps its own convergence test and
    leaves the batch when it stops; results equal plain solves bit for bit.
    ``add``/``step`` feed a running batch (solvers admitted at multiples of
    100 iterations, finished ones handed back)."""

    def __init__(self, solvers=()):
        self.lib = load()
        self.solvers = list(solvers)
        self._by_handle = {s.h.value: s for s in self.solvers}
        arr = (C.c_void_p * max(1, len(self.solvers)))(*[s.h.value for s in self.solvers])
        b = C.c_void_p()
        rc = self.lib.gral_pf_batch_create(arr, len(self.solvers), C.byref(b))
        if rc != 0:
            raise RuntimeError(f"gral_pf_batch_create failed ({rc})")
        self.b = b

    def _check(self, rc):
        if rc != 0:
            msg = self.lib.gral_pf_batch_last_error(self.b)
            raise RuntimeError(f"gral_pf_batch call failed ({rc}): {msg.decode() if msg else ''}")

    def iterate(self, n, check_convergence=False):
        st = (Stats * len(self.solvers))()
        self._check(self.lib.gral_pf_batch_iterate(self.b, int(n), int(bool(check_convergence)), st))
        return [_stats(x) for x in st]

    def solve(self):
        st = (Stats * len(self.solvers))()
        self._check(self.lib.gral_pf_batch_solve(self.b, st))
        return [_stats(x) for x in st]

    def add(self, solver):
        """Queue a prepared solver; it joins at the next admission point."""
        self._check(self.lib.gral_pf_batch_add(self.b, solver.h))
        self._by_handle[solver.h.value] = solver

    def size(self):
        a, q = C.c_int32(0), C.c_int32(0)
        self._check(self.lib.gral_pf_batch_size(self.b, C.byref(a), C.byref(q)))
        return a.value, q.value

    def step(self):
        """Advance to the next multiple of 100 iterations; returns [(solver, stats)]
        of the situations that stopped there (already finished)."""
        cap = max(1, len(self._by_handle))
        done = (C.c_void_p * cap)()
        st = (Stats * cap)()
        n = C.c_int32(0)
        self._check(self.lib.gral_pf_batch_step(self.b, C.byref(n), done, st, cap))
        out = []
        for q in range(n.value):
            sv = self._by_handle.pop(done[q])
            out.append((sv, _stats(st[q])))
        return out

    def timing(self):
        t = Timing()
        self._check(self.lib.gral_pf_batch_last_timing(self.b, C.byref(t)))
        return {n: getattr(t, n) for n, _ in Timing._fields_}

    def profile_families(self, n):
        arr = (FamilyTime * len(FAMILIES))()
        self._check(self.lib.gral_pf_batch_profile_iterations(self.b, int(n), arr, len(FAMILIES)))
        return {f: {"ms": arr[q].ms, "launches": arr[q].launches, "cells": arr[q].cells}
                for q, f in enumerate(FAMILIES) if arr[q].launches > 0}

    def close(self):
        if getattr(self, "b", None):
            self.lib.gral_pf_batch_destroy(self.b)
            self.b = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def zip_write(path, entry, data: bytes):
    """One-entry deflated zip archive (the container of the flow-field files)."""
    lib = load()
    rc = lib.gral_pf_zip_write(os.fsencode(path), entry.encode(), data, len(data))
    if rc != 0:
        raise RuntimeError(f"gral_pf_zip_write failed ({rc})")


def _setup_check(lib, rc, what):
    if rc != 0:
        msg = lib.gral_pf_setup_last_error()
        raise RuntimeError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")


def subdomains(cutk, nii, njj, dxk, dyk, factor=15, advdom=None):
    """Prognostic sub-domains around the buildings (ReadGeometryAndDomains.cs:321-349), on the GPU."""
    lib = load()
    cut = _f32(cutk)
    adv = np.zeros((nii + 2, njj + 2), np.uint8) if advdom is None else np.array(advdom, dtype=np.uint8, order="C")
    _setup_check(lib, lib.gral_pf_subdomains(int(nii), int(njj), float(dxk), float(dyk), int(factor), _ptr(cut),
                                             _ptr(adv, U8P)), "gral_pf_subdomains")
    return adv


def reference_point(advdom):
    """Spiral search of the reference point (MicroscaleTerrainSearchRefPoint.cs:27-103)."""
    lib = load()
    adv = np.ascontiguousarray(advdom, dtype=np.uint8)
    ri, rj = C.c_int32(0), C.c_int32(0)
    rc = lib.gral_pf_reference_point(adv.shape[0] - 2, adv.shape[1] - 2, _ptr(adv, U8P), C.byref(ri), C.byref(rj))
    if rc not in (0, -4):
        _setup_check(lib, rc, "gral_pf_reference_point")
    return (ri.value, rj.value), rc == 0


def flat_init(nii, njj, nkk, vertcellsff, dxk, dyk, dzk, hokart, cutk, met, uk=None, vk=None, wk=None, advdom=None,
              kkart=None, ahk=None, bui_height=None):
    """MicroscaleFlat.Calculate() up to the flow-field call (MicroscaleFlat.cs:29-345), on the GPU.
    met: dict(n_profile, meas_height, obs_u, obs_v, ob_length, ol_column=None, near_wall=0).
    Returns dict(uk, vk, wk, kkart, ahk, bui_height, kadvmax)."""
    lib = load()
    shp3 = (nii + 2, njj + 2, nkk + 2)
    shp2 = (nii + 2, njj + 2)
    out = {"uk": np.zeros(shp3, np.float32) if uk is None else np.array(uk, dtype=np.float32, order="C"),
           "vk": np.zeros(shp3, np.float32) if vk is None else np.array(vk, dtype=np.float32, order="C"),
           "wk": np.zeros(shp3, np.float32) if wk is None else np.array(wk, dtype=np.float32, order="C"),
           "kkart": np.zeros(shp2, np.int16) if kkart is None else np.array(kkart, dtype=np.int16, order="C"),
           "ahk": np.zeros(shp2, np.float32) if ahk is None else np.array(ahk, dtype=np.float32, order="C"),
           "bui_height": np.zeros(shp2, np.float32) if bui_height is None else
           np.array(bui_height, dtype=np.float32, order="C")}
    keep = {"dzk": _f32(dzk), "hok": _f32(hokart), "cut": _f32(cutk), "H": _f32(met["meas_height"]),
            "U": _f32(met["obs_u"]), "V": _f32(met["obs_v"]),
            "ol": None if met.get("ol_column") is None else _f32(met["ol_column"]),
            "adv": None if advdom is None else np.ascontiguousarray(advdom, dtype=np.uint8)}
    g = Grid(nii, njj, nkk, 2, vertcellsff, dxk, dyk, _ptr(keep["dzk"]), _ptr(keep["hok"]))
    m = FlatMet(int(met["n_profile"]), _ptr(keep["H"]), _ptr(keep["U"]), _ptr(keep["V"]), float(met["ob_length"]),
                _ptr(keep["ol"]), int(met.get("near_wall", 0)))
    kad = C.c_int32(0)
    _setup_check(lib, lib.gral_pf_flat_init(C.byref(g), C.byref(m), _ptr(keep["cut"]), _ptr(keep["adv"], U8P),
                                            _ptr(out["uk"]), _ptr(out["vk"]), _ptr(out["wk"]),
                                            _ptr(out["kkart"], I16P), _ptr(out["ahk"]), _ptr(out["bui_height"]),
                                            C.byref(kad)), "gral_pf_flat_init")
    out["kadvmax"] = kad.value
    return out


def terrain_init(nii, njj, nkk, vertcellsff, dxk, dyk, dzk, hokart, cutk, advdom, gramm, ahmin, uk, vk, wk, kkart,
                 ahk, bui_height, near_wall=0):
    """MicroscaleTerrain.Calculate() from a GRAMM field (MicroscaleTerrain.cs:57-316), on the GPU.
    gramm: dict(nx, ny, nz, ddx, ddy, delta_i, delta_j, ahe, uwin, vwin, wwin, zsp, ah, zko1, zko2).
    Returns dict(uk, vk, wk, kkart, ahk, bui_height, ahmin, kadvmax, wind10)."""
    lib = load()
    out = {"uk": np.array(uk, dtype=np.float32, order="C"), "vk": np.array(vk, dtype=np.float32, order="C"),
           "wk": np.array(wk, dtype=np.float32, order="C"), "kkart": np.array(kkart, dtype=np.int16, order="C"),
           "ahk": np.array(ahk, dtype=np.float32, order="C"),
           "bui_height": np.array(bui_height, dtype=np.float32, order="C")}
    keep = {n: _f32(gramm[n]) for n in ("ahe", "uwin", "vwin", "wwin", "zsp", "ah")}
    keep.update(dzk=_f32(dzk), hok=_f32(hokart), cut=_f32(cutk), adv=np.ascontiguousarray(advdom, dtype=np.uint8))
    g = Grid(nii, njj, nkk, 2, vertcellsff, dxk, dyk, _ptr(keep["dzk"]), _ptr(keep["hok"]))
    gr = Gramm(int(gramm["nx"]), int(gramm["ny"]), int(gramm["nz"]), float(gramm["ddx"]), float(gramm["ddy"]),
               int(gramm["delta_i"]), int(gramm["delta_j"]), _ptr(keep["ahe"]), _ptr(keep["uwin"]),
               _ptr(keep["vwin"]), _ptr(keep["wwin"]), _ptr(keep["zsp"]), _ptr(keep["ah"]), float(gramm["zko1"]),
               float(gramm["zko2"]))
    am = C.c_float(float(ahmin))
    kad = C.c_int32(0)
    w10 = (C.c_double * 3)()
    _setup_check(lib, lib.gral_pf_terrain_init(C.byref(g), C.byref(gr), _ptr(keep["cut"]), _ptr(keep["adv"], U8P),
                                               int(near_wall), _ptr(out["uk"]), _ptr(out["vk"]), _ptr(out["wk"]),
                                               _ptr(out["kkart"], I16P), _ptr(out["ahk"]), _ptr(out["bui_height"]),
                                               C.byref(am), C.byref(kad), w10), "gral_pf_terrain_init")
    out.update(ahmin=np.float32(am.value), kadvmax=kad.value, wind10=(w10[0], w10[1], int(w10[2])))
    return out


def terrain_init_topo(nii, njj, nkk, vertcellsff, dxk, dyk, dzk, hokart, cutk, advdom, gramm, ahmin, ahk_ori,
                      ahk_ori_min, uk, vk, wk, kkart, ahk, bui_height, near_wall=0):
    """The terrain set-up with the original GRAL topography (MicroscaleTerrain.cs:318-598), on the GPU.
    gramm as for terrain_init (zsp, ah, zko not read); ahk_ori [nii+2][njj+2], ahk_ori_min its minimum
    as GRAL reads it.  Returns dict(uk, vk, wk, kkart, ahk, bui_height, ahmin, kadvmax, wind10)."""
    lib = load()
    out = {"uk": np.array(uk, dtype=np.float32, order="C"), "vk": np.array(vk, dtype=np.float32, order="C"),
           "wk": np.array(wk, dtype=np.float32, order="C"), "kkart": np.array(kkart, dtype=np.int16, order="C"),
           "ahk": np.array(ahk, dtype=np.float32, order="C"),
           "bui_height": np.array(bui_height, dtype=np.float32, order="C")}
    keep = {n: _f32(gramm[n]) for n in ("ahe", "uwin", "vwin", "wwin")}
    keep.update(dzk=_f32(dzk), hok=_f32(hokart), cut=_f32(cutk), adv=np.ascontiguousarray(advdom, dtype=np.uint8),
                ori=_f32(ahk_ori))
    g = Grid(nii, njj, nkk, 2, vertcellsff, dxk, dyk, _ptr(keep["dzk"]), _ptr(keep["hok"]))
    gr = Gramm(int(gramm["nx"]), int(gramm["ny"]), int(gramm["nz"]), float(gramm["ddx"]), float(gramm["ddy"]),
               int(gramm["delta_i"]), int(gramm["delta_j"]), _ptr(keep["ahe"]), _ptr(keep["uwin"]),
               _ptr(keep["vwin"]), _ptr(keep["wwin"]), None, None, 0.0, 0.0)
    am = C.c_float(float(ahmin))
    kad = C.c_int32(0)
    w10 = (C.c_double * 3)()
    _setup_check(lib, lib.gral_pf_terrain_init_topo(C.byref(g), C.byref(gr), _ptr(keep["ori"]), float(ahk_ori_min),
                                                    _ptr(keep["cut"]), _ptr(keep["adv"], U8P), int(near_wall),
                                                    _ptr(out["uk"]), _ptr(out["vk"]), _ptr(out["wk"]),
                                                    _ptr(out["kkart"], I16P), _ptr(out["ahk"]), _ptr(out["bui_height"]),
                                                    C.byref(am), C.byref(kad), w10), "gral_pf_terrain_init_topo")
    out.update(ahmin=np.float32(am.value), kadvmax=kad.value, wind10=(w10[0], w10[1], int(w10[2])))
    return out


def _stats(st):
    return {n: getattr(st, n) for n, _ in Stats._fields_}


def run(case, iterations=None):
    """Convenience: solve ``case`` on the GPU; returns dict of fields + 'stats'."""
    with PrognosticFlowfield(case) as pf:
        stats = pf.solve(iterations)
        out = pf.download()
    out["stats"] = stats
    return out


def run_batch(cases, iterations):
    """Convenience: advance ``cases`` as one lock-step batch for ``iterations``
    SIMPLE iterations (fixed count); returns one dict of fields + 'stats' per case."""
    solvers = [PrognosticFlowfield(c) for c in cases]
    try:
        for s in solvers:
            s.prepare(iterations)
        with Batch(solvers) as b:
            b.iterate(iterations)
        out = []
        for s in solvers:
            st = s.finish()
            f = s.download()
            f["stats"] = st
            out.append(f)
    finally:
        for s in solvers:
            s.close()
    return out


def solve_stream(cases, capacity, on_done=None, ahead=None, workers=3, on_step=None):
    """Solve a stream of situations with one batch fed while it runs: at most
    ``capacity`` situations are in the batch at a time (active or waiting
    for admission); up to ``ahead`` more (default capacity // 2) are set up
    and prepared by ``workers`` host threads while the GPU iterates, so that
    retired situations are replaced at the next admission point.  The first
    step waits until the batch is full.  ``cases`` yields Case objects or
    callables returning one, in the order they should be solved (longest
    first keeps the situations of one batch alike in size, so that no launch
    waits long for one large situation).  ``on_done(case, solver, stats)`` is
    called for each finished situation before its solver is closed (download
    the fields there); ``on_step(timing)`` after each step gets the step's
    ``Batch.timing()``.  The return value is the list of stats in completion
    order."""
    import collections
    import concurrent.futures as cf

    def build(item):
        case = item() if callable(item) else item
        sv = PrognosticFlowfield(case)
        sv.prepare()
        return sv

    if ahead is None:
        ahead = max(1, capacity // 2)
    src = iter(cases)
    more = True
    futs = collections.deque()
    out = []
    first = True
    with cf.ThreadPoolExecutor(max_workers=max(1, workers)) as ex, Batch() as batch:
        while True:
            active, queued = batch.size()
            while more and active + queued + len(futs) < capacity + ahead:
                try:
                    futs.append(ex.submit(build, next(src)))
                except StopIteration:
                    more = False
            # admit prepared situations while room is left: at the start until
            # the batch is full, later those that are ready (one when idle)
            while futs and active + queued < capacity and (futs[0].done() or first or active + queued == 0):
                batch.add(futs.popleft().result())
                queued += 1
            first = False
            if active + queued == 0:
                if not futs and not more:
                    break
                continue
            finished = batch.step()
            if on_step is not None:
                on_step(batch.timing())
            for sv, st in finished:
                if on_done is not None:
                    on_done(sv.case, sv, st)
                out.append(st)
                sv.close()
    return out
