"""C-ABI behaviour that needs no GPU: argument checks and error codes."""
import ctypes as C
import os

import numpy as np
import pytest

from gral_amd import domain, pf


def _lib():
    if not os.path.exists(pf.LIB_PATH):
        pytest.skip("libgral_pf.so not built")
    return pf.load()


def test_null_and_bad_grid_rejected():
    lib = _lib()
    h = C.c_void_p()
    assert lib.gral_pf_create(None, C.byref(h)) == -1
    dz = np.ones(10, dtype=np.float32)
    g = pf.Grid(3, 3, 8, 4, 2, 2.0, 2.0, dz.ctypes.data_as(pf.FP), dz.ctypes.data_as(pf.FP))
    assert lib.gral_pf_create(C.byref(g), C.byref(h)) == -1      # nii < 5
    g = pf.Grid(10, 10, 8, 9, 2, 2.0, 2.0, dz.ctypes.data_as(pf.FP), dz.ctypes.data_as(pf.FP))
    assert lib.gral_pf_create(C.byref(g), C.byref(h)) == -1      # kadvmax > nkk


def test_no_device_is_reported_not_crashed():
    # on a host without a GPU a valid grid fails with GRAL_PF_EDEVICE (-2)
    try:
        import torch
        if torch.cuda.device_count() > 0:
            pytest.skip("a GPU is present")
    except Exception:
        pass
    lib = _lib()
    c = domain.tiny_case(n=9)
    g = pf.Grid(c.nii, c.njj, c.nkk, c.kadvmax, c.vertcellsff, c.dxk, c.dyk,
                c.dzk.ctypes.data_as(pf.FP), c.hokart.ctypes.data_as(pf.FP))
    h = C.c_void_p()
    assert lib.gral_pf_create(C.byref(g), C.byref(h)) == -2
