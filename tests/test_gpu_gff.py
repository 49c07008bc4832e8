"""GPU parity of the flow-field file packer (GFFWrite.cs): the payload the
GPU packs from the handle's fields -- the solver's device-resident levels plus
the host levels above them -- against the oracle's payload of the same fields
after download, byte for byte, for the three writing modes and the *.tke file;
and the zip file gral_pf_gff_write leaves, read back with Python's zipfile."""
import os
import zipfile

import numpy as np
import pytest

from gral_amd import domain
from oracle import gff_oracle as G

META = dict(wind_dir=235.0, wind_vel=3.5, stab_class=5)


def _oracle(case, f, kind):
    return G.payload(kind, f["uk"], f["vk"], f["wk"], case.kkart, case.nii, case.njj, case.nkk, case.dxk,
                     META["wind_dir"], META["wind_vel"], META["stab_class"], turb=f.get("turb"))


def _check(s, case, kinds):
    f = s.download()
    for kind in kinds:
        print(f"[gff] {case.name} {case.nii}x{case.njj}x{case.nkk} model={case.turbulence_model} kind={kind}", flush=True)
        got = s.gff_payload(kind, **META)
        want = _oracle(case, f, kind)
        assert len(got) == len(want), (kind, len(got), len(want))
        a, b = np.frombuffer(got[32:], "<i2"), np.frombuffer(want[32:], "<i2")
        bad = np.flatnonzero(a != b)
        assert got[:32] == want[:32], kind
        assert bad.size == 0, (kind, bad[:8].tolist())


def _two_buildings(model, vcf=10):
    bl = [domain.Building(3, 4, 3, 5, 6.5), domain.Building(6, 7, 4, 6, 3.2)]
    return domain.build_case("two", 10, 10, bl, vertcellsff=vcf, direction=235.0, turbulence_model=model, iterations=3)


@pytest.mark.gpu
def test_uploaded_fields_before_prepare():
    from gral_amd import pf
    c = _two_buildings(1)
    with pf.PrognosticFlowfield(c) as s:
        _check(s, c, (0, 1, 2))
        with pytest.raises(RuntimeError):
            s.gff_payload(3, **META)           # no k-eps situation


@pytest.mark.gpu
@pytest.mark.parametrize("model", [1, 2])
def test_after_iterations_without_download(model):
    # the payload comes from the device fields while the host copy is stale
    from gral_amd import pf
    c = _two_buildings(model)
    with pf.PrognosticFlowfield(c) as s:
        s.prepare(4)
        s.iterate(4)
        kinds = (0, 1, 2, 3) if model == 2 else (0, 1, 2)
        pays = {k: s.gff_payload(k, **META) for k in kinds}
        s.finish()
        f = s.download()
        for k in kinds:
            assert pays[k] == _oracle(c, f, k), k


@pytest.mark.gpu
def test_after_diagnostic():
    from gral_amd import pf
    c = _two_buildings(2)
    with pf.PrognosticFlowfield(c) as s:
        s.prepare(3)
        s.iterate(3)
        _check(s, c, (0, 2, 3))
        s.diagnostic(2)
        _check(s, c, (0, 1, 2, 3))


@pytest.mark.gpu
def test_config1_kinds():
    from gral_amd import pf
    c = domain.config1(turbulence_model=2, iterations=5)
    with pf.PrognosticFlowfield(c) as s:
        s.solve(5)
        _check(s, c, (0, 1, 2, 3))


@pytest.mark.gpu
def test_write_zip_files(tmp_path):
    from gral_amd import pf
    c = _two_buildings(2)
    with pf.PrognosticFlowfield(c) as s:
        s.solve(3)
        for kind, name in ((0, "00001.gff"), (1, "00002.gff"), (2, "00003.gff"), (3, "00001.tke")):
            p = os.path.join(str(tmp_path), name)
            print(f"[gff] write kind={kind} {name}", flush=True)
            s.write_gff(p, kind, **META)
            with zipfile.ZipFile(p) as z:
                assert z.testzip() is None
                assert z.namelist() == [name]
                assert z.read(name) == s.gff_payload(kind, **META)
        rd = G.read_payload(zipfile.ZipFile(os.path.join(str(tmp_path), "00003.gff")).read("00003.gff"), c.kkart)
        assert (rd["nii"], rd["njj"], rd["nkk"], rd["flag"]) == (c.nii, c.njj, c.nkk, -3)
