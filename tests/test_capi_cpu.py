"""C-ABI checks that need no GPU: the library loads, exports every function the header
declares, and its pure-host lookups follow the reference's selection tables
(ionicModel.C:76-94, ionicSelector.C:60-66, BuenoOrovio_2008.H:24-30)."""
import ctypes as C

import numpy as np
import pytest

from cardiacfoam_amd import _capi


def test_library_exports_header_symbols():
    assert _capi.missing_symbols() == []
    assert b"gfx950" in _capi.lib().cg_version()


def test_model_and_tissue_lookup():
    L = _capi.lib()
    assert [L.cg_model_id(k) for k in (b"BuenoOrovio", b"TNNP", b"ORd", b"AlievPanfilov")] == [0, 1, 2, 3]
    assert L.cg_model_id(b"Courtemanche") == 5
    assert L.cg_model_id(b"PerisYague_2022") == 6
    assert L.cg_model_id(b"Stewart") == 7
    assert L.cg_model_id(b"Grandi") == 8
    assert L.cg_model_id(b"ToRORd_dynCl") == 9
    assert L.cg_model_id(b"Gaur") == 10
    assert L.cg_model_id(b"Fabbri") == -1
    assert [L.cg_tissue_flag(t) for t in (b"epicardialCells", b"mCells", b"endocardialCells",
                                          b"myocyte")] == [1, 2, 3, 4]
    assert L.cg_tissue_flag(b"purkinje") == -1
    assert [L.cg_model_nstates(m) for m in range(11)] == [4, 17, 41, 2, 4, 21, 22, 20, 41, 45,
                                                          29]
    assert L.cg_model_nstates(11) == -1
    buf = C.create_string_buffer(32)
    names = []
    for i in range(4):
        _capi.check(L.cg_model_state_name(0, i, buf, 32))
        names.append(buf.value)
    assert names == [b"u", b"v", b"w", b"s"]
    _capi.check(L.cg_model_state_name(2, 40, buf, 32))
    assert buf.value == b"ryr_Jrelp"
    # Courtemanche_1998.H:24-46 state names, first and last
    _capi.check(L.cg_model_state_name(5, 0, buf, 32))
    assert buf.value == b"membrane_V"
    _capi.check(L.cg_model_state_name(5, 20, buf, 32))
    assert buf.value == b"cajsr_w"
    with pytest.raises(_capi.CgError):
        _capi.check(L.cg_model_state_name(1, 17, buf, 32))


def test_mesh_ldu_matching(orc):
    L = _capi.lib()
    m = orc.OracleMesh(5, 4, 3, 1e-4, 1e-4, 1e-4)
    own = np.ascontiguousarray(m.owner, dtype=np.int32)
    nei = np.ascontiguousarray(m.neighbour, dtype=np.int32)
    _capi.check(L.cg_mesh_match_ldu(5, 4, 3, len(own), own.ctypes.data_as(C.c_void_p),
                                    nei.ctypes.data_as(C.c_void_p)))
    nei2 = nei.copy()
    nei2[[3, 4]] = nei2[[4, 3]]
    with pytest.raises(_capi.CgError):
        _capi.check(L.cg_mesh_match_ldu(5, 4, 3, len(own), own.ctypes.data_as(C.c_void_p),
                                        nei2.ctypes.data_as(C.c_void_p)))


def test_ionic_new_errors_without_device():
    import cardiacfoam_amd as cf
    with pytest.raises(cf.FatalIOError):
        cf.IonicModel.New({"ionicModel": "ToRORd_dynCl", "tissue": "myocyte"}, 10, 1e-5)
    with pytest.raises(cf.FatalIOError):
        cf.IonicModel.New({"ionicModel": "ORd", "tissue": "epicardialCells"}, 10, 1e-5)
    with pytest.raises(cf.FatalIOError):
        cf.IonicModel.New({"ionicModel": "Courtemanche", "tissue": "endocardialCells"}, 10, 1e-5)
    with pytest.raises(cf.FatalIOError):
        cf.IonicModel.New({"ionicModel": "TNNP"}, 10, 1e-5)


def test_adjoint_stale_backward_check():
    """The host side of the adjoint refuses a backward for any forward run but the
    latest (no device needed)."""
    from cardiacfoam_amd.adjoint import TissueAdjoint
    adj = TissueAdjoint.__new__(TissueAdjoint)
    adj.generation = 3
    adj.check_backward(3)
    with pytest.raises(RuntimeError, match="replaced"):
        adj.check_backward(2)
    adj.generation = -1                       # a forward run that failed
    with pytest.raises(RuntimeError):
        adj.check_backward(3)


def test_model_units_follow_the_reference():
    """cg_model_units: BuenoOrovio.C:119-122 / BuenoOrovio.H:78-81 ((Vm*1000 + 84)/85.7,
    85.7 u - 84, ms), AlievPanfilov.C:107-122 / AlievPanfilov.H:78-81 ((Vm*1000 + 80)/100,
    100 u - 80, t*1000/12.9), the mV models (Vm*1000, ms) and the manufactured model (V =
    Vm, s: monodomainFDAManufactured.C:98; Grandi.C:121-124 as the other mV models)."""
    from cardiacfoam_amd.ionic import model_units
    assert model_units("BuenoOrovio") == (1000.0, 84.0, 85.7, 1000.0)
    assert model_units("AlievPanfilov") == (1000.0, 80.0, 100.0, 1000.0 / 12.9)
    for key in ("TNNP", "ORd", "Courtemanche", "PerisYague_2022", "Stewart", "Grandi",
                "ToRORd_dynCl", "Gaur"):
        assert model_units(key) == (1000.0, 0.0, 1.0, 1000.0), key
    assert model_units("monodomainFDAManufactured") == (1.0, 0.0, 1.0, 1.0)
    with pytest.raises(_capi.CgError):
        model_units(11)
