"""The C ABI library loads and exports every symbol include/tan_mi355x.h
declares; host-side constants match the header.  No compute calls (CPU)."""
import ctypes
import os
import re

import tan_amd
from tan_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _header_functions():
    txt = open(os.path.join(ROOT, "include", "tan_mi355x.h")).read()
    return sorted(set(re.findall(r"\b(tanx_[a-z_0-9]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    names = _header_functions()
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(_lib.exported_symbols())


def test_version_string():
    lib = _lib.load()
    assert b"gfx950" in lib.tanx_version()


def test_header_enums_match_python_mirror():
    txt = open(os.path.join(ROOT, "include", "tan_mi355x.h")).read()
    vals = dict((k, int(v, 0)) for k, v in re.findall(r"(TANX_[A-Z_0-9]+)\s*=\s*(0x[0-9A-Fa-f]+|\d+)", txt))
    assert vals["TANX_WRONG_STATE"] == tan_amd.AMF_WRONG_STATE == 15
    assert vals["TANX_NOT_FOUND"] == tan_amd.AMF_NOT_FOUND == 11
    assert vals["TANX_TAN_UNSUPPORTED_VERSION"] == tan_amd.AMF_TAN_UNSUPPORTED_VERSION == 43
    assert vals["TANX_METHOD_FHT_UNIFORM_HEAD_TAIL"] == tan_amd.TAN_CONVOLUTION_METHOD_FHT_UNIFORM_HEAD_TAIL
    assert vals["TANX_METHOD_USE_PROCESS_FINALIZE"] == 0x8000
    assert vals["TANX_FFT_C2R_PLANAR_BACKWARD"] == tan_amd.TAN_FFT_C2R_PLANAR_TRANSFORM_DIRECTION_BACKWARD


def test_null_handles_rejected_without_device():
    lib = _lib.load()
    assert lib.tanx_conv_process(None, None, None, 0, None, None) == tan_amd.AMF_NOT_INITIALIZED
    assert lib.tanx_ctx_destroy(None) == tan_amd.AMF_INVALID_POINTER
    assert lib.tanx_conv_get_info(None, None) == tan_amd.AMF_INVALID_POINTER
