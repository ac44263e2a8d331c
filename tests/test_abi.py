"""The C ABI library loads and exports every entry point include/rbw.h declares;
without a device, creation fails with a status code and a message (no
exception crosses the ABI, no CPU fallback is taken)."""
import ctypes as C
import os
import re

import pytest

from rainbow_amd import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = []
    for h in sorted(os.listdir(os.path.join(ROOT, "include"))):
        if h.endswith(".h"):
            src = open(os.path.join(ROOT, "include", h)).read()
            names += re.findall(r"^\s*(?:rbw_status|void|const char\s*\*)\s+(rbw_\w+)\s*\(", src, re.M)
    return names


def test_header_declares_the_abi():
    names = declared_functions()
    assert {"rbw_create", "rbw_destroy", "rbw_last_error", "rbw_upload_scene", "rbw_render",
            "rbw_trace_rays"} <= set(names)


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(abi.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_python_signatures_cover_the_header():
    assert set(declared_functions()) <= set(abi.RBW_SIGNATURES)


def test_struct_sizes_match_header_layout():
    # plain 4-byte fields, no padding surprises between C and ctypes
    assert C.sizeof(abi.rbw_material) == 52
    assert C.sizeof(abi.rbw_entity) == 32
    assert C.sizeof(abi.rbw_hit) == 20
    assert C.sizeof(abi.rbw_camera_desc) == 16 * 4 * 2 + 4 * 2 + 4 * 4 + 4 * 2


def _has_gpu():
    lib = C.CDLL("libamdhip64.so")
    n = C.c_int(0)
    return lib.hipGetDeviceCount(C.byref(n)) == 0 and n.value > 0


def test_create_without_device_reports_error():
    if _has_gpu():
        pytest.skip("a GPU is present")
    from rainbow_amd.tracer import RbwError, Tracer
    with pytest.raises(RbwError) as e:
        Tracer(0)
    assert e.value.code != abi.RBW_OK
    assert str(e.value)


def test_null_arguments_are_rejected():
    lib = abi.load_library()
    assert lib.rbw_create(None, None) == abi.RBW_ERR_INVALID_ARGUMENT
    assert lib.rbw_last_error()
    assert lib.rbw_upload_scene(None, None) != 0
    assert lib.rbw_render(None, None, None, None, None) != 0
