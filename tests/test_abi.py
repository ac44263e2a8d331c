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


STRUCTS = ["rbw_material", "rbw_entity", "rbw_medium", "rbw_scene_desc", "rbw_camera_desc", "rbw_guiding_opts",
           "rbw_render_params", "rbw_stats", "rbw_device_opts", "rbw_hit", "rbw_job", "rbw_batch_opts",
           "rbw_job_stats", "rbw_batch_stats"]


def test_struct_layouts_match_the_c_compiler(tmp_path):
    """sizeof and the offset of every field as the C compiler lays out include/rbw.h,
    against the ctypes mirror"""
    import subprocess
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "rbw.h"', "int main(void) {"]
    for name in STRUCTS:
        lines.append(f'    printf("{name} %zu\\n", sizeof({name}));')
        for fname, _ in getattr(abi, name)._fields_:
            lines.append(f'    printf("{name}.{fname} %zu\\n", offsetof({name}, {fname}));')
    lines += ["    return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), "-o", str(exe), str(src)], check=True)
    got = dict(l.split() for l in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.splitlines())
    for name in STRUCTS:
        t = getattr(abi, name)
        assert int(got[name]) == C.sizeof(t), name
        for fname, _ in t._fields_:
            assert int(got[f"{name}.{fname}"]) == getattr(t, fname).offset, (name, fname)


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
