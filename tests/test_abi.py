"""CPU checks of the C-ABI library: it loads, exports every entry point that
include/oibvh.h declares, the ctypes table covers the header, and calls that
need no GPU behave (no compute is launched here)."""
import ctypes as C
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HDR = os.path.join(ROOT, "include", "oibvh.h")


def header_functions():
    txt = open(HDR).read()
    return sorted(set(re.findall(r"OIBVH_API\s+[^;(]*?\b(oibvh_\w+)\s*\(", txt)))


@pytest.fixture(scope="module")
def lib():
    from oibvh_amd import _lib

    if not os.path.exists(_lib.LIB_PATH):
        from oibvh_amd import build

        build.build()
    return _lib.load()


def test_header_declares_api():
    fns = header_functions()
    assert len(fns) >= 25
    assert "oibvh_tree_build" in fns and "oibvh_scene_detect" in fns


def test_library_exports_every_declared_symbol(lib):
    from oibvh_amd import _lib

    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r"\bT\s+(oibvh_\w+)", out))
    missing = [f for f in header_functions() if f not in exported]
    assert not missing, missing
    for f in header_functions():
        assert hasattr(lib, f)


def test_ctypes_table_matches_header():
    from oibvh_amd import _lib

    assert sorted(n for n, _, _ in _lib.API) == header_functions()


def test_no_internal_symbols_leak(lib):
    from oibvh_amd import _lib

    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True, text=True).stdout
    assert "oi_launch_" not in out


def test_null_handles_are_rejected(lib):
    assert lib.oibvh_tree_build(None) == 1
    assert lib.oibvh_scene_detect(None, 0, 1, 0) == 1
    v = C.c_uint32()
    assert lib.oibvh_tree_get_depth(None, C.byref(v)) == 1
    assert lib.oibvh_ctx_create(0, None, None) == 1


def test_facade_header_compiles():
    """include/oibvh.hpp is plain C++17 over the C header."""
    src = '#include "oibvh.hpp"\nint main(){ return 0; }\n'
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-I", os.path.join(ROOT, "include"), "-x", "c++", "-"],
                       input=src, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def header_struct_fields(name):
    """Field names of `typedef struct { ... } name;` in include/oibvh.h, in
    order (comments dropped)."""
    txt = re.sub(r"/\*.*?\*/", "", open(HDR).read(), flags=re.S)
    body = re.search(r"typedef struct \{([^}]*)\}\s*" + name + r"\s*;", txt).group(1)
    return re.findall(r"\b(\w+)\s*;", body)


@pytest.mark.parametrize("cname,pyname", [("oibvh_tuning", "Tuning"), ("oibvh_detect_stats", "DetectStats"),
                                          ("oibvh_scene_timing", "SceneTiming")])
def test_ctypes_structs_match_header(cname, pyname):
    """The ctypes mirrors of the header's structs have the same fields in the
    same order (the size follows: every field is a fixed-width type)."""
    from oibvh_amd import _lib

    py = getattr(_lib, pyname)
    assert [f for f, _ in py._fields_] == header_struct_fields(cname)
    src = '#include "oibvh.h"\n#include <stdio.h>\nint main(){ printf("%zu", sizeof(' + cname + ')); return 0; }\n'
    exe = os.path.join(ROOT, "_build", "abi_sizeof_" + cname)
    os.makedirs(os.path.dirname(exe), exist_ok=True)
    r = subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), "-x", "c", "-", "-o", exe], input=src,
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    assert int(subprocess.run([exe], capture_output=True, text=True).stdout) == C.sizeof(py)
