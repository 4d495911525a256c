"""The C ABI library loads and exports what include/v3dm.h declares (CPU only: no
compute call is made without a GPU)."""
import ctypes as C
import os
import re
import subprocess

import pytest

import v3dmatch
from v3dmatch import _native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(v3dm_[a-z0-9_]+)\s*\(", text)))


HEADERS = ["v3dm.h", "v3dm_ransac.h"]


@pytest.mark.parametrize("header,at_least", [("v3dm.h", 20), ("v3dm_ransac.h", 5)])
def test_every_declared_symbol_is_exported(header, at_least):
    lib = C.CDLL(_native.lib_path())
    names = declared(header)
    assert len(names) >= at_least
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


@pytest.mark.parametrize("header", HEADERS)
def test_python_binding_covers_the_header(header):
    assert set(declared(header)) <= set(_native.SIGNATURES)


def test_create_without_gpu_reports_an_error():
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("a GPU is present")
    except ImportError:
        pass
    lib = v3dmatch.load_library()
    h = C.c_void_p()
    rc = lib.v3dm_create(C.byref(h), 0)
    assert rc != 0
    assert lib.v3dm_last_error(None)
    with pytest.raises(v3dmatch.V3DMError):
        v3dmatch.SIFT_Matching(0)


def test_null_handle_calls_fail_cleanly():
    lib = v3dmatch.load_library()
    assert lib.v3dm_allocate(None, 32, 32, 0) == 1
    assert lib.v3dm_set_left(None, 32, None) == 1
    assert lib.v3dm_find_putative_matches_async(None, C.c_float(0.9), C.c_float(0.8)) == 1
    lib.v3dm_destroy(None)


@pytest.mark.parametrize("header", HEADERS)
@pytest.mark.parametrize("compiler,std,src", [("gcc", "-std=c99", "x.c"), ("g++", "-std=c++11", "x.cpp")])
def test_header_compiles_standalone(tmp_path, header, compiler, std, src):
    f = tmp_path / src
    f.write_text(f'#include "{header}"\nint main(void) {{ return 0; }}\n')
    r = subprocess.run([compiler, std, "-Wall", "-Werror", "-fsyntax-only", "-I", os.path.join(ROOT, "include"),
                        str(f)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_dropin_header_builds_and_links(tmp_path):
    import sys
    sys.path.insert(0, os.path.join(ROOT, "tests", "cpp"))
    from build_dropin import build
    exe = build(str(tmp_path))
    assert os.path.exists(exe)
