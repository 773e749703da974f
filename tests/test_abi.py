"""The C-ABI libraries load and export every symbol their headers declare
(CPU-only: nothing is called that needs a GPU)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "gpuattract_2.0_amd", "lib")


def declared(header, prefix):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(" + prefix + r"\w+)\s*\(", text)))


@pytest.mark.parametrize("header,lib,prefix", [
    ("attract_hip.h", "libattract_hip.so", "att_"),
    ("attract_gen.h", "libattract_gen.so", "attg_"),
    ("attract_serve.h", "libattract_service.so", "att_server_"),
    ("attract_io.h", "libattract_service.so", "att_gridfile_"),
    ("attract_formats.h", "libattract_service.so", "att_"),
])
def test_exports(header, lib, prefix):
    names = declared(header, prefix)
    assert len(names) >= 4
    L = ctypes.CDLL(os.path.join(LIB, lib))
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_build_info():
    L = ctypes.CDLL(os.path.join(LIB, "libattract_hip.so"))
    L.att_build_info.restype = ctypes.c_char_p
    assert b"gfx950" in L.att_build_info()


def test_no_cpu_fallback_when_library_missing(monkeypatch):
    from gpuattract_amd import hip
    monkeypatch.setattr(hip, "LIBPATH", "/nonexistent/libattract_hip.so")
    monkeypatch.setattr(hip, "_lib", None)
    with pytest.raises(RuntimeError, match="no CPU fallback"):
        hip.lib()
