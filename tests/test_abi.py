"""The C-ABI libraries load and export every entry point their headers
declare (no compute calls: this runs without a GPU)."""
import ctypes
import os
import re

import pytest

from uvlm_amd import _native as N

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(uvlm_\w+)\s*\(", text)))


@pytest.mark.parametrize("header,lib", [("uvlm_gpu.h", N.GPU_LIB), ("uvlm_host.h", N.HOST_LIB)])
def test_library_exports_header(header, lib):
    names = declared(header)
    assert names, header
    L = ctypes.CDLL(lib)
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_binding_lists_match_headers():
    assert sorted(N.GPU_SYMBOLS) == declared("uvlm_gpu.h")
    assert sorted(N.HOST_SYMBOLS) == declared("uvlm_host.h")


def test_missing_library_fails_loudly(tmp_path, monkeypatch):
    monkeypatch.setattr(N, "_gpu", None)
    monkeypatch.setattr(N, "GPU_LIB", str(tmp_path / "nope.so"))
    with pytest.raises(N.MissingLibrary):
        N.gpu_lib()


def test_gpu_library_is_gfx950():
    """The shared object carries a gfx950 code object."""
    data = open(N.GPU_LIB, "rb").read()
    assert b"gfx950" in data
