"""The C ABIs load and export every symbol their headers declare (no GPU
calls)."""
import ctypes
import os
import re

from conftest import ROOT


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    return sorted(set(re.findall(r"^[A-Za-z_][\w \*]*?\b(ap26h?_\w+)\(", text, re.M)))


def test_gpu_library_exports_header():
    from ap26_amd import gpu
    names = declared("ap26_gpu.h")
    assert "ap26_search_k" in names and "ap26_prp_batch" in names
    lib = ctypes.CDLL(gpu.LIB_PATH)
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(gpu.SIGNATURES), set(names) ^ set(gpu.SIGNATURES)


def test_host_library_exports_header():
    from ap26_amd import host
    names = declared("ap26_host.h")
    lib = ctypes.CDLL(host.LIB_PATH)
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(host.SIGNATURES)


def test_version_and_strerror_without_gpu():
    from ap26_amd import gpu
    lib = gpu.load()
    assert b"gfx950" in lib.ap26_version()
    assert lib.ap26_strerror(-4) == b"argument out of range"
