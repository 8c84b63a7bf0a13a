"""The C-ABI library loads and exports every entry point the header declares
(CPU-only: no compute call is made here)."""
import ctypes as ct
import os
import re

import pytest

import forcesmith_amd as fa

HDR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "forcesmith_amd.h")


def declared():
    txt = open(HDR).read()
    return sorted(set(re.findall(r"\b(fs_[a-z_0-9]+)\s*\(", txt)))


def test_header_and_mirror_agree():
    assert declared() == sorted(fa.EXPORTS)


def test_library_exports_every_declared_symbol():
    lib = fa.load_library()
    for name in declared():
        assert hasattr(lib, name), name
    assert lib.fs_abi_version() == 1


def test_null_context_is_an_error_not_a_crash():
    lib = fa.load_library()
    assert lib.fs_prepare(None) == -1
    assert b"null context" in lib.fs_last_error(None)


def test_missing_library_fails_loudly(tmp_path):
    saved = fa._lib
    try:
        fa._lib = None
        with pytest.raises(FileNotFoundError):
            fa.load_library(str(tmp_path / "nope.so"))
    finally:
        fa._lib = saved
