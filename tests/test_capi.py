"""The C-ABI library (CPU checks, no GPU calls that compute)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = [os.path.join(ROOT, "include", h) for h in ("exo_sht.h", "exo_si.h")]


def declared_symbols(path):
    text = open(path).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(exo_[a-z0-9_]+)\s*\(", text)))


def test_header_symbols_exported():
    from exoplasim_amd import _lib

    lib = _lib.lib()
    syms = sorted(set(sum((declared_symbols(h) for h in HEADERS), [])))
    assert len(syms) >= 20
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing
    # every exported entry point has a ctypes prototype in the Python mirror
    assert set(syms) == set(_lib.PROTOTYPES), set(syms) ^ set(_lib.PROTOTYPES)


def test_strerror_and_argument_errors():
    from exoplasim_amd import _lib

    lib = _lib.lib()
    assert lib.exo_strerror(0) == b"ok"
    assert b"truncation" in lib.exo_strerror(-2)
    # argument validation happens before any device work
    assert lib.exo_sht_plan_create(None, 42, None, None, None, None, 1.0) == -1
    assert lib.exo_sht_plan_destroy(None) == -1
    assert lib.exo_legini(30, 48, 10, None, None, None, None, 1.0) == -2
    assert lib.exo_legini(42, 64, 0, None, None, None, None, 1.0) == -2


def test_no_fallback_without_library(monkeypatch, tmp_path):
    from exoplasim_amd import _lib

    monkeypatch.setattr(_lib, "_lib", None)
    monkeypatch.setattr(_lib, "LIB_PATH", str(tmp_path / "missing.so"))
    with pytest.raises(_lib.ExoError):
        _lib.lib()
