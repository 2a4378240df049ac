"""The C-ABI library (CPU checks, no GPU calls that compute)."""
import ctypes
import os
import re

import numpy as np
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


def test_other_headers_exported():
    """every entry point of the radiation and dynamics headers is exported"""
    from exoplasim_amd import _lib

    lib = _lib.lib()
    for h in ("exo_rad.h", "exo_dyn.h", "exo_mp.h", "exo_out.h"):
        syms = declared_symbols(os.path.join(ROOT, "include", h))
        assert len(syms) >= 1, h
        missing = [s for s in syms if not hasattr(lib, s)]
        assert not missing, (h, missing)


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
    # ensemble schedule: no ensemble, no output
    v = ctypes.c_int()
    assert lib.exo_sht_ens_get_schedule(None, ctypes.byref(v)) == -1
    assert lib.exo_sht_ens_set_schedule(None, 0) == -1


def test_no_fallback_without_library(monkeypatch, tmp_path):
    from exoplasim_amd import _lib

    monkeypatch.setattr(_lib, "_lib", None)
    monkeypatch.setattr(_lib, "LIB_PATH", str(tmp_path / "missing.so"))
    with pytest.raises(_lib.ExoError):
        _lib.lib()


def _fake_plan(ntru=42):
    """A Plan object with the dimensions of ntru and no device state, for
    the checks that run before any address reaches the library."""
    from exoplasim_amd import sht

    p = sht.Plan.__new__(sht.Plan)
    p._h = None
    p.ntru = ntru
    p.nlat, p.nlon, p.ncsp = sht.dims(ntru)
    p.nrsp = 2 * p.ncsp
    return p


def test_batched_buffer_checks():
    # wrong element type, shape, member count or field range raise ExoError
    # before a device address is taken; right buffers on the host reach
    # the device check
    import torch
    from exoplasim_amd import sht

    p = _fake_plan(42)
    nlev, nmem = 20, 3
    nfl, ngp = 3 * nlev + 1, 5 * nlev + 2
    sp32 = torch.zeros((nmem, nfl, p.nrsp), dtype=torch.float32)
    gp32 = torch.zeros((nmem, ngp, p.nlat, p.nlon), dtype=torch.float32)
    sp64 = sp32.double()

    def raises(fn, text):
        with pytest.raises(sht.ExoError, match=text):
            fn()

    # float32 buffers of a split-BF16 round trip handed to the FP64 mode
    raises(lambda: p.roundtrip("f64", sp32, gp32, sp32.clone(), nlev), "float64")
    raises(lambda: p.roundtrip("bf16x3", sp64, gp32, sp32, nlev), "float32")
    # buffers of another level count or truncation
    raises(lambda: p.roundtrip("bf16x3", sp32, gp32, sp32.clone(), nlev + 1), "outside")
    raises(lambda: p.roundtrip("bf16x3", sp32[:, :, :100].contiguous(), gp32, sp32, nlev),
           "shape")
    raises(lambda: p.roundtrip("bf16x3", sp32, gp32[:2], sp32, nlev), "members")
    raises(lambda: p.inverse_all("bf16x3", sp32, gp32, 0), "nlev")
    # field ranges past the buffer and member counts above it
    raises(lambda: p.inverse("bf16x3", sp32, gp32, 10, src0=nfl - 5), "outside")
    raises(lambda: p.inverse("bf16x3", sp32, gp32, 10, dst0=ngp - 1), "outside")
    raises(lambda: p.inverse("bf16x3", sp32, gp32, 1, nmem=4), "members")
    raises(lambda: p.inverse("bf16x3", sp32, gp32, 1, table="qc"), "table")
    raises(lambda: p.inverse_uv("f32", sp32, gp32, nlev, 0, nfl - 3, 0, nlev), "outside")
    raises(lambda: p.inverse_uv("f32", sp32, gp32, nlev, 0, nlev, ngp - 5, 0), "outside")
    raises(lambda: p.direct("f32", gp32, sp32, nfl, src0=0, dst0=1), "outside")
    raises(lambda: p.direct("f32mfma", gp32, sp32, -1), "outside")
    # consistent buffers get as far as the device check
    raises(lambda: p.roundtrip("bf16x3", sp32, gp32, sp32.clone(), nlev), "device")
    raises(lambda: p.roundtrip("f64", sp64, gp32.double(), sp64.clone(), nlev), "device")
    raises(lambda: p.inverse("f32", sp32, gp32, nfl), "device")
    raises(lambda: p.direct("f32", gp32, sp32, nfl), "device")


def test_dropin_size_checks():
    from exoplasim_amd import sht

    saved = dict(sht._state)
    try:
        sht._state.update(ntru=21, nlat=32, nlev=10)
        short = np.zeros((9, 2 * 253))
        with pytest.raises(sht.ExoError, match="NLEV"):
            sht.dv2uv(short, np.zeros((10, 2 * 253)))
    finally:
        sht._state.update(saved)


def test_new_entry_points_argument_errors():
    """the multi-process, output, dynamics and guarded-emulator entry points
    reject bad arguments, and accept empty work, before any device work"""
    import ctypes

    from exoplasim_amd import dynamics, multiproc, output

    mp, out, dyn = multiproc._lib(), output._lib(), dynamics._lib()
    ERR = -1
    # mpsumsc: no processes; no fields is nothing to do
    assert mp.exo_sht_sumsc(0, 0, 4, 1, None, None, None) == ERR
    assert mp.exo_sht_sumsc(0, 2, 4, 0, None, None, None) == 0
    assert mp.exo_sht_sumsc(0, 2, 4, 1, None, None, None) == ERR
    assert mp.exo_sht_sumsc(7, 2, 4, 1, None, None, None) == ERR  # mode
    # mrtrunc / mrexpand: the low truncation above the high one
    assert mp.exo_sht_mrtrunc(0, 21, 42, 1, None, 0, None, 0, None) == ERR
    assert mp.exo_sht_mrexpand(0, 42, 21, 1, None, 0, None, 0, None) == ERR
    assert mp.exo_sht_mrtrunc(0, 42, 21, 0, None, 0, None, 0, None) == 0
    # partial fc2sp without a plan
    assert mp.exo_sht_fc2sp_part(None, 0, 1, None, 0, 0, 0, 8, None, 0, 0, 506, 1, None) == ERR
    # writesp: mode and truncation checks, empty work
    assert out.exo_sht_writesp(9, 21, 1, None, 0, 1.0, None, None, 0, None) == ERR
    assert out.exo_sht_writesp(0, -1, 1, None, 0, 1.0, None, None, 0, None) == ERR
    assert out.exo_sht_writesp(0, 21, 0, None, 0, 1.0, None, None, 0, None) == 0
    assert out.exo_sht_writesp(0, 21, 1, None, 0, 1.0, None, None, 0, None) == ERR
    # dynamics: no handle, too few levels; the work-array size of gridpointa
    h = ctypes.c_void_p()
    d = (ctypes.c_double * 4)()
    assert dyn.exo_dyn_create(None, 4, d, d, d, d, d, d, d, d, 0.286, 0.6, 0.6, 1e5,
                              7.292e-5) == ERR
    assert dyn.exo_dyn_create(ctypes.byref(h), 1, d, d, d, d, d, d, d, d, 0.286, 0.6, 0.6,
                              1e5, 7.292e-5) == ERR
    assert dyn.exo_dyn_destroy(None) == ERR
    assert dyn.exo_gridpointa_fields(20) == 18 * 20 + 5
    assert dyn.exo_gridpointa_fields(1) == ERR
    assert dyn.exo_gridpointa(None, None, 0, 1, 1, None, 0, None, 0, None, 0, None) == ERR
