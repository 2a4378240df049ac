"""The host-side routines of the Fortran drop-in (exoplasim_amd/fortran/
exo_transforms.f90) that do no GPU work: ql and qld (src/gaussmod.f90:46-86,
the normalised Legendre polynomial and the reciprocal of its derivative)
and fftini's length check (src/fftmod.f90:93-122).  The shim is compiled
with amdflang into a shared library and called through ctypes; ql and qld
are held to scipy's Legendre polynomials and, where oracle/_ref is built, to
the reference's own ql_ and qld_."""
import ctypes
import os
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FORT = os.path.join(ROOT, "exoplasim_amd", "fortran")
REF = os.path.join(ROOT, "oracle", "_ref", "libref_t21.so")


@pytest.fixture(scope="module")
def shim(tmp_path_factory):
    fc = os.environ.get("EXO_FC") or shutil.which("amdflang") or "/opt/rocm/bin/amdflang"
    if not os.path.exists(fc):
        pytest.skip("amdflang not found")
    lib = os.path.join(ROOT, "exoplasim_amd", "libexosht.so")
    if not os.path.exists(lib):
        pytest.skip("libexosht.so not built")
    d = tmp_path_factory.mktemp("shim")
    out = str(d / "libshim.so")
    srcs = [os.path.join(FORT, "exo_legmod_capi.f90"), os.path.join(FORT, "exo_transforms.f90"),
            os.path.join(ROOT, "tests", "fortran", "abort_stub.f90")]
    subprocess.run([fc, "-O2", "-fPIC", "-shared", f"-J{d}", *srcs, "-o", out,
                    f"-L{os.path.dirname(lib)}", "-lexosht",
                    f"-Wl,-rpath,{os.path.dirname(lib)}"], check=True, capture_output=True)
    return ctypes.CDLL(out)


def _fn(lib, name):
    f = getattr(lib, name)
    f.restype = ctypes.c_double
    f.argtypes = [ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_double)]
    return lambda k, p: f(ctypes.byref(ctypes.c_int(k)), ctypes.byref(ctypes.c_double(p)))


def test_ql_qld_closed_forms(shim):
    from scipy.special import eval_legendre

    ql, qld = _fn(shim, "ql_"), _fn(shim, "qld_")
    rng = np.random.default_rng(3)
    for k in (0, 1, 2, 5, 16, 32, 64, 96, 192):
        for p in rng.uniform(-0.999, 0.999, 8):
            want = np.sqrt((2 * k + 1) / 2.0) * eval_legendre(k, p)
            assert abs(ql(k, p) - want) <= 1e-12 * max(1.0, abs(want)), (k, p)
            if k >= 1:
                # 1 / (d ql / dp), the derivative by a central difference
                h = 1e-6
                d = (np.sqrt((2 * k + 1) / 2.0) * (eval_legendre(k, p + h) - eval_legendre(k, p - h))
                     / (2 * h))
                assert abs(qld(k, p) * d - 1.0) < 1e-5, (k, p)


@pytest.mark.skipif(not os.path.exists(REF), reason="oracle/_ref not built (needs the reference tree)")
def test_ql_qld_vs_reference(shim):
    ref = ctypes.CDLL(REF)
    ql, qld = _fn(shim, "ql_"), _fn(shim, "qld_")
    rql, rqld = _fn(ref, "ql_"), _fn(ref, "qld_")
    for k in (1, 2, 7, 32, 64, 96, 192):
        for p in np.linspace(-0.97, 0.95, 11):  # off the derivative's zero at 0
            assert abs(ql(k, p) - rql(k, p)) <= 1e-12 * max(1.0, abs(rql(k, p))), (k, p)
            assert abs(qld(k, p) - rqld(k, p)) <= 1e-9 * abs(rqld(k, p)), (k, p)
        # at Gauss nodes, where inigau's Newton steps use them
        x, _ = np.polynomial.legendre.leggauss(k)
        for p in x[: (k + 1) // 2]:
            assert abs(ql(k, p) - rql(k, p)) <= 1e-12, (k, p)
            assert abs(qld(k, p) - rqld(k, p)) <= 1e-9 * abs(rqld(k, p)), (k, p)


def test_fftini_accepts_supported_lengths(shim):
    f = shim.fftini_
    f.argtypes = [ctypes.POINTER(ctypes.c_int)]
    for n in (64, 128, 192, 384):
        f(ctypes.byref(ctypes.c_int(n)))  # returns; an unsupported n stops
