"""Pins of the post-processor wrapper oracle (oracle.PyfftOracle, restating
exoplasim/pyfft.f90's sp2gp, gp2sp, sp3fc, gp3fc, fc3gp, spvgp, gpvsp and
its physics filter) against the reference's own pyfft.f90 compiled here:

* the -fdefault-real-8 build ("r8"): every literal and the filter in double
  precision; the oracle in kind "r8" must agree with it to rounding;
* the default-real build ("r4", what the post-processor's f2py step
  compiles): the sqrt 2 literal and the filter are single precision.  Its
  FFT twiddles are single precision too (fftini, pyfft.f90:1203-1208, has
  `del` and `angle` implicitly default real), which neither the oracle nor
  the GPU path repeats: the functions that run an FFT differ from it by that
  rounding (below 3e-7 relative, the same gap as between the reference's own
  r4 and r8 builds, checked here too); sp3fc, which runs no FFT, and the
  filter agree with it to rounding.

Golden vectors (tests/golden/reference_pyfft_t21.npz, written by
tests/golden/make_golden.py from both builds) and, where oracle/_ref is
built, the live libraries.
"""
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
from make_golden import pyfft_inputs  # noqa: E402

GOLD = os.path.join(HERE, "golden", "reference_pyfft_t21.npz")
REF_DIR = os.path.join(os.path.dirname(HERE), "oracle", "_ref")
TWIDDLE_GAP = 3e-7  # single-precision FFT twiddles of the default-real build


def rel(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / np.linalg.norm(np.ravel(b)))


def _outputs(p, ntru, nlat, nlev, nfs, fft=True):
    sd, sz, gu, gv, rd, cd = pyfft_inputs(ntru, nlat, nlev)
    out = {"sp2gp": p.sp2gp(ntru, nlat, sd, nfs), "gp2sp": p.gp2sp(ntru, gu, nfs),
           "sp3fc": p.sp3fc(ntru, nlat, sd, nfs)}
    out["spvgp_u"], out["spvgp_v"] = p.spvgp(ntru, sd, sz, rd, nfs)
    out["gpvsp_d"], out["gpvsp_z"] = p.gpvsp(ntru, gu, gv, cd, nfs)
    out["gp3fc"], out["fc3gp"] = p.gp3fc(gu), p.fc3gp(gu)
    return out


def test_pyfft_oracle_vs_golden(oracle):
    from oracle import PyfftOracle

    g = np.load(GOLD)
    r4 = PyfftOracle(oracle, "r4")
    for nfs in (0, 1):
        o = _outputs(r4, 21, 32, 1, nfs)
        k = f"r4_nfs{nfs}_"
        assert rel(o["sp3fc"], g[k + "sp3fc"]) < 1e-15
        for name in ("sp2gp", "gp2sp", "spvgp_u", "spvgp_v", "gpvsp_d", "gpvsp_z"):
            e = rel(o[name], g[k + name])
            print(f"r4 nfs={nfs} {name}: {e:.1e}")
            assert e < TWIDDLE_GAP, (nfs, name, e)
    assert rel(o["gp3fc"], g["r4_gp3fc"]) < TWIDDLE_GAP
    assert rel(o["fc3gp"], g["r4_fc3gp"]) < TWIDDLE_GAP
    # the filter of the default-real build, bit for bit
    for t in (21, 42, 63, 127):
        np.testing.assert_array_equal(r4.sfilt(t, 1), g[f"r4_sfilt_t{t}"])
    # the double-precision build: the same wrappers to rounding
    r8 = PyfftOracle(oracle, "r8")
    sd, sz, gu, gv, rd, cd = pyfft_inputs(21, 32, 1)
    assert rel(r8.sp2gp(21, 32, sd, 1), g["r8_nfs1_sp2gp"]) < 1e-14
    assert rel(r8.gp2sp(21, gu, 1), g["r8_nfs1_gp2sp"]) < 1e-14


def test_pyfft_closed_forms(oracle):
    """What the wrappers mean, independent of the reference: gp2sp inverts
    sp2gp on band-limited fields (the sqrt 2 factors cancel), sp2gp of the
    (m=0, n=0) mode is the constant sqrt(0.5) sqrt 2 = 1, and the filter
    damps mode n by exp(-8 (n/NTRU)^8) with the reference's 1-based n."""
    from oracle import PyfftOracle

    p = PyfftOracle(oracle, "r8")
    sd, _, _, _, _, _ = pyfft_inputs(42, 64, 2)
    assert rel(p.gp2sp(42, p.sp2gp(42, 64, sd)), sd) < 1e-12
    one = np.zeros((1, 22 * 23))
    one[0, 0] = 1.0
    assert np.abs(p.sp2gp(21, 32, one) - 1.0).max() < 1e-14
    f = p.sfilt(21, 1)
    n1 = np.arange(1, 23)
    assert np.allclose(f, np.exp(-8.0 * (n1 / 21.0) ** 8), rtol=1e-14, atol=0)
    # a planted mistake: 0-based n in the filter must fail this check
    assert not np.allclose(np.exp(-8.0 * ((n1 - 1) / 21.0) ** 8), f, rtol=1e-6)


@pytest.mark.skipif(not os.path.exists(os.path.join(REF_DIR, "libref_pyfft_r4.so")),
                    reason="oracle/_ref not built (needs the reference tree)")
@pytest.mark.parametrize("ntru,nlat", [(21, 32), (42, 64)])
def test_pyfft_oracle_vs_live_reference(oracle, ntru, nlat):
    from oracle import PyfftOracle, RefPyfft

    for kind in ("r8", "r4"):
        ref, ora = RefPyfft(kind), PyfftOracle(oracle, kind)
        for nfs in (0, 1):
            a = _outputs(ora, ntru, nlat, 2, nfs)
            b = _outputs(ref, ntru, nlat, 2, nfs)
            for name in a:
                e = rel(a[name], b[name])
                tol = 1e-14 if (kind == "r8" or name == "sp3fc") else TWIDDLE_GAP
                assert e < tol, (kind, nfs, name, e)
        if kind == "r4":
            for t, nl in ((21, 32), (42, 64), (63, 96), (127, 192)):
                np.testing.assert_array_equal(ora.sfilt(t, 1), ref.sfilt(t, nl, 1))
    # the reference's own builds differ by the default-real twiddles
    sd = pyfft_inputs(ntru, nlat, 2)[0]
    gap = rel(RefPyfft("r4").sp2gp(ntru, nlat, sd), RefPyfft("r8").sp2gp(ntru, nlat, sd))
    assert 1e-8 < gap < TWIDDLE_GAP
