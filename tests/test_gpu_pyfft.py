"""The post-processor's transform module (exoplasim_amd.pyfft, the GPU
form of exoplasim/pyfft.f90 as pyburn calls it) against the oracle's
restatement of the same wrappers (oracle.PyfftOracle, kind "r4": the
default-real sqrt 2 literal and filter) and the reference's default-real
build's golden vectors.

Arrays go in and come out in pyburn's Fortran conventions: spectral
(NRSP, NLEV), grid (NLON, NLAT, NLEV), Fourier (2, NLON/2, NLAT, NLEV).
Tolerances: 1e-12 relative L2 against the oracle (FP64 on both sides);
3e-7 against the reference's default-real build, whose FFT twiddles are
single precision (the gap between its own default-real and double builds).
"""
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
from make_golden import pyfft_inputs  # noqa: E402

pytestmark = pytest.mark.gpu
GOLD = os.path.join(HERE, "golden", "reference_pyfft_t21.npz")


def rel(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / np.linalg.norm(np.ravel(b)))


def F(a):
    """C-ordered (nlev, ...) oracle array -> the Fortran-ordered shape."""
    return np.asfortranarray(np.transpose(a))


@pytest.mark.parametrize("ntru,nlat,nlev", [(21, 32, 3), (42, 64, 2), (63, 96, 2)])
@pytest.mark.parametrize("nfs", [0, 1])
def test_pyfft_vs_oracle(oracle, gpu, ntru, nlat, nlev, nfs):
    from oracle import PyfftOracle

    from exoplasim_amd import pyfft

    print(f"pyfft T{ntru} L{nlev} nfs={nfs}", flush=True)
    o = PyfftOracle(oracle, "r4")
    sd, sz, gu, gv, rd, cd = pyfft_inputs(ntru, nlat, nlev)
    nlon = 2 * nlat
    np.testing.assert_array_equal(pyfft.sfilt(ntru, nfs), o.sfilt(ntru, nfs))
    g = pyfft.sp2gp(F(sd), nlat, nlon, ntru, nfs)
    assert g.shape == (nlon, nlat, nlev) and g.flags.f_contiguous
    assert rel(g, F(o.sp2gp(ntru, nlat, sd, nfs))) < 1e-12
    s = pyfft.gp2sp(F(gu), nlat, nlon, ntru, nfs)
    assert s.shape == (2 * (ntru + 1) * (ntru + 2) // 2, nlev)
    assert rel(s, F(o.gp2sp(ntru, gu, nfs))) < 1e-12
    fc = pyfft.sp3fc(F(sd), nlat, nlon, ntru, nfs)
    assert fc.shape == (2, nlon // 2, nlat, nlev)
    want = o.sp3fc(ntru, nlat, sd, nfs).reshape(nlev, nlat, nlon // 2, 2)
    assert rel(fc, F(want)) < 1e-12
    u, v = pyfft.spvgp(F(sd), F(sz), rd, nlon, ntru, nfs)
    ou, ov = o.spvgp(ntru, sd, sz, rd, nfs)
    assert rel(u, F(ou)) < 1e-12 and rel(v, F(ov)) < 1e-12
    d, z = pyfft.gpvsp(F(gu), F(gv), cd, ntru, nfs)
    od, oz = o.gpvsp(ntru, gu, gv, cd, nfs)
    assert rel(d, F(od)) < 1e-12 and rel(z, F(oz)) < 1e-12
    assert rel(pyfft.gp3fc(F(gu)), F(o.gp3fc(gu))) < 1e-12
    assert rel(pyfft.fc3gp(F(gu)), F(o.fc3gp(gu))) < 1e-12
    # the post-processor's round trip (sqrt 2 cancels)
    if nfs == 0:
        assert rel(pyfft.gp2sp(g, nlat, nlon, ntru, 0), F(sd)) < 1e-11


def test_pyfft_vs_reference_golden(gpu):
    from exoplasim_amd import pyfft

    gold = np.load(GOLD)
    ntru, nlat, nlev = 21, 32, 1
    sd, sz, gu, gv, rd, cd = pyfft_inputs(ntru, nlat, nlev)
    for nfs in (0, 1):
        k = f"r4_nfs{nfs}_"
        got = {"sp2gp": pyfft.sp2gp(F(sd), nlat, 2 * nlat, ntru, nfs),
               "gp2sp": pyfft.gp2sp(F(gu), nlat, 2 * nlat, ntru, nfs)}
        got["spvgp_u"], got["spvgp_v"] = pyfft.spvgp(F(sd), F(sz), rd, 2 * nlat, ntru, nfs)
        got["gpvsp_d"], got["gpvsp_z"] = pyfft.gpvsp(F(gu), F(gv), cd, ntru, nfs)
        for name, x in got.items():
            e = rel(x, F(gold[k + name]))
            print(f"nfs={nfs} {name} vs default-real reference {e:.1e}", flush=True)
            assert e < 3e-7, (nfs, name, e)
        fc = pyfft.sp3fc(F(sd), nlat, 2 * nlat, ntru, nfs)
        want = gold[k + "sp3fc"].reshape(nlev, nlat, nlat, 2)
        assert rel(fc, F(want)) < 1e-13


def test_pyfft_rejects_unsupported(gpu):
    from exoplasim_amd import pyfft
    from exoplasim_amd._lib import ExoError

    with pytest.raises(ExoError, match="not supported"):
        pyfft.sp2gp(np.zeros((2 * 86 * 87 // 2, 1), order="F"), 128, 256, 85, 0)
    with pytest.raises(ExoError):
        pyfft.sp2gp(np.zeros((10, 1)), 32, 64, 21, 0)


@pytest.mark.parametrize("mode", ["f64", "f32"])
def test_writesp_and_the_postprocessor_chain(oracle, gpu, mode):
    """writesp's output values (exoplasim_amd.output.writesp, the model's
    spectral output format, src/outmod.f90:141-169) against the oracle's
    restatement, bit for bit; then the post-processor's chain on them:
    pyfft.sp2gp of the written coefficients, rolled by half the longitudes
    as pyburn does (exoplasim/pyburn.py:711-713), is the model's own grid of
    the field (sp2fc + fc2gp), up to the single-precision sqrt 2 literal."""
    import torch

    from oracle import writesp as ora_writesp

    from exoplasim_amd import pyfft
    from exoplasim_amd.output import writesp

    ntru, nlat, nlev = 42, 64, 3
    sp = pyfft_inputs(ntru, nlat, nlev)[0]
    real = np.float64 if mode == "f64" else np.float32
    dt = torch.float64 if mode == "f64" else torch.float32
    poff = np.array([0.0, 2.5, -1.0])
    z = writesp(mode, ntru, torch.from_numpy(sp.astype(real)).cuda(), 7.0,
                torch.from_numpy(poff).cuda()).cpu().numpy()
    for k in range(nlev):
        np.testing.assert_array_equal(z[k], ora_writesp(sp[k].astype(real), ntru, 7.0, poff[k],
                                                        real))
    # the chain, on the scale-1, offset-free output
    z1 = writesp(mode, ntru, torch.from_numpy(sp.astype(real)).cuda(), 1.0).cpu().numpy()
    g = pyfft.sp2gp(F(z1.astype(np.float64)), nlat, 2 * nlat, ntru, 0)
    g = np.roll(np.transpose(g, (2, 1, 0)), nlat, axis=-1)
    t = oracle.legini(ntru, nlat)
    one = np.ones(ntru + 1)
    want = oracle.fc2gp(np.stack([oracle.sp2fc(ntru, nlat, 2 * nlat, t["qi"], one, x)
                                  for x in sp]))
    assert rel(g, want) < (1e-7 if mode == "f64" else 3e-7)
