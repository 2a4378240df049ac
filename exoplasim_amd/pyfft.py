"""The post-processor's transform module (exoplasim/pyfft.f90, imported by
exoplasim/pyburn.py as ``pyfft``) served by the MI355X kernels.

Same function names, argument order and array conventions as the f2py
module pyburn calls (exoplasim/pyburn.py:606,711,810,840,867,1067,1172):
arrays are float64 in Fortran order, spectral fields (NRSP, NLEV), grid
fields (NLON, NLAT, NLEV), Fourier fields (2, NLON/2, NLAT, NLEV).  The
reference's own conventions are kept:

* its legini (pyfft.f90:8-155) has no planetary vorticity in dv2uv and one
  physics filter for both directions, sfilt(n) = exp(-8 (n/NTRU)^8) for
  n = 1 .. NTRU+1 when nfs = 1 (pyfft.f90:150-153), evaluated in default
  (single) precision as the f2py build does;
* fc2gp multiplies by the default-real literal 1.4142135623730951
  (pyfft.f90:891), gp2sp and gpvsp divide by it (pyfft.f90:1162, 1112), so
  the factor is sqrt(2) rounded to single precision.

Every transform runs on the GPU (FP64 tables and arithmetic); truncations
21, 42, 63 and 127 on their Gaussian grids.  There is no CPU path.
"""
from __future__ import annotations

import numpy as np

from . import sht
from ._lib import ExoError

# the default-real literal 1.4142135623730951 of pyfft.f90
SQRT2 = float(np.float32(1.4142135623730951))

_plans: dict = {}


def sfilt(ntru: int, nfs: int) -> np.ndarray:
    """pyfft.f90:150-153: sfilt(n) = (1-nfs) + nfs exp(-8 (real(n)/NTRU)**8),
    n = 1 .. NTRU+1, index n-1 of the result, in single precision as the
    default-real build evaluates it: x**8 as seven successive products, then
    exp rounded to single precision."""
    if nfs not in (0, 1):
        raise ExoError(f"nfs = {nfs}, expected 0 or 1")
    out = np.ones(ntru + 1)
    if nfs == 1:
        x = np.arange(1, ntru + 2, dtype=np.float32) / np.float32(ntru)
        p = x.copy()
        for _ in range(7):
            p = p * x
        y = np.float32(-8.0) * p
        out = np.exp(y.astype(np.float64)).astype(np.float32).astype(np.float64)
    return out


def _plan(ntru: int, nlat: int, nfs: int) -> sht.Plan:
    if sht.NLAT_OF.get(ntru) != nlat:
        raise ExoError(f"T{ntru} on {nlat} latitudes is not supported "
                       f"(truncations {sorted(sht.NLAT_OF)} on their Gaussian grids)")
    key = (ntru, int(nfs))
    if key not in _plans:
        f = sfilt(ntru, int(nfs))
        _plans[key] = sht.Plan(ntru, skgpsp=f, skspgp=f, plavor=0.0)
    return _plans[key]


def _dev(a):
    import torch

    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).cuda()


def _host(t) -> np.ndarray:
    import torch

    torch.cuda.synchronize()
    return t.cpu().numpy()


def _levels(a, nrows: int, what: str) -> np.ndarray:
    """(nrows, nlev) Fortran array -> C-contiguous (nlev, nrows)."""
    a = np.asarray(a, dtype=np.float64)
    if a.ndim == 1:
        a = a[:, None]
    if a.ndim != 2 or a.shape[0] != nrows:
        raise ExoError(f"{what}: shape {a.shape}, expected ({nrows}, nlev)")
    return np.ascontiguousarray(a.T)


def _grid(a, what: str):
    """(nlon, nlat[, nlev]) Fortran array -> C-contiguous (nlev, nlat, nlon)."""
    a = np.asarray(a, dtype=np.float64)
    if a.ndim == 2:
        a = a[:, :, None]
    if a.ndim != 3:
        raise ExoError(f"{what}: shape {a.shape}, expected (nlon, nlat, nlev)")
    return np.ascontiguousarray(np.transpose(a, (2, 1, 0)))


def _fortran_grid(g: np.ndarray) -> np.ndarray:
    """(nlev, nlat, nlon) -> the (nlon, nlat, nlev) Fortran-ordered array."""
    return np.asfortranarray(np.transpose(g, (2, 1, 0)))


def inigau(nlat: int):
    """Gaussian abscissas and weights (pyfft.f90:697-732 = gaussmod)."""
    return sht.inigau(nlat)


def sp2gp(sp, nlat: int, nlon: int, ntru: int, nfs: int = 0) -> np.ndarray:
    """pyfft.f90:971-1005: sp2fc then fc2gp (times SQRT2) of every level."""
    p = _plan(ntru, nlat, nfs)
    if nlon != p.nlon:
        raise ExoError(f"sp2gp: nlon {nlon}, expected {p.nlon}")
    s = _levels(sp, p.nrsp, "sp2gp sp")
    nlev = s.shape[0]
    ds = _dev(s)[None]
    g = ds.new_zeros((1, nlev, p.nlat, p.nlon))
    p.inverse("f64", ds, g, nlev)
    g *= SQRT2
    return _fortran_grid(_host(g[0]))


def gp2sp(gp, nlat: int, nlon: int, ntru: int, nfs: int = 0) -> np.ndarray:
    """pyfft.f90:1141-1174: gp / SQRT2, gp2fc, fc2sp of every level."""
    p = _plan(ntru, nlat, nfs)
    g = _grid(gp, "gp2sp gp")
    if g.shape[1:] != (p.nlat, p.nlon) or nlon != p.nlon:
        raise ExoError(f"gp2sp: grid {g.shape[1:]}, expected ({p.nlat}, {p.nlon})")
    nlev = g.shape[0]
    dg = _dev(g)[None] / SQRT2
    s = dg.new_zeros((1, nlev, p.nrsp))
    p.direct("f64", dg, s, nlev)
    return np.asfortranarray(_host(s[0]).T)


def sp3fc(sp, nlat: int, nlon: int, ntru: int, nfs: int = 0) -> np.ndarray:
    """pyfft.f90:344-370: sp2fc of every level, fc(2, NLON/2, NLAT, NLEV)."""
    p = _plan(ntru, nlat, nfs)
    if nlon != p.nlon:
        raise ExoError(f"sp3fc: nlon {nlon}, expected {p.nlon}")
    s = _levels(sp, p.nrsp, "sp3fc sp")
    nlev = s.shape[0]
    ds = _dev(s)[None]
    fc = ds.new_zeros((1, nlev, p.nlat, p.nlon))
    p.sp2fc("f64", ds, fc, nlev)
    f = _host(fc[0]).reshape(nlev, p.nlat, p.nlon // 2, 2)
    return np.asfortranarray(np.transpose(f, (3, 2, 1, 0)))


def _fft3(a, direction: int, what: str) -> np.ndarray:
    g = _grid(a, what)
    ntru = {v: k for k, v in sht.NLAT_OF.items()}.get(g.shape[1])
    if ntru is None or g.shape[2] != 2 * g.shape[1]:
        raise ExoError(f"{what}: grid {g.shape[1:]} has no supported FFT length")
    p = _plan(ntru, g.shape[1], 0)
    d = _dev(g)
    p.fft("f64", d, direction)
    if direction > 0:
        d *= SQRT2
    return _fortran_grid(_host(d))


def gp3fc(gp) -> np.ndarray:
    """pyfft.f90:938-963: gp2fc of every level (no scaling)."""
    return _fft3(gp, -1, "gp3fc")


def fc3gp(fc) -> np.ndarray:
    """pyfft.f90:905-930: fc2gp (times SQRT2) of every level."""
    return _fft3(fc, +1, "fc3gp")


def spvgp(sd, sz, rdcostheta, nlon: int, ntru: int, nfs: int = 0):
    """pyfft.f90:1012-1069: dv2uv (no planetary vorticity), fc2gp (times
    SQRT2), then times rdcostheta(lat); returns (gu, gv)."""
    rd = np.ascontiguousarray(rdcostheta, dtype=np.float64)
    p = _plan(ntru, rd.size, nfs)
    if nlon != p.nlon:
        raise ExoError(f"spvgp: nlon {nlon}, expected {p.nlon}")
    d = _levels(sd, p.nrsp, "spvgp sd")
    z = _levels(sz, p.nrsp, "spvgp sz")
    if d.shape != z.shape:
        raise ExoError("spvgp: sd and sz of different shapes")
    nlev = d.shape[0]
    ds = _dev(np.concatenate([z, d]))[None]
    g = ds.new_zeros((1, 2 * nlev, p.nlat, p.nlon))
    p.inverse_uv("f64", ds, g, nlev, vor0=0, div0=nlev, u0=0, v0=nlev)
    g *= SQRT2
    g *= _dev(rd)[None, None, :, None]
    h = _host(g[0])
    return _fortran_grid(h[:nlev]), _fortran_grid(h[nlev:])


def gpvsp(gu, gv, costhetadr, ntru: int, nfs: int = 0):
    """pyfft.f90:1076-1134: (gu, gv) times costhetadr(lat) / SQRT2, gp2fc,
    uv2dv (folded); returns (sd, sz)."""
    u = _grid(gu, "gpvsp gu")
    v = _grid(gv, "gpvsp gv")
    if u.shape != v.shape:
        raise ExoError("gpvsp: gu and gv of different shapes")
    cd = np.ascontiguousarray(costhetadr, dtype=np.float64)
    p = _plan(ntru, u.shape[1], nfs)
    if cd.size != p.nlat or u.shape[2] != p.nlon:
        raise ExoError(f"gpvsp: grid {u.shape[1:]} / {cd.size} latitudes, expected "
                       f"({p.nlat}, {p.nlon})")
    nlev = u.shape[0]
    w = _dev(cd)[None, None, :, None]
    fu = (_dev(u)[None] * w) / SQRT2
    fv = (_dev(v)[None] * w) / SQRT2
    p.fft("f64", fu, -1)
    p.fft("f64", fv, -1)
    pd = fu.new_zeros((1, nlev, p.nrsp))
    pz = fu.new_zeros((1, nlev, p.nrsp))
    p.uv2dv("f64", fu, fv, pd, pz, nlev)
    return np.asfortranarray(_host(pd[0]).T), np.asfortranarray(_host(pz[0]).T)
