"""Host-side mirror of PlaSim's spectral output normalisation on the GPU
(include/exo_out.h): writesp (exoplasim/plasim/src/outmod.f90:141-169)
without the file write, on torch tensors that live on the GPU."""
from __future__ import annotations

from ctypes import POINTER, c_double, c_float, c_int, c_long, c_void_p

from ._lib import ExoError, check, lib
from .sht import _ptr, _stream, mode_id

_PROTO = {"exo_sht_writesp": (c_int, [c_int, c_int, c_int, c_void_p, c_long, c_double,
                                      c_void_p, c_void_p, c_long, c_void_p])}


def _lib():
    L = lib()
    if not getattr(L, "_out_proto", False):
        for name, (res, args) in _PROTO.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        L._out_proto = True
    return L


def writesp(mode, ntru: int, pf, pscale: float, poff=None):
    """the single-precision output values zf = pf * spnorm * pscale (+ poff
    on the first value) of every row of pf (nfield, NRSP): double rows for
    mode "f64", float otherwise; poff a device tensor of nfield doubles or
    None."""
    import torch

    nrsp = (ntru + 1) * (ntru + 2)
    dt = torch.float64 if mode_id(mode) == 0 else torch.float32
    if pf.dim() != 2 or pf.shape[1] != nrsp or pf.dtype != dt:
        raise ExoError(f"writesp: {pf.dtype} {tuple(pf.shape)}, expected {dt} (nfield, {nrsp})")
    if poff is not None and (poff.dtype != torch.float64 or poff.numel() != pf.shape[0]):
        raise ExoError("writesp: poff must hold one double per field")
    zf = torch.empty((pf.shape[0], nrsp), dtype=torch.float32, device=pf.device)
    check(_lib().exo_sht_writesp(mode_id(mode), ntru, pf.shape[0], _ptr(pf), nrsp, pscale,
                                 _ptr(poff) if poff is not None else None, _ptr(zf), nrsp,
                                 _stream()), "writesp")
    return zf
