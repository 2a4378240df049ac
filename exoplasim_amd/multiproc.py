"""Host-side mirror of the multi-process (NPRO > 1) pieces of PlaSim's
transform path on one GPU (include/exo_mp.h): the unfolded partial fc2sp
of one process's latitudes (exoplasim/plasim/src/legmod.f90:191-202),
mpsumsc's sum and scatter (src/mpimod.f90:279-291) and mrtrunc / mrexpand
(src/mpimod_multi.f90:117-153), on torch tensors that live on the GPU.
"""
from __future__ import annotations

from ctypes import c_int, c_long, c_void_p

from ._lib import ExoError, check, lib
from .sht import _fields, _ptr, _stream, mode_id

_PROTO = {
    "exo_sht_fc2sp_part": (c_int, [c_void_p, c_int, c_int, c_void_p, c_long, c_int, c_int,
                                   c_int, c_void_p, c_long, c_int, c_int, c_int, c_void_p]),
    "exo_sht_sumsc": (c_int, [c_int, c_int, c_int, c_int, c_void_p, c_void_p, c_void_p]),
    "exo_sht_mrtrunc": (c_int, [c_int, c_int, c_int, c_int, c_void_p, c_long, c_void_p, c_long,
                                c_void_p]),
    "exo_sht_mrexpand": (c_int, [c_int, c_int, c_int, c_int, c_void_p, c_long, c_void_p,
                                 c_long, c_void_p]),
}


def _lib():
    L = lib()
    if not getattr(L, "_mp_proto", False):
        for name, (res, args) in _PROTO.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        L._mp_proto = True
    return L


def _dt(mode):
    import torch

    return torch.float64 if mode_id(mode) == 0 else torch.float32


def nspp_of(ntru: int, nproc: int) -> int:
    """modes per process, NSPP = ceil(NRSP / NPRO) (src/plasimmod.f90:112)"""
    nrsp = (ntru + 1) * (ntru + 2)
    return (nrsp + nproc - 1) // nproc


def fc2sp_part(plan, mode, fc, sp, nfield, lat0, nlpp, src0=0, dst0=0):
    """Partial direct Legendre sums of the latitudes [lat0, lat0+nlpp) of
    Fourier fields fc (nmem, >= src0+nfield, nlat, nlon) into sp (nmem,
    >= dst0+nfield, nesp), nesp >= nrsp (zero past nrsp)."""
    nmem = fc.shape[0]
    dt = _dt(mode)
    _fields("fc2sp_part fc", fc, dt, (plan.nlat, plan.nlon), nmem, (src0, nfield))
    if sp.dim() != 3 or sp.shape[2] < plan.nrsp:
        raise ExoError(f"fc2sp_part sp: shape {tuple(sp.shape)}, expected (nmem, nfield, nesp "
                       f">= {plan.nrsp})")
    _fields("fc2sp_part sp", sp, dt, (sp.shape[2],), nmem, (dst0, nfield))
    check(_lib().exo_sht_fc2sp_part(plan._h, mode_id(mode), nmem, _ptr(fc), fc[0].numel(), src0,
                                    lat0, nlpp, _ptr(sp), sp[0].numel(), dst0, sp.shape[2],
                                    nfield, _stream()), "fc2sp_part")


def sumsc(mode, psf, nspp):
    """mpsumsc: psf (nproc, nfield, nproc*nspp) partial sums -> psp (nproc,
    nfield, nspp), process r's slice of the sum over processes."""
    import torch

    nproc, nfield, nesp = psf.shape
    if nesp != nproc * nspp or psf.dtype != _dt(mode):
        raise ExoError(f"sumsc: psf {psf.dtype} {tuple(psf.shape)}, expected NESP = "
                       f"{nproc} x {nspp}")
    psp = torch.empty((nproc, nfield, nspp), dtype=psf.dtype, device=psf.device)
    check(_lib().exo_sht_sumsc(mode_id(mode), nproc, nspp, nfield, _ptr(psf), _ptr(psp),
                               _stream()), "sumsc")
    return psp


def _ncsp(nt):
    return (nt + 1) * (nt + 2) // 2


def mrtrunc(mode, sphi, nthi, ntlo, splo):
    """mrtrunc of every row of sphi (nfield, >= 2 NCSP(nthi)) into splo
    (nfield, >= 2 NCSP(ntlo))"""
    if sphi.shape[-1] < 2 * _ncsp(nthi) or splo.shape[-1] < 2 * _ncsp(ntlo) or \
            sphi.shape[0] != splo.shape[0] or sphi.dtype != _dt(mode) or \
            splo.dtype != _dt(mode):
        raise ExoError("mrtrunc: shapes or types do not match the truncations and mode")
    check(_lib().exo_sht_mrtrunc(mode_id(mode), nthi, ntlo, sphi.shape[0], _ptr(sphi),
                                 sphi.shape[-1], _ptr(splo), splo.shape[-1], _stream()),
          "mrtrunc")


def mrexpand(mode, splo, ntlo, sphi, nthi):
    """mrexpand of every row of splo into sphi (in place on sphi)"""
    if sphi.shape[-1] < 2 * _ncsp(nthi) or splo.shape[-1] < 2 * _ncsp(ntlo) or \
            sphi.shape[0] != splo.shape[0] or sphi.dtype != _dt(mode) or \
            splo.dtype != _dt(mode):
        raise ExoError("mrexpand: shapes or types do not match the truncations and mode")
    check(_lib().exo_sht_mrexpand(mode_id(mode), ntlo, nthi, splo.shape[0], _ptr(splo),
                                  splo.shape[-1], _ptr(sphi), sphi.shape[-1], _stream()),
          "mrexpand")
