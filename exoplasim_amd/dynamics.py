"""Host-side mirror of the grid-point dynamics between the transforms
(include/exo_dyn.h): calcgp (exoplasim/plasim/src/plasim.f90:2587-2750)
and the Fourier-space d/dlambda of ln(ps) (gpmt, src/plasim.f90:2425-2430),
on torch tensors that live on the GPU.  Everything runs in libexosht.so;
there is no CPU path.
"""
from __future__ import annotations

import ctypes
from ctypes import POINTER, byref, c_double, c_int, c_long, c_void_p

import numpy as np

from ._lib import ExoError, check, lib
from .sht import _ptr, _stream, mode_id

CALCGP_IN = ("gu", "gv", "gt", "gd", "gz", "gq", "dq")
CALCGP_OUT = ("gtn", "gqn", "guz", "gvz", "gphi", "dw")
PRODUCTS = ("gut", "gvt", "gke", "guq", "gvq")


class CalcgpFields(ctypes.Structure):
    """exo_calcgp_fields of include/exo_dyn.h"""
    _fields_ = [(k, c_void_p) for k in CALCGP_IN + ("gp", "gpj", "gpm", "rcsq") + CALCGP_OUT
                + ("gvp",) + PRODUCTS]


_PROTO = {
    "exo_dyn_create": (c_int, [POINTER(c_void_p), c_int] + [POINTER(c_double)] * 8
                       + [c_double] * 5),
    "exo_dyn_destroy": (c_int, [c_void_p]),
    "exo_dyn_calcgp": (c_int, [c_void_p, c_int, c_int, c_long, c_int, c_int, c_int,
                               POINTER(CalcgpFields), c_void_p]),
    "exo_dyn_gpmt": (c_int, [c_int, c_int, c_long, c_void_p, c_void_p, c_void_p]),
    "exo_gridpointa_fields": (c_int, [c_int]),
    "exo_gridpointa": (c_int, [c_void_p, c_void_p, c_int, c_int, c_int, c_void_p, c_long,
                               c_void_p, c_long, c_void_p, c_long, c_void_p]),
}


def work_slots(nlev: int) -> dict:
    """first field (units of NHOR) of each grid field of the gridpointa
    work array (include/exo_dyn.h)"""
    n = nlev
    a = 7 * n + 4
    return {"gz": 0, "gd": n, "gt": 2 * n, "gq": 3 * n, "gp": 4 * n, "gpj": 4 * n + 1,
            "gpm": 4 * n + 2, "gu": 4 * n + 3, "gv": 5 * n + 3, "dq": 6 * n + 3,
            "rcsq": 7 * n + 3, "gtn": a, "gqn": a + n, "guz": a + 2 * n, "gvz": a + 3 * n,
            "gut": a + 4 * n, "gvt": a + 5 * n, "gke": a + 6 * n, "guq": a + 7 * n,
            "gvq": a + 8 * n, "gvp": a + 9 * n, "gphi": a + 9 * n + 1, "dw": a + 10 * n + 1}


def _lib():
    L = lib()
    if not getattr(L, "_dyn_proto", False):
        for name, (res, args) in _PROTO.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        L._dyn_proto = True
    return L


def _d(a) -> ctypes.Array:
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a.ctypes.data_as(POINTER(c_double)), a


class Dynamics:
    """The vertical set-up of calcgp (initsi, src/plasim.f90:1709-1770) on the
    GPU.  c and g are (nlev, nlev) matrices in mathematical order (row,
    column)."""

    def __init__(self, nlev, sigmah, dsigma, rdsig, t0, t01s2, tkp, c, g, akap, rdbrv, adv,
                 psurf, ww):
        keep = [_d(x) for x in (sigmah, dsigma, rdsig, t0, t01s2, tkp)]
        # column-major as the model's arrays: c(i, j) at [j * nlev + i]
        keep += [_d(np.asarray(m, dtype=np.float64).T) for m in (c, g)]
        h = c_void_p()
        check(_lib().exo_dyn_create(byref(h), int(nlev), *[k[0] for k in keep], akap, rdbrv,
                                    adv, psurf, ww), "dyn_create")
        self._h = h
        self.nlev = int(nlev)

    @classmethod
    def from_vertical(cls, v: dict, rdbrv, adv, psurf, ww):
        """from a dict with sigmah dsigma rdsig t0 t01s2 tkp c g akap"""
        return cls(len(v["t0"]), v["sigmah"], v["dsigma"], v["rdsig"], v["t0"], v["t01s2"],
                   v["tkp"], v["c"], v["g"], v["akap"], rdbrv, adv, psurf, ww)

    def close(self):
        if getattr(self, "_h", None):
            _lib().exo_dyn_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def calcgp(self, mode, f: dict, nqspec: int = 1, exp_gp: bool = False, nmem: int = 1,
               member_stride: int = 0):
        """calcgp on device tensors f (gu gv gt gd gz gq [dq]: (nlev, nhor) of
        one member, gp gpj gpm rcsq: (nhor,), outputs gtn gqn guz gvz gphi dw
        (nlev, nhor) and gvp (nhor,), optional products gut gvt gke guq gvq).
        For nmem > 1 every field of member k lies k * member_stride elements
        after the given one."""
        import torch

        dt = torch.float64 if mode_id(mode) == 0 else torch.float32
        nhor = f["gp"].numel() if nmem == 1 else f["gp"].shape[-1]
        need = list(CALCGP_IN[:6]) + ["gp", "gpj", "gpm", "rcsq"] + list(CALCGP_OUT) + ["gvp"]
        if nqspec == 0:
            need.append("dq")
        for k in need:
            if k not in f:
                raise ExoError(f"calcgp: field {k} missing")
        for k, t in f.items():
            if t is None:
                continue
            if t.dtype != dt or not t.is_cuda or not t.is_contiguous():
                raise ExoError(f"calcgp: {k} must be a contiguous {dt} device tensor")
            per = self.nlev * nhor if k in CALCGP_IN + CALCGP_OUT + PRODUCTS else nhor
            span = per + (nmem - 1) * member_stride
            have = t.untyped_storage().nbytes() // t.element_size() - t.storage_offset()
            if t.numel() < per or have < span:
                raise ExoError(f"calcgp: {k} reaches {have} values, {span} needed")
        fs = CalcgpFields(**{k: (_ptr(f[k]) if f.get(k) is not None else None)
                             for k, _ in CalcgpFields._fields_})
        check(_lib().exo_dyn_calcgp(self._h, mode_id(mode), nmem, member_stride, nhor, nqspec,
                                    int(bool(exp_gp)), byref(fs), _stream()), "calcgp")


    def gridpointa(self, plan, mode, sp_state, work, sp_tend, nqspec: int = 1):
        """The transform group of one dynamics step (exo_gridpointa): sp_state
        and sp_tend (nmem, 4*nlev+1, nrsp), work (nmem, fields, nlat, nlon)
        with fields = exo_gridpointa_fields(nlev), device tensors of the
        mode's type."""
        import torch

        dt = torch.float64 if mode_id(mode) == 0 else torch.float32
        nf = _lib().exo_gridpointa_fields(self.nlev)
        nsp = 4 * self.nlev + 1
        nmem = sp_state.shape[0]
        shapes = {"sp_state": (sp_state, (nmem, nsp, plan.nrsp)),
                  "sp_tend": (sp_tend, (nmem, nsp, plan.nrsp)),
                  "work": (work, (nmem, nf, plan.nlat, plan.nlon))}
        for name, (t, shape) in shapes.items():
            if tuple(t.shape) != shape or t.dtype != dt:
                raise ExoError(f"gridpointa {name}: {t.dtype} {tuple(t.shape)}, expected "
                               f"{dt} {shape}")
        check(_lib().exo_gridpointa(plan._h, self._h, mode_id(mode), nmem, nqspec,
                                    _ptr(sp_state), sp_state[0].numel(), _ptr(work),
                                    work[0].numel(), _ptr(sp_tend), sp_tend[0].numel(),
                                    _stream()), "gridpointa")

    def alloc_gridpointa(self, plan, nmem: int, mode):
        """zeroed device arrays (sp_state, work, sp_tend) of gridpointa"""
        import torch

        dt = torch.float64 if mode_id(mode) == 0 else torch.float32
        nf = _lib().exo_gridpointa_fields(self.nlev)
        nsp = 4 * self.nlev + 1
        z = lambda *s: torch.zeros(s, dtype=dt, device="cuda")  # noqa: E731
        return (z(nmem, nsp, plan.nrsp), z(nmem, nf, plan.nlat, plan.nlon),
                z(nmem, nsp, plan.nrsp))


def gpmt(mode, fc, out=None):
    """Fourier coefficients of d/dlambda of every row of fc (rows of NLON:
    re, im of m = 0 .. NLON/2 - 1), src/plasim.f90:2425-2430."""
    out = fc if out is None else out
    if fc.shape != out.shape or fc.dtype != out.dtype:
        raise ExoError("gpmt: in and out of different shapes or types")
    nlon = fc.shape[-1]
    check(_lib().exo_dyn_gpmt(mode_id(mode), nlon, fc.numel() // nlon, _ptr(fc), _ptr(out),
                              _stream()), "gpmt")
    return out
