"""Semi-implicit spectral solve on the GPU in FP64 (include/exo_si.h):
makebm (src/plasim.f90:1683-1703) and the divergence / ln(ps) / temperature
update of spectrala (src/plasim.f90:2817-2876)."""
from __future__ import annotations

from ctypes import byref, c_void_p

import numpy as np

from ._lib import check, lib
from .sht import _cd, _ptr, _stream


class SemiImplicit:
    """g, tau: (nlev, nlev) in math (row, column) order; t0, dsigma (nlev)."""

    def __init__(self, nlev: int, ntru: int, t0, dsigma, g, tau):
        c = lambda a: np.ascontiguousarray(a, dtype=np.float64)  # noqa: E731
        self._keep = [c(t0), c(dsigma), c(np.asarray(g).T), c(np.asarray(tau).T)]
        h = c_void_p()
        check(lib().exo_si_create(byref(h), nlev, ntru, *[_cd(a) for a in self._keep]),
              "si_create")
        self._h = h
        self.nlev, self.ntru = nlev, ntru
        self.nsp = (ntru + 1) * (ntru + 2)

    def close(self):
        if getattr(self, "_h", None):
            lib().exo_si_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def makebm(self, delt: float):
        check(lib().exo_si_makebm(self._h, delt, _stream()), "si_makebm")

    def bm1(self) -> np.ndarray:
        """bm1 as (ntru, nlev, nlev) in math (row, column) order."""
        b = np.zeros((self.ntru, self.nlev, self.nlev))
        check(lib().exo_si_bm1(self._h, _cd(b)), "si_bm1")
        return np.transpose(b, (0, 2, 1)).copy()

    def apply(self, delt, sop, spm, atm, adm, stt, sdt, spt):
        """Device tensors: sop, spm, spt (nsp); atm, adm, stt, sdt (nlev, nsp);
        stt, sdt, spt are updated in place."""
        import torch

        fp64 = int(stt.dtype == torch.float64)
        check(lib().exo_si_apply(self._h, fp64, delt, _ptr(sop), _ptr(spm), _ptr(atm),
                                 _ptr(adm), _ptr(stt), _ptr(sdt), _ptr(spt), _stream()),
              "si_apply")
