"""exoplasim_amd -- MI355X-native spectral-transform core of PlaSim/ExoPlaSim.

The grid-point <-> spectral round trip of the reference's legmod/fftmod
(exoplasim/plasim/src/legmod.f90, fftmod.f90, fft991mod.f90, gaussmod.f90)
as hand-written HIP kernels for gfx950 behind a C ABI (include/exo_sht.h),
with a Fortran ISO_C_BINDING shim exporting the original subroutine names
and this Python mirror of the same interface.
"""
from . import sht  # noqa: F401
from ._lib import ExoError, lib  # noqa: F401

__all__ = ["sht", "lib", "ExoError"]
