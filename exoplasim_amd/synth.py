"""Seeded synthetic spectral fields of the benchmark configurations.

Coefficients of every field-level are N(0,1) in real and imaginary part,
scaled by (n+1)^-slope, with the imaginary part of m=0 zeroed (a real field);
fields are ordered vorticity, divergence, temperature (nlev levels each) and
ln(ps).  Spectral order is the reference's (m-major, n inner,
src/legmod.f90:101-116).  Member k of configuration c uses numpy PCG64 with
seed 1000*c + k.
"""
from __future__ import annotations

import numpy as np


def mn_of(ntru: int):
    m = np.concatenate([np.full(ntru + 1 - mm, mm) for mm in range(ntru + 1)])
    n = np.concatenate([np.arange(mm, ntru + 1) for mm in range(ntru + 1)])
    return m, n


def spectral_fields(ntru: int, nfl: int, seed: int, slope: float = 1.5,
                    wave1: float = 0.0) -> np.ndarray:
    """(nfl, 2*ncsp) float64 coefficients of one member."""
    m, n = mn_of(ntru)
    ncsp = m.size
    rng = np.random.default_rng(seed)
    z = rng.standard_normal((nfl, ncsp, 2))
    z *= ((n + 1.0) ** -slope)[None, :, None]
    z[:, m == 0, 1] = 0.0
    if wave1:
        # tidally-locked day-night pattern: Re of (m=1, n=1)
        w11 = np.nonzero((m == 1) & (n == 1))[0][0]
        z[:, w11, 0] += wave1
    return z.reshape(nfl, 2 * ncsp)


def ensemble(ntru: int, nlev: int, nmem: int, cfg: int, slope: float = 1.5,
             wave1: float = 0.0) -> np.ndarray:
    nfl = 3 * nlev + 1
    return np.stack([spectral_fields(ntru, nfl, 1000 * cfg + k, slope, wave1)
                     for k in range(nmem)])
