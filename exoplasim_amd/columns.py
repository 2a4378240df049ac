"""Seeded synthetic atmospheric columns of the learned long-wave workload.

Each column (nlev sigma levels of the neqsig = 0 model setup,
src/plasim.f90:1553-1567) carries:
  * temperature: surface T_s ~ U(200, 330) K, lapse rate U(4, 8) K/km over
    hypsometric heights z = (R T / g) ln(1/sigma), clipped to [150, 350] K;
  * specific humidity log-uniform in [1e-7, 2e-2] kg/kg per level;
  * surface pressure U(0.1, 10) bar;
  * CO2 log-uniform in [1, 1e5] ppmv, uniform in the column;
  * clear sky (dcc = dql = 0), no ozone, ocean surface (dls = 0).
"""
from __future__ import annotations

import numpy as np

GA = 9.80665
GASCON = 287.0


def sigma_levels(nlev: int):
    """sigma, sigmah, dsigma of the neqsig = 0 configuration."""
    zsk = np.arange(1, nlev + 1) / nlev
    sigmah = 0.75 * zsk + 1.75 * zsk ** 3 - 1.5 * zsk ** 4
    dsigma = np.diff(np.concatenate([[0.0], sigmah]))
    sigma = np.concatenate([[0.5 * sigmah[0]], 0.5 * (sigmah[:-1] + sigmah[1:])])
    return sigma, sigmah, dsigma


def make_columns(ncol: int, nlev: int, seed: int) -> dict:
    rng = np.random.default_rng(seed)
    sigma, _, _ = sigma_levels(nlev)
    ts = rng.uniform(200.0, 330.0, ncol)
    lapse = rng.uniform(4.0, 8.0, ncol) * 1e-3  # K/m
    # height from the hypsometric relation at the surface temperature
    z = GASCON * ts[:, None] / GA * np.log(1.0 / sigma[None, :])
    t = np.clip(ts[:, None] - lapse[:, None] * z, 150.0, 350.0)
    dt = np.concatenate([t, ts[:, None]], axis=1)
    q = np.exp(rng.uniform(np.log(1e-7), np.log(2e-2), (ncol, nlev)))
    ps = rng.uniform(0.1, 10.0, ncol) * 1e5
    co2 = np.exp(rng.uniform(0.0, np.log(1e5), ncol))
    return {"dt": dt, "dq": q, "dp": ps, "dqco2": np.repeat(co2[:, None], nlev, axis=1),
            "dqo3": np.zeros((ncol, nlev)), "dcc": np.zeros((ncol, nlev)),
            "dql": np.zeros((ncol, nlev)), "dls": np.zeros(ncol)}
