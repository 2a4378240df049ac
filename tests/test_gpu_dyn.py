"""The grid-point dynamics on the GPU (exoplasim_amd.dynamics: calcgp and
the Fourier-space d/dlambda gpmt, include/exo_dyn.h) against the oracle's
calcgp (oracle/exo_oracle_dyn.c, pinned against the reference's compiled
calcgp in test_oracle_pins.py) on the same inputs.

Levels 10, 20 and 40 run the kernels with per-level arrays in registers,
15 the run-time level count.  Tolerances, relative L2 per output field:
1e-12 in FP64, 2e-5 in FP32 (inputs rounded to FP32; the geopotential
and sigma-dot sums cancel to about 1/20 of their terms).
"""
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
from make_golden import PLASIM_SCAL, plasim_inputs  # noqa: E402

pytestmark = pytest.mark.gpu
TOL = {"f64": 1e-12, "f32": 2e-5}


# the model holds temperatures divided by ct = (plarad ww)^2 / gascon
# (src/plasim.f90:1319-1320, :1659): a basic state of 220-288 K is about 0.3
CT = (6371220.0 * 7.292e-5) ** 2 / 287.0


def model_vertical(oracle, nlev):
    return oracle.vertical(nlev, t0=np.linspace(220.0, 288.0, nlev) / CT)


def rel(a, b):
    a, b = np.ravel(a), np.ravel(b)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _dev(x, dt):
    import torch

    return torch.from_numpy(np.ascontiguousarray(x)).to("cuda", dt)


@pytest.mark.parametrize("nlev", [10, 20, 40, 15])
@pytest.mark.parametrize("mode", ["f64", "f32"])
@pytest.mark.parametrize("nq", [1, 0])
def test_calcgp_vs_oracle(oracle, gpu, nlev, mode, nq):
    import torch

    from exoplasim_amd.dynamics import Dynamics

    print(f"calcgp L{nlev} {mode} nqspec={nq}", flush=True)
    nhor, nmem = 1000, 2
    v = model_vertical(oracle, nlev)
    dyn = Dynamics.from_vertical(v, *PLASIM_SCAL)
    dt = torch.float64 if mode == "f64" else torch.float32
    try:
        # two members with different inputs, each field of member k at
        # k * stride; ln ps goes in, exp_gp turns it into ps/psurf
        ins = [plasim_inputs(nlev, nhor, 21)[1]]
        f2 = {k: (x[::-1].copy() if x.ndim == 1 else x[::-1, ::-1].copy())
              for k, x in ins[0].items()}
        f2["gp"] = np.exp(-0.5 * np.log(ins[0]["gp"]))
        f2["rcsq"] = ins[0]["rcsq"]
        ins.append(f2)
        level = ("gu", "gv", "gt", "gd", "gz", "gq", "dq", "gtn", "gqn", "guz", "gvz", "gphi",
                 "dw", "gut", "gvt", "gke", "guq", "gvq")
        single = ("gp", "gpj", "gpm", "rcsq", "gvp")
        offs, off = {}, 0
        for k in level + single:
            n = nlev * nhor if k in level else nhor
            offs[k] = (off, n)
            off += n
        stride = off + 7  # members apart by more than their fields
        host = np.full(nmem * stride, np.nan)
        for m in range(nmem):
            for k, x in ins[m].items():
                o, n = offs[k]
                host[m * stride + o:m * stride + o + n] = (np.log(x) if k == "gp" else x).ravel()
        buf = _dev(host, dt)
        views = {k: buf[o:o + n] for k, (o, n) in offs.items()}
        dyn.calcgp(mode, views, nqspec=nq, exp_gp=True, nmem=nmem, member_stride=stride)
        torch.cuda.synchronize()
        out = buf.double().cpu().numpy()
        for m in range(nmem):
            f = {k: x.astype(np.float32).astype(np.float64) if mode == "f32" else x
                 for k, x in ins[m].items()}
            want = oracle.calcgp(f, v, PLASIM_SCAL, nq)
            got = {}
            for k, (o, n) in offs.items():
                x = out[m * stride + o:m * stride + o + n]
                got[k] = x if n == nhor else x.reshape(-1, nhor)
            assert rel(got["gp"], f["gp"]) < (1e-15 if mode == "f64" else 1e-6)
            for k, w in want.items():
                e = rel(got[k], w)
                assert e < TOL[mode], (m, k, e)
            u, vv, t, q = f["gu"], f["gv"], f["gt"], f["gq"]
            assert rel(got["gut"], u * t) < TOL[mode]
            assert rel(got["gvt"], vv * t) < TOL[mode]
            assert rel(got["gke"], u * u + vv * vv + want["gphi"]) < TOL[mode]
            if nq == 1:
                assert rel(got["guq"], u * q) < TOL[mode]
                assert rel(got["gvq"], vv * q) < TOL[mode]
            else:
                assert np.isnan(got["guq"]).all() and np.isnan(got["gvq"]).all()
    finally:
        dyn.close()


@pytest.mark.parametrize("mode", ["f64", "f32"])
def test_gpmt(gpu, mode):
    """gpmt (src/plasim.f90:2425-2430) against its definition: the Fourier
    coefficients of d/dlambda, i m times those of the field."""
    import torch

    from exoplasim_amd.dynamics import gpmt

    dt = torch.float64 if mode == "f64" else torch.float32
    nlon, nrow = 128, 64
    x = np.random.default_rng(3).standard_normal((nrow, nlon))
    out = gpmt(mode, _dev(x, dt), torch.empty((nrow, nlon), dtype=dt, device="cuda"))
    m = np.arange(nlon // 2)
    want = np.empty_like(x)
    want[:, 0::2] = -x[:, 1::2] * m
    want[:, 1::2] = x[:, 0::2] * m
    assert rel(out.double().cpu().numpy(), want) < (1e-15 if mode == "f64" else 1e-7)
    # in place
    y = _dev(x, dt)
    gpmt(mode, y)
    assert rel(y.double().cpu().numpy(), want) < (1e-15 if mode == "f64" else 1e-7)
