"""The device-resident gridpointa transform group (exo_gridpointa,
include/exo_dyn.h) against the oracle's gridpointa (Oracle.gridpointa,
pinned against the reference's compiled gridpointa in
test_oracle_gridpointa.py): the spectral tendencies szt sdt stt sqt spt of
two members per call, and the grid fields the group leaves in its work
array.

Tolerances, relative L2 per field: 1e-12 in FP64; 2e-5 in the FP32 modes
(FP32 grid-point arithmetic and FP32 interface arrays).
"""
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
from make_golden import PLASIM_SCAL, gridpointa_humidity, gridpointa_state  # noqa: E402

pytestmark = pytest.mark.gpu
TOL = {"f64": 1e-12, "f32": 2e-5, "bf16x3": 2e-5}


# the model holds temperatures divided by ct = (plarad ww)^2 / gascon
# (src/plasim.f90:1319-1320, :1659): a basic state of 220-288 K is about 0.3
CT = (6371220.0 * 7.292e-5) ** 2 / 287.0


def model_vertical(oracle, nlev):
    return oracle.vertical(nlev, t0=np.linspace(220.0, 288.0, nlev) / CT)


def rel(a, b):
    a, b = np.ravel(a), np.ravel(b)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.mark.parametrize("ntru,nlev,mode", [(21, 10, "f64"), (21, 10, "f32"), (21, 10, "bf16x3"),
                                            (42, 20, "f64"), (42, 20, "bf16x3"),
                                            (63, 20, "f64")])
@pytest.mark.parametrize("nq", [1, 0])
def test_gridpointa_vs_oracle(oracle, gpu, ntru, nlev, mode, nq):
    import torch

    from exoplasim_amd import sht
    from exoplasim_amd.dynamics import Dynamics, work_slots

    print(f"gridpointa T{ntru} L{nlev} {mode} nqspec={nq}", flush=True)
    v = model_vertical(oracle, nlev)
    plan = sht.Plan(ntru)
    dyn = Dynamics.from_vertical(v, *PLASIM_SCAL)
    try:
        nmem = 2
        nhor = plan.nlat * plan.nlon
        states = [gridpointa_state(ntru, nlev, 0.9 + 0.4 * k) for k in range(nmem)]
        hums = [gridpointa_humidity(nlev, nhor) for _ in range(nmem)]
        hums[1] = (hums[1][0][::-1].copy(), hums[1][1][::-1].copy())
        st, work, tend = dyn.alloc_gridpointa(plan, nmem, mode)
        dt = st.dtype
        st.copy_(torch.from_numpy(np.stack(states)).to(dt))
        sl = work_slots(nlev)
        w = work.view(nmem, -1, nhor)
        for k in range(nmem):
            w[k, sl["gq"]:sl["gq"] + nlev] = torch.from_numpy(hums[k][0]).to(dt)
            w[k, sl["dq"]:sl["dq"] + nlev] = torch.from_numpy(hums[k][1]).to(dt)
        tend.fill_(float("nan"))
        dyn.gridpointa(plan, mode, st, work, tend, nqspec=nq)
        torch.cuda.synchronize()
        got = tend.double().cpu().numpy()
        wk = w.double().cpu().numpy()
        for k in range(nmem):
            sk = states[k]
            if mode != "f64":
                sk = sk.astype(np.float32).astype(np.float64)
            want, f = oracle.gridpointa(ntru, nlev, sk, v, PLASIM_SCAL, nq, gq=hums[k][0],
                                        dq=hums[k][1])
            for i, name in enumerate(("szt", "sdt", "stt", "sqt")):
                part = slice(i * nlev, (i + 1) * nlev)
                if name == "sqt" and nq == 0:
                    assert np.isnan(got[k, part]).all()  # left as it was
                    continue
                e = rel(got[k, part], want[part])
                print(f"  member {k} {name} {e:.1e}", flush=True)
                assert e < TOL[mode], (k, name, e)
            assert rel(got[k, 4 * nlev], want[4 * nlev]) < TOL[mode]
            for name in ("gz", "gd", "gt", "gu", "gv", "gphi", "dw"):
                x = wk[k, sl[name]:sl[name] + nlev]
                assert rel(x, f[name]) < TOL[mode], (k, name, rel(x, f[name]))
            for name in ("gp", "gpj", "gpm", "rcsq"):
                assert rel(wk[k, sl[name]], f[name]) < TOL[mode], (k, name)
    finally:
        dyn.close()
        plan.close()
