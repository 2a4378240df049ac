"""exo_sht_roundtrip_noalloc against exo_sht_roundtrip through the C ABI.

The no-allocation round trip runs the same kernels on the library's fixed
per-GPU scratch array, in chunks of members whose scratch fits it, so its
outputs must equal exo_sht_roundtrip's bit for bit.  Each case makes two
calls in a row, the second with fewer members, and checks both; the cases
cover a split into several chunks (T42 L20 split-BF16, 96 members: the
pre-split operand of about 70 members fits), one member per chunk (T127 L40
split-BF16: pre-split operand plus three latitude-pair blocks of partial
direct sums), the partial sums alone (T63 L20 FP64) and the size the array
cannot hold (T127 L40 FP64, 96 MB of partial sums per member).
"""
import numpy as np
import pytest

from exoplasim_amd import sht
from exoplasim_amd._lib import ExoError
from exoplasim_amd.synth import ensemble

pytestmark = pytest.mark.gpu


def _run_both(plan, mode, nlev, sp):
    import torch

    nmem = sp.shape[0]
    a = plan.alloc_roundtrip(nmem, nlev, mode)
    b = plan.alloc_roundtrip(nmem, nlev, mode)
    a[0].copy_(torch.from_numpy(sp).to(a[0].dtype))
    b[0].copy_(a[0])
    # outputs start as NaN, so that a member the call misses shows
    for t in (a[1], a[2], b[1], b[2]):
        t.fill_(float("nan"))
    plan.roundtrip(mode, a[0], a[1], a[2], nlev)
    plan.roundtrip_noalloc(mode, b[0], b[1], b[2], nlev)
    torch.cuda.synchronize()
    return a, b


def _same(a, b, what):
    import torch

    assert not torch.isnan(b[1]).any(), (what, "grid")
    assert not torch.isnan(b[2]).any(), (what, "spectral")
    assert torch.equal(a[1], b[1]), (what, "grid", float((a[1] - b[1]).abs().max()))
    assert torch.equal(a[2], b[2]), (what, "spectral", float((a[2] - b[2]).abs().max()))


@pytest.mark.parametrize("ntru,nlev,mode,nmem1,nmem2,slope,wave1", [
    (42, 20, "bf16x3", 96, 5, 1.5, 0.0),
    (127, 40, "bf16x3", 3, 1, 2.0, 10.0),
    (63, 20, "f64", 20, 3, 1.5, 0.0),
])
def test_noalloc_equals_roundtrip(oracle, gpu, ntru, nlev, mode, nmem1, nmem2, slope, wave1):
    plan = sht.Plan(ntru)
    try:
        for call, nmem in enumerate((nmem1, nmem2)):
            print(f"noalloc T{ntru} L{nlev} {mode} call {call + 1}: {nmem} members", flush=True)
            sp = ensemble(ntru, nlev, nmem, cfg=4 if ntru == 127 else 2, slope=slope,
                          wave1=wave1)
            if call == 1:
                sp = sp[::-1].copy()  # other data than the first call's leading members
            a, b = _run_both(plan, mode, nlev, sp)
            _same(a, b, f"T{ntru} {mode} call {call + 1}")
        # and the last call's first member against the oracle
        if ntru != 127:
            g_ref, s_ref = oracle.roundtrip(ntru, nlev, sp[0])
            tol = 1e-12 if mode == "f64" else 1e-6
            g = b[1][0].double().cpu().numpy()
            s = b[2][0].double().cpu().numpy()
            assert np.linalg.norm(g - g_ref) / np.linalg.norm(g_ref) < tol
            assert np.linalg.norm(s - s_ref) / np.linalg.norm(s_ref) < tol
    finally:
        plan.close()


def test_noalloc_rejects_oversize(gpu):
    """T127 L40 in FP64 needs 6 latitude-pair blocks x 121 field-levels x
    16512 reals x 8 B = 96 MB of partial sums per member: more than the
    array holds, refused before anything is launched."""
    plan = sht.Plan(127)
    try:
        nlev = 40
        a = plan.alloc_roundtrip(1, nlev, "f64")
        print("noalloc T127 L40 f64: expect EXO_ERR_SCRATCH", flush=True)
        with pytest.raises(ExoError, match="scratch buffer too small"):
            plan.roundtrip_noalloc("f64", a[0], a[1], a[2], nlev)
    finally:
        plan.close()
