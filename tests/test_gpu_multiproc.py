"""The multi-process pieces on the GPU (exoplasim_amd.multiproc,
include/exo_mp.h) against the oracle: each process's unfolded partial
fc2sp (ora_fc2sp_unfolded), their mpsumsc reduce-scatter (oracle.sumsc)
and mrtrunc / mrexpand (oracle.mrtrunc / oracle.mrexpand), and the
identity the reference's parallel run relies on: reduce-scattered partial
sums equal the single-process direct transform.

Tolerances: 1e-13 relative L2 in FP64 (sums in another order), 1e-6 in
FP32; mrtrunc / mrexpand are copies and must be exact.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
TOL = {"f64": 1e-13, "f32": 1e-6}


def rel(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / np.linalg.norm(np.ravel(b)))


@pytest.mark.parametrize("ntru,nproc", [(21, 4), (42, 2), (127, 6)])
@pytest.mark.parametrize("mode", ["f64", "f32"])
def test_partials_and_sumsc(oracle, gpu, ntru, nproc, mode):
    import torch

    from exoplasim_amd import multiproc, sht

    print(f"fc2sp partials T{ntru} NPRO={nproc} {mode}", flush=True)
    g, s = oracle.filters(ntru, 3, 1, 1)  # Lander-Hoskins on both directions
    p = sht.Plan(ntru, skgpsp=g, skspgp=s)
    try:
        nlat, nlon, nrsp = p.nlat, p.nlon, p.nrsp
        nlpp = nlat // nproc
        nspp = multiproc.nspp_of(ntru, nproc)
        nesp = nspp * nproc
        nfield = 2
        t = oracle.legini(ntru, nlat)
        fc = np.random.default_rng(ntru).standard_normal((1, nfield, nlat, nlon))
        dt = p.dtype_of(mode)
        dfc = torch.from_numpy(fc).to("cuda", dt)
        # every process's partial sums, stacked as mpsumsc receives them
        psf = torch.full((nproc, nfield, nesp), float("nan"), dtype=dt, device="cuda")
        for r in range(nproc):
            part = psf[r:r + 1]
            multiproc.fc2sp_part(p, mode, dfc, part, nfield, r * nlpp, nlpp)
        got = psf.double().cpu().numpy()
        for r in range(nproc):
            for f in range(nfield):
                want = oracle.fc2sp_unfolded(ntru, nlon, t["qc"][r * nlpp:(r + 1) * nlpp], g,
                                             fc[0, f, r * nlpp:(r + 1) * nlpp])
                assert rel(got[r, f, :nrsp], want) < TOL[mode], (r, f)
                assert not got[r, f, nrsp:].any()
        psp = multiproc.sumsc(mode, psf, nspp).double().cpu().numpy()
        from oracle import sumsc

        assert rel(psp, sumsc(got, nspp)) < TOL[mode]
        # the gathered slices are the single-process folded transform
        full = psp.transpose(1, 0, 2).reshape(nfield, nesp)[:, :nrsp]
        for f in range(nfield):
            assert rel(full[f], oracle.fc2sp(ntru, nlat, nlon, t["qc"], g, fc[0, f])) < \
                TOL[mode] * 10
    finally:
        p.close()


@pytest.mark.parametrize("nthi,ntlo", [(42, 21), (127, 63)])
@pytest.mark.parametrize("mode", ["f64", "f32"])
def test_mrtrunc_mrexpand(gpu, nthi, ntlo, mode):
    import torch

    from oracle import mrexpand, mrtrunc

    from exoplasim_amd import multiproc

    dt = torch.float64 if mode == "f64" else torch.float32
    rng = np.random.default_rng(nthi + ntlo)
    nhi, nlo = (nthi + 1) * (nthi + 2), (ntlo + 1) * (ntlo + 2)
    x = rng.standard_normal((3, nhi))
    if mode == "f32":
        x = x.astype(np.float32).astype(np.float64)
    dx = torch.from_numpy(x).to("cuda", dt)
    lo = torch.full((3, nlo), float("nan"), dtype=dt, device="cuda")
    multiproc.mrtrunc(mode, dx, nthi, ntlo, lo)
    y = lo.double().cpu().numpy()
    for f in range(3):
        assert np.array_equal(y[f], mrtrunc(x[f], nthi, ntlo))
    base = rng.standard_normal((3, nhi))
    if mode == "f32":
        base = base.astype(np.float32).astype(np.float64)
    hi = torch.from_numpy(base).to("cuda", dt)
    multiproc.mrexpand(mode, lo, ntlo, hi, nthi)
    z = hi.double().cpu().numpy()
    for f in range(3):
        assert np.array_equal(z[f], mrexpand(y[f], ntlo, base[f], nthi))
