"""The device-resident Fourier-space transforms of a plan (exo_sht_sp2fc,
exo_sht_dv2uv, exo_sht_fc2sp, exo_sht_fft, exo_sht_uv2dv, exo_sht_qtend,
exo_sht_mktend) against the oracle, on two members at once with member
strides larger than the fields they touch (a spare field-level before and
after), so that a member or level offset error shows.

Tolerances: 1e-12 relative L2 in FP64, 1e-6 in the FP32 modes (the
multi-input direct transforms sum in FP32 on the vector ALUs there).
"""
import numpy as np
import pytest

from exoplasim_amd import sht

pytestmark = pytest.mark.gpu
TOL = {"f64": 1e-12, "f32": 1e-6, "bf16x3": 1e-6}


def rel(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / np.linalg.norm(np.ravel(b)))


def _seed(shape, seed):
    return np.random.default_rng(seed).standard_normal(shape)


@pytest.mark.parametrize("ntru,nlev", [(21, 10), (42, 4), (127, 2)])
@pytest.mark.parametrize("mode", ["f64", "f32", "bf16x3"])
def test_fourier_space_api(oracle, gpu, ntru, nlev, mode):
    import torch

    from exoplasim_amd.synth import spectral_fields

    print(f"fourier api T{ntru} L{nlev} {mode}", flush=True)
    p = sht.Plan(ntru)
    try:
        nlat, nlon, ncsp = p.nlat, p.nlon, p.ncsp
        tab = oracle.legini(ntru, nlat)
        one = np.ones(ntru + 1)
        dt = p.dtype_of(mode)
        dev = torch.device("cuda")
        nmem, nf = 2, nlev + 2  # one spare field-level before and after
        sp = np.stack([spectral_fields(ntru, nf, 50 + k) for k in range(nmem)])
        spd = torch.from_numpy(sp).to(dev, dt)
        # sp2fc / sp2fcdmu of levels 1..nlev into Fourier fields 1..nlev
        for table in ("qi", "qj"):
            fc = torch.zeros((nmem, nf, nlat, nlon), dtype=dt, device=dev)
            p.sp2fc(mode, spd, fc, nlev, src0=1, dst0=1, table=table)
            got = fc.double().cpu().numpy()
            for k in range(nmem):
                want = np.stack([oracle.sp2fc(ntru, nlat, nlon, tab[table], one, sp[k, 1 + j])
                                 for j in range(nlev)])
                assert rel(got[k, 1:1 + nlev], want) < TOL[mode], (table, k)
            assert not got[:, 0].any() and not got[:, -1].any()
        # dv2uv: vorticity levels 0.., divergence levels 1.. (overlapping runs)
        fc = torch.zeros((nmem, 2 * nlev, nlat, nlon), dtype=dt, device=dev)
        p.dv2uv(mode, spd, fc, nlev, vor0=0, div0=1, u0=0, v0=nlev)
        got = fc.double().cpu().numpy()
        for k in range(nmem):
            pu, pv = oracle.dv2uv(ntru, nlat, nlon, nlev, tab["qu"], tab["qv"], one, sht.PLAVOR,
                                  sp[k, 1:1 + nlev], sp[k, 0:nlev])
            assert rel(got[k, :nlev], pu) < TOL[mode] and rel(got[k, nlev:], pv) < TOL[mode]
        # Fourier inputs for the direct transforms
        f = [_seed((nmem, nlev, nlat, nlon), 70 + i) for i in range(6)]
        fd = [torch.from_numpy(x).to(dev, dt).contiguous() for x in f]
        out = torch.zeros((nmem, nf, 2 * ncsp), dtype=dt, device=dev)
        p.fc2sp(mode, fd[0], out, nlev, src0=0, dst0=1)
        got = out.double().cpu().numpy()
        for k in range(nmem):
            want = np.stack([oracle.fc2sp(ntru, nlat, nlon, tab["qc"], one, f[0][k, j])
                             for j in range(nlev)])
            assert rel(got[k, 1:1 + nlev], want) < TOL[mode]
        sd = [torch.zeros((nmem, nlev, 2 * ncsp), dtype=dt, device=dev) for _ in range(3)]
        p.uv2dv(mode, fd[0], fd[1], sd[0], sd[1], nlev)
        p.qtend(mode, sd[2], fd[0], fd[1], fd[2], nlev)
        got = [x.double().cpu().numpy() for x in sd]
        for k in range(nmem):
            pd, pz = oracle.uv2dv(ntru, nlat, nlon, nlev, tab["qe"], tab["qm"], one, f[0][k],
                                  f[1][k])
            assert rel(got[0][k], pd) < TOL[mode] and rel(got[1][k], pz) < TOL[mode]
            q = oracle.qtend(ntru, nlat, nlon, nlev, tab["qc"], tab["qe"], tab["qm"], one,
                             f[0][k], f[1][k], f[2][k])
            assert rel(got[2][k], q) < TOL[mode]
        p.mktend(mode, sd[0], sd[1], sd[2], *fd, nlev)
        got = [x.double().cpu().numpy() for x in sd]
        for k in range(nmem):
            want = oracle.mktend(ntru, nlat, nlon, nlev, tab["qc"], tab["qe"], tab["qm"],
                                 tab["qq"], one, *[x[k] for x in f])
            for g, w in zip(got, want):
                assert rel(g[k], w) < TOL[mode]
        # in-place FFTs of every row
        rows = torch.from_numpy(f[3]).to(dev, dt).contiguous()
        p.fft(mode, rows, -1)
        assert rel(rows.double().cpu().numpy(), oracle.gp2fc(f[3])) < TOL[mode]
        p.fft(mode, rows, +1)
        assert rel(rows.double().cpu().numpy(),
                   oracle.fc2gp(oracle.gp2fc(f[3]))) < TOL[mode]
    finally:
        p.close()
