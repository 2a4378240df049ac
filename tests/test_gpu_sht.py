"""GPU parity of the HIP spectral transforms against the oracle.

Every test runs the HIP path through the C ABI (libexosht.so) and compares
with the FP64 CPU oracle (oracle/exo_oracle.c, itself pinned against the
reference's compiled Fortran and closed forms in test_oracle.py) on the same
seeded inputs.

Tolerances (relative L2 per field, max-abs relative to the field's max):
  f64                      1e-12   (FP64 arithmetic, different summation order)
  f32, f32mfma, bf16x3     1e-6    (the bound of the split-BF16 target,
                                    FP32 interface arrays)
"""
import numpy as np
import pytest

from exoplasim_amd import sht
from exoplasim_amd.synth import ensemble

pytestmark = pytest.mark.gpu

TOL = {"f64": 1e-12, "f32": 1e-6, "f32mfma": 1e-6, "bf16x3": 1e-6}
TRUNCS = [(21, 10), (42, 20), (63, 20), (127, 4)]


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64).ravel()
    b = np.asarray(b, dtype=np.float64).ravel()
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.fixture(scope="module")
def plans(gpu):
    ps = {}
    yield ps
    for p in ps.values():
        p.close()


def get_plan(plans, ntru):
    if ntru not in plans:
        plans[ntru] = sht.Plan(ntru)
    return plans[ntru]


@pytest.mark.parametrize("ntru", [21, 42, 63, 127])
def test_grid(plans, oracle, ntru):
    p = get_plan(plans, ntru)
    sid, gwd = p.grid()
    osid, ogwd = oracle.inigau(p.nlat)
    print(f"T{ntru} grid", flush=True)
    # GPU cos/acos differ from the host libm by an ulp.  Newton's iteration
    # stops once its step is below 1e-16 in absolute value
    # (src/gaussmod.f90:30), so each side's abscissa lies within that step
    # plus one rounding (1.1e-16 near 1) of the root: 4.2e-16 apart at most
    dz = np.abs(sid - osid)
    assert dz.max() <= 4.2e-16, dz.max()
    # The weight formula (src/gaussmod.f90:37-38) is (1 - z^2) / ql(k-1, z)^2
    # evaluated at the root z, so it carries the abscissa's rounding with the
    # gain |d ln w / dz|, up to 2k / (1 - z^2): one ulp of z near the pole
    # moves the weight by 1e-12 (NLAT 32) to 3e-10 (NLAT 192).  Each GPU
    # weight is held to that first-order propagation of its own abscissa
    # difference, plus 2e-15 NLAT for the ulps of cos/acos inside ql (random
    # one-ulp perturbations of each move the weights by up to 1.7e-14 at NLAT
    # 32 and 1.5e-13 at NLAT 192).
    k = p.nlat
    x = np.polynomial.legendre.leggauss(k)[0][::-1]
    lnw = lambda z: np.log(1.0 - z * z) - 2.0 * np.log(  # noqa: E731
        np.abs(np.polynomial.legendre.legval(z, [0] * (k - 1) + [1])))
    h = 1e-7
    gain = np.abs((lnw(x + h) - lnw(x - h)) / (2 * h))
    bound = 2.0 * gain * dz + 2e-15 * k
    ew = np.abs(gwd - ogwd) / ogwd
    print(f"  max |sid| error {dz.max():.2e}, weights rel {ew.max():.2e} "
          f"(bound at that latitude {bound[np.argmax(ew)]:.2e})", flush=True)
    assert (ew <= bound).all(), (ew.max(), np.argmax(ew), bound[np.argmax(ew)])
    # and the quadrature they define integrates constants exactly
    assert abs(gwd.sum() - 2.0) < 1e-13


@pytest.mark.parametrize("ntru", [21, 42, 63, 127])
def test_tables(oracle, gpu, ntru):
    # tables built on the GPU from the model's own sid/gwd, as legini does
    nlat = sht.dims(ntru)[0]
    osid, ogwd = oracle.inigau(nlat)
    p = sht.Plan(ntru, sid=osid, gwd=ogwd)
    print(f"T{ntru} tables", flush=True)
    t = p.tables()
    p.close()
    ot = oracle.legini(ntru, nlat, osid, ogwd)["_all"]
    for i, name in enumerate(sht.TABLES):
        err = np.abs(t[i] - ot[i]).max() / np.abs(ot[i]).max()
        assert err < 1e-14, (name, err)


def _oracle_roundtrip(oracle, ntru, nlev, sp, sk=None):
    gps, sps = [], []
    tab = oracle.legini(ntru, sht.dims(ntru)[0])["_all"]
    for k in range(sp.shape[0]):
        g, s = oracle.roundtrip(ntru, nlev, sp[k], tables=tab, sk=sk)
        gps.append(g)
        sps.append(s)
    return np.stack(gps), np.stack(sps)


def _check_roundtrip(p, oracle, mode, ntru, nlev, nmem=2, cfg=2, slope=1.5, wave1=0.0,
                     refs=None, identity=True):
    import torch

    sp = ensemble(ntru, nlev, nmem, cfg=cfg, slope=slope, wave1=wave1)
    g_ref, s_ref = refs if refs is not None else _oracle_roundtrip(oracle, ntru, nlev, sp)
    sp_in, gp, sp_out = p.alloc_roundtrip(nmem, nlev, mode)
    sp_in.copy_(torch.from_numpy(sp).to(sp_in.dtype))
    gp.fill_(float("nan"))
    sp_out.fill_(float("nan"))
    p.roundtrip(mode, sp_in, gp, sp_out, nlev)
    torch.cuda.synchronize()
    g = gp.double().cpu().numpy()
    s = sp_out.double().cpu().numpy()
    tol = TOL[mode]
    nfl = 3 * nlev + 1
    names = {"vor": range(0, nlev), "div": range(nlev, 2 * nlev),
             "tmp": range(2 * nlev, 3 * nlev), "lnps": [nfl - 1], "dmu": [nfl],
             "u": range(nfl + 1, nfl + 1 + nlev), "v": range(nfl + 1 + nlev, nfl + 1 + 2 * nlev)}
    worst = {}
    for k in range(nmem):
        for name, idx in names.items():
            idx = list(idx)
            e = rel_l2(g[k, idx], g_ref[k, idx])
            worst["grid." + name] = max(worst.get("grid." + name, 0.0), e)
            assert e < tol, (k, name, "grid", e)
        for name in ("vor", "div", "tmp", "lnps"):
            idx = list(names[name])
            e = rel_l2(s[k, idx], s_ref[k, idx])
            worst["spectral." + name] = max(worst.get("spectral." + name, 0.0), e)
            assert e < tol, (k, name, "spectral", e)
    print("  per-field rel L2: " + ", ".join(f"{k} {v:.1e}" for k, v in worst.items()),
          flush=True)
    # the round trip reproduces its band-limited input (without filters)
    if identity:
        assert rel_l2(s, sp) < max(tol, 1e-12) * 10


@pytest.mark.parametrize("mode", ["f64", "f32", "f32mfma", "bf16x3"])
@pytest.mark.parametrize("ntru,nlev", TRUNCS)
def test_roundtrip_vs_oracle(plans, oracle, gpu, mode, ntru, nlev):
    print(f"roundtrip T{ntru} L{nlev} {mode}", flush=True)
    p = get_plan(plans, ntru)
    p.set_tiles(mode, 0, 0, 0)
    _check_roundtrip(p, oracle, mode, ntru, nlev)


@pytest.mark.parametrize("mode", ["bf16x3", "f32mfma"])
def test_roundtrip_t127_l40_tidal(plans, oracle, gpu, mode):
    # the T127 L40 workload: spectral slope (n+1)^-2 with a wave-1 day-night
    # pattern of amplitude 10 on Re(m=1, n=1) of every field-level; every
    # field of both members within 1e-6 rel L2 of the FP64 oracle, for the
    # default tiles and the three variants
    print(f"roundtrip T127 L40 slope 2 wave-1 {mode}", flush=True)
    p = get_plan(plans, 127)
    sp = ensemble(127, 40, 2, cfg=4, slope=2.0, wave1=10.0)
    refs = _oracle_roundtrip(oracle, 127, 40, sp)
    try:
        for var in (0, 1, 2, 3):
            p.set_tiles(mode, var, var, var)
            _check_roundtrip(p, oracle, mode, 127, 40, nmem=2, cfg=4, slope=2.0, wave1=10.0,
                             refs=refs)
    finally:
        p.set_tiles(mode, 0, 0, 0)


@pytest.mark.parametrize("nf", [1, 2, 3, 4])
@pytest.mark.parametrize("mode", ["f64", "bf16x3", "f32mfma"])
def test_roundtrip_physics_filters(oracle, gpu, nf, mode):
    # legini's physics filters (src/legmod.f90:123-155: Cesaro, exponential,
    # Lander-Hoskins, Riesz-2) on both directions, folded into the GPU
    # tables: the T42 L20 round trip against the oracle with the same filters
    ntru, nlev = 42, 20
    print(f"roundtrip T{ntru} L{nlev} nfilter={nf} {mode}", flush=True)
    g, s = oracle.filters(ntru, nf, 1, 1)
    p = sht.Plan(ntru, skgpsp=g, skspgp=s)
    try:
        sp = ensemble(ntru, nlev, 2, cfg=2)
        refs = _oracle_roundtrip(oracle, ntru, nlev, sp, sk=(g, s))
        _check_roundtrip(p, oracle, mode, ntru, nlev, refs=refs, identity=False)
    finally:
        p.close()


@pytest.mark.parametrize("nf", [1, 3])
def test_dropin_physics_filters(oracle, gpu, nf):
    # the drop-in entry points after legini with filters, FP64 (the model's
    # MOST_PREC=8 build): sp2fc, sp2fcdmu, dv2uv take skspgp; fc2sp, uv2dv,
    # mktend take skgpsp
    print(f"drop-in T21 nfilter={nf}", flush=True)
    ntru, nlev = 21, 10
    nlat, nlon, _ = sht.dims(ntru)
    g, s = oracle.filters(ntru, nf, 1, 1)
    sht.legini(ntru, nlat, nlev, skgpsp=g, skspgp=s)
    tab = oracle.legini(ntru, nlat)
    sp = ensemble(ntru, nlev, 1, cfg=8)[0]
    fc = sht.sp2fc(sp[:2])
    for f in range(2):
        assert rel_l2(fc[f], oracle.sp2fc(ntru, nlat, nlon, tab["qi"], s, sp[f])) < 1e-13
    assert rel_l2(sht.sp2fcdmu(sp[-1])[0],
                  oracle.sp2fc(ntru, nlat, nlon, tab["qj"], s, sp[-1])) < 1e-13
    back = sht.fc2sp(fc)
    for f in range(2):
        assert rel_l2(back[f], oracle.fc2sp(ntru, nlat, nlon, tab["qc"], g, fc[f])) < 1e-13
    pu, pv = sht.dv2uv(sp[nlev:2 * nlev], sp[:nlev])
    ou, ov = oracle.dv2uv(ntru, nlat, nlon, nlev, tab["qu"], tab["qv"], s, sht.PLAVOR,
                          sp[nlev:2 * nlev], sp[:nlev])
    assert rel_l2(pu, ou) < 1e-13 and rel_l2(pv, ov) < 1e-13
    rng = np.random.default_rng(nf)
    f6 = [rng.standard_normal((nlev, nlat, nlon)) for _ in range(6)]
    d, t, z = sht.mktend(*f6)
    o = oracle.mktend(ntru, nlat, nlon, nlev, tab["qc"], tab["qe"], tab["qm"], tab["qq"], g, *f6)
    for x, y in zip((d, t, z), o):
        assert rel_l2(x, y) < 1e-13
    pd, pz = sht.uv2dv(f6[0], f6[1])
    od, oz = oracle.uv2dv(ntru, nlat, nlon, nlev, tab["qe"], tab["qm"], g, f6[0], f6[1])
    assert rel_l2(pd, od) < 1e-13 and rel_l2(pz, oz) < 1e-13
    sht.legini(ntru, nlat, nlev)  # back to no filter


@pytest.mark.parametrize("mode", ["f64", "f32", "f32mfma", "bf16x3"])
@pytest.mark.parametrize("ntru,nlev", [(21, 10), (42, 20), (127, 4)])
@pytest.mark.parametrize("var", [1, 2])
def test_roundtrip_tile_variant(oracle, gpu, mode, ntru, nlev, var):
    # variant 1: half the field-levels per block; variant 2: 512-thread
    # blocks; all three fused kernels.  nlev=20 also leaves a partial last
    # block for the odd field counts
    print(f"roundtrip tiles={var} T{ntru} L{nlev} {mode}", flush=True)
    p = sht.Plan(ntru)
    try:
        p.set_tiles(mode, var, var, var)
        assert p.tiles(mode) == (var, var, var)
        _check_roundtrip(p, oracle, mode, ntru, nlev, nmem=3)
        # the single-kind inverse kernels (sp2fc, sp2fcdmu, dv2uv) with the
        # same tiles against the grid fields of the oracle's round trip
        import torch

        sp = ensemble(ntru, nlev, 2, cfg=6)
        g_ref, _ = _oracle_roundtrip(oracle, ntru, nlev, sp)
        sp_in, gp, _ = p.alloc_roundtrip(2, nlev, mode)
        sp_in.copy_(torch.from_numpy(sp).to(sp_in.dtype))
        gp.fill_(float("nan"))
        nfl = 3 * nlev + 1
        p.inverse(mode, sp_in, gp, nfl, 0, 0, "qi")
        p.inverse(mode, sp_in, gp, 1, nfl - 1, nfl, "qj")
        p.inverse_uv(mode, sp_in, gp, nlev, 0, nlev, nfl + 1, nfl + 1 + nlev)
        torch.cuda.synchronize()
        g = gp.double().cpu().numpy()
        for k in range(2):
            assert rel_l2(g[k], g_ref[k]) < TOL[mode], (k, rel_l2(g[k], g_ref[k]))
    finally:
        p.close()


@pytest.mark.parametrize("mode", ["f64", "bf16x3", "f32mfma"])
@pytest.mark.parametrize("ntru,nlev", [(21, 10), (42, 20), (127, 4)])
def test_inverse_all(plans, oracle, gpu, mode, ntru, nlev):
    # the one-launch inverse half of the round trip, both tile variants
    import torch

    print(f"inverse_all T{ntru} L{nlev} {mode}", flush=True)
    p = get_plan(plans, ntru)
    nmem = 3
    sp = ensemble(ntru, nlev, nmem, cfg=5)
    g_ref = np.stack([oracle.roundtrip(ntru, nlev, sp[k])[0] for k in range(nmem)])
    sp_in, gp, _ = p.alloc_roundtrip(nmem, nlev, mode)
    sp_in.copy_(torch.from_numpy(sp).to(sp_in.dtype))
    for var in (0, 1, 2, 3):
        p.set_tiles(mode, var, 0, 0)
        gp.fill_(float("nan"))
        p.inverse_all(mode, sp_in, gp, nlev)
        torch.cuda.synchronize()
        g = gp.double().cpu().numpy()
        for k in range(nmem):
            e = rel_l2(g[k], g_ref[k])
            assert e < TOL[mode], (var, k, e)
    p.set_tiles(mode, 0, 0, 0)


def test_autotune(oracle, gpu):
    print("autotune T42 L20 bf16x3", flush=True)
    p = sht.Plan(42)
    try:
        v = p.autotune("bf16x3", 16, 20)
        assert all(x in (0, 1, 2, 3) for x in v) and p.tiles("bf16x3") == v
        _check_roundtrip(p, oracle, "bf16x3", 42, 20)
        with pytest.raises(sht.ExoError):
            p.set_tiles("bf16x3", 4, 0, 0)
    finally:
        p.close()


def test_dropin_f64_t21(oracle, gpu):
    print("drop-in T21 f64", flush=True)
    ntru, nlev = 21, 10
    nlat, nlon, ncsp = sht.dims(ntru)
    sht.legini(ntru, nlat, nlev)
    sid, gwd = oracle.inigau(nlat)
    tab = oracle.legini(ntru, nlat, sid, gwd)
    one = np.ones(ntru + 1)
    rng = np.random.default_rng(7)
    sp = ensemble(ntru, nlev, 1, cfg=1)[0]
    fc = sht.sp2fc(sp[:3])
    for f in range(3):
        ref = oracle.sp2fc(ntru, nlat, nlon, tab["qi"], one, sp[f])
        assert rel_l2(fc[f], ref) < 1e-13
    fcd = sht.sp2fcdmu(sp[-1])
    ref = oracle.sp2fc(ntru, nlat, nlon, tab["qj"], one, sp[-1])
    assert rel_l2(fcd[0], ref) < 1e-13
    back = sht.fc2sp(fc)
    for f in range(3):
        ref = oracle.fc2sp(ntru, nlat, nlon, tab["qc"], one, fc[f])
        assert rel_l2(back[f], ref) < 1e-13
    pd = sp[nlev:2 * nlev]
    pz = sp[:nlev].copy()
    pz_before = pz.copy()
    pu, pv = sht.dv2uv(pd, pz)
    ou, ov = oracle.dv2uv(ntru, nlat, nlon, nlev, tab["qu"], tab["qv"], one, sht.PLAVOR, pd, pz)
    assert np.array_equal(pz, pz_before)
    assert rel_l2(pu, ou) < 1e-13 and rel_l2(pv, ov) < 1e-13
    rows = rng.standard_normal((2 * nlat + 1, nlon))
    f_gpu = sht.gp2fc(rows)
    f_ref = oracle.gp2fc(rows)
    assert rel_l2(f_gpu, f_ref) < 1e-14
    g_gpu = sht.fc2gp(f_ref)
    g_ref = oracle.fc2gp(f_ref)
    assert rel_l2(g_gpu, g_ref) < 1e-14


def test_dropin_f32_t42(oracle, gpu):
    print("drop-in T42 f32", flush=True)
    ntru, nlev = 42, 20
    nlat, nlon, ncsp = sht.dims(ntru)
    sht.legini(ntru, nlat, nlev)
    tab = oracle.legini(ntru, nlat)
    one = np.ones(ntru + 1)
    sp = ensemble(ntru, nlev, 1, cfg=2)[0]
    fc = sht.sp2fc(sp[:4].astype(np.float32))
    assert fc.dtype == np.float32
    for f in range(4):
        ref = oracle.sp2fc(ntru, nlat, nlon, tab["qi"], one, sp[f])
        assert rel_l2(fc[f], ref) < 1e-6
    back = sht.fc2sp(fc)
    assert rel_l2(back, sp[:4]) < 2e-6
    g = sht.fc2gp(fc)
    f2 = sht.gp2fc(g)
    ref = oracle.gp2fc(oracle.fc2gp(fc.astype(np.float64)))
    assert rel_l2(f2, ref) < 1e-6


def test_error_paths(gpu):
    # unsupported truncation and calls before legini map to error codes
    with pytest.raises(sht.ExoError):
        sht.Plan(30)
    with pytest.raises(sht.ExoError):
        sht.legini(42, 50, 10)


@pytest.mark.parametrize("ntru,nlev,dtype,tol", [(21, 10, np.float64, 1e-13),
                                                 (42, 20, np.float32, 2e-6),
                                                 (63, 3, np.float64, 1e-13),
                                                 (127, 2, np.float32, 2e-6)])
def test_dropin_multi_direct(oracle, gpu, ntru, nlev, dtype, tol):
    print(f"uv2dv/qtend/mktend T{ntru} {np.dtype(dtype).name}", flush=True)
    nlat, nlon, ncsp = sht.dims(ntru)
    sht.legini(ntru, nlat, nlev)
    tab = oracle.legini(ntru, nlat)
    one = np.ones(ntru + 1)
    rng = np.random.default_rng(ntru)
    f = [rng.standard_normal((nlev, nlat, nlon)) for _ in range(6)]
    fd = [x.astype(dtype) for x in f]
    pd, pz = sht.uv2dv(fd[0], fd[1])
    od, oz = oracle.uv2dv(ntru, nlat, nlon, nlev, tab["qe"], tab["qm"], one, f[0], f[1])
    assert rel_l2(pd, od) < tol and rel_l2(pz, oz) < tol
    q = sht.qtend(fd[0], fd[1], fd[2])
    oq = oracle.qtend(ntru, nlat, nlon, nlev, tab["qc"], tab["qe"], tab["qm"], one, f[0], f[1], f[2])
    assert rel_l2(q, oq) < tol
    d, t, z = sht.mktend(*fd)
    o = oracle.mktend(ntru, nlat, nlon, nlev, tab["qc"], tab["qe"], tab["qm"], tab["qq"], one, *f)
    for x, y in zip((d, t, z), o):
        assert rel_l2(x, y) < tol
