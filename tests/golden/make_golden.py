"""Write the golden vectors of the spectral-transform path from the
reference's own Fortran, compiled by oracle/Makefile into oracle/_ref/.

Run in a tree where oracle/_ref/ is built (needs the reference sources):
    make -C oracle ref && python tests/golden/make_golden.py
It writes tests/golden/*.npz: inputs are closed-form deterministic
functions (golden_input) so only outputs are stored; arrays are subsets
where the full output would be large.
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def golden_input(shape, salt: float) -> np.ndarray:
    """Deterministic test input without a random generator:
    x_i = sin(0.7 i + salt) * cos(0.013 i^1.5 + 2 salt)."""
    i = np.arange(int(np.prod(shape)), dtype=np.float64)
    return (np.sin(0.7 * i + salt) * np.cos(0.013 * i ** 1.5 + 2.0 * salt)).reshape(shape)


def spectral_input(ntru, nfl, salt):
    """golden_input shaped as spectral field-levels, Im(m=0) zeroed and
    damped with (n+1)^-1.5."""
    m = np.concatenate([np.full(ntru + 1 - mm, mm) for mm in range(ntru + 1)])
    n = np.concatenate([np.arange(mm, ntru + 1) for mm in range(ntru + 1)])
    x = golden_input((nfl, m.size, 2), salt)
    x *= ((n + 1.0) ** -1.5)[None, :, None]
    x[:, m == 0, 1] = 0.0
    return x.reshape(nfl, 2 * m.size)


LAT_SUBSET = [0, 5, 15, 16, 31]  # rows of the T21 tables kept (N and S)


def main():
    from oracle import RefFft991, RefPyfft, RefT21

    r = RefT21()
    out = {}
    out["sid"], out["gwd"] = r.sid, r.gwd
    out["tables"] = r.tables[:, LAT_SUBSET, :]
    out["plavor"] = np.array([r.plavor])
    nl = r.nlev
    sp = spectral_input(21, 3 * nl + 1, 0.3)
    out["sp2fc_f0"] = r.sp2fc(sp[0])
    out["sp2fcdmu_lnps"] = r.sp2fcdmu(sp[-1])
    out["fc2sp_f0"] = r.fc2sp(out["sp2fc_f0"])
    pu, pv = r.dv2uv(sp[nl:2 * nl], sp[:nl])
    out["dv2uv_u01"], out["dv2uv_v01"] = pu[:2], pv[:2]
    f = [golden_input((nl, r.nlat, r.nlon), 1.1 + 0.5 * k) for k in range(6)]
    pd, pz = r.uv2dv(f[0], f[1])
    out["uv2dv_d01"], out["uv2dv_z01"] = pd[:2], pz[:2]
    out["qtend_01"] = r.qtend(f[0], f[1], f[2])[:2]
    d, t, z = r.mktend(*f)
    out["mktend_d01"], out["mktend_t01"], out["mktend_z01"] = d[:2], t[:2], z[:2]
    for n in (64, 128, 384):
        g = golden_input((3, n), 0.1 * n)
        out[f"gp2fc_{n}"] = r.gp2fc(g)
        out[f"fc2gp_{n}"] = r.fc2gp(g)
    f9 = RefFft991()
    for n in (192, 128):
        g = golden_input((3, n), 0.2 * n)
        out[f"fft991_gp2fc_{n}"] = f9.gp2fc(g)
        out[f"fft991_fc2gp_{n}"] = f9.fc2gp(g)
    p = RefPyfft()
    tab42 = p.legini(42, 64)
    out["t42_tables_lat0_7"] = tab42[:, [0, 7], :]
    sp42 = spectral_input(42, 1, 0.9)[0]
    fc42 = p.sp2fc(42, 64, sp42)
    out["t42_sp2fc"] = fc42
    out["t42_fc2sp"] = p.fc2sp(42, 64, fc42)
    path = os.path.join(HERE, "reference_t21_t42.npz")
    np.savez_compressed(path, **out)
    print(path, os.path.getsize(path), "bytes")
    filter_golden(r, sp, nl)
    plasim_golden()
    lwr_golden()
    pyfft_golden()
    gridpointa_golden()


GPA_ROWS = [0, 4, 9]  # levels of each tendency field kept in the golden file


def gridpointa_state(ntru, nlev, salt=0.9):
    """the spectral state sz sd st sq sp of the gridpointa pins: small
    deterministic fields on the planetary vorticity, as a model state"""
    nfl = 4 * nlev + 1
    x = 0.01 * spectral_input(ntru, nfl, salt)
    x[:nlev, 2] += 1.63299310207  # plavor at Re(m=0, n=1) of the absolute vorticity
    x[3 * nlev:4 * nlev] *= 0.1   # humidity
    return x


def gridpointa_humidity(nlev, nhor):
    """grid-point gq and dq of the nqspec = 0 pin"""
    return (1e-3 * np.abs(golden_input((nlev, nhor), 5.3)),
            1e-3 * golden_input((nlev, nhor), 5.9))


def gridpointa_golden():
    """tendencies of the reference's gridpointa (T21 L10) for nqspec 1 and 0"""
    from oracle import Oracle, RefGridpointa

    r = RefGridpointa()
    n = r.nlev
    v = plasim_vertical(Oracle(), n)
    st = gridpointa_state(21, n)
    gq, dq = gridpointa_humidity(n, r.nlat * r.nlon)
    out = {}
    for nq in (1, 0):
        tend, grid = r.run(st, v, PLASIM_SCAL, nq, gq=gq, dq=dq)
        rows = [f * n + j for f in range(4) for j in GPA_ROWS] + [4 * n]
        out[f"tend{nq}"] = tend[rows]
        out[f"gp{nq}"] = grid["gp"]
    path = os.path.join(HERE, "reference_gridpointa_t21.npz")
    np.savez_compressed(path, **out)
    print(path, os.path.getsize(path), "bytes")


def pyfft_inputs(ntru, nlat, nlev):
    """inputs of the post-processor wrapper pins: spectral levels, grid
    levels and the 1/cos and cos latitude factors pyburn passes
    (exoplasim/pyburn.py:1040-1067, 1150-1172 use radius / cos and cos /
    radius; unit radius here)"""
    from oracle import Oracle

    sid, _ = Oracle().inigau(nlat)
    coslat = np.sqrt(1.0 - sid * sid)
    sd = spectral_input(ntru, nlev, 1.7)
    sz = spectral_input(ntru, nlev, 2.3)
    gu = golden_input((nlev, nlat, 2 * nlat), 3.1)
    gv = golden_input((nlev, nlat, 2 * nlat), 3.7)
    return sd, sz, gu, gv, 1.0 / coslat, coslat


def pyfft_golden():
    """sp2gp, gp2sp, sp3fc, gp3fc, fc3gp, spvgp, gpvsp of exoplasim/pyfft.f90
    in its default-real build (the post-processor's own) at T21, one level,
    with and without the physics filter; sp2gp and gp2sp also from the
    -fdefault-real-8 build with the filter."""
    from oracle import RefPyfft

    out = {}
    ntru, nlat, nlev = 21, 32, 1
    sd, sz, gu, gv, rd, cd = pyfft_inputs(ntru, nlat, nlev)
    for kind in ("r4", "r8"):
        p = RefPyfft(kind)
        for nfs in ((0, 1) if kind == "r4" else (1,)):
            k = f"{kind}_nfs{nfs}_"
            out[k + "sp2gp"] = p.sp2gp(ntru, nlat, sd, nfs)
            out[k + "gp2sp"] = p.gp2sp(ntru, gu, nfs)
            if kind == "r8":
                continue
            out[k + "sp3fc"] = p.sp3fc(ntru, nlat, sd, nfs)
            out[k + "spvgp_u"], out[k + "spvgp_v"] = p.spvgp(ntru, sd, sz, rd, nfs)
            out[k + "gpvsp_d"], out[k + "gpvsp_z"] = p.gpvsp(ntru, gu, gv, cd, nfs)
        if kind == "r4":
            out["r4_gp3fc"] = p.gp3fc(gu)
            out["r4_fc3gp"] = p.fc3gp(gu)
            for t, nl in ((21, 32), (42, 64), (63, 96), (127, 192)):
                out[f"r4_sfilt_t{t}"] = p.sfilt(t, nl, 1)
    path = os.path.join(HERE, "reference_pyfft_t21.npz")
    np.savez_compressed(path, **out)
    print(path, os.path.getsize(path), "bytes")


# physics filters of legini (src/legmod.f90:123-155): nfilter 1..4 with both
# directions switched on and the namelist defaults of the other parameters
FILTER_PARAMS = dict(kappa=8.0, nexp=8, n0=15.0)


def filter_golden(r, sp, nl):
    from ctypes import c_double, c_int

    from oracle import _p

    out = {}
    for nf in (1, 2, 3, 4):
        g = np.zeros(r.ntru + 1)
        s = np.zeros(r.ntru + 1)
        r.lib.ref_t21_set_filter(c_int(nf), c_int(1), c_int(1), c_double(FILTER_PARAMS["kappa"]),
                                 c_int(FILTER_PARAMS["nexp"]), c_double(FILTER_PARAMS["n0"]),
                                 _p(g), _p(s))
        out[f"f{nf}_skgpsp"], out[f"f{nf}_skspgp"] = g, s
        fc = r.sp2fc(sp[0])
        out[f"f{nf}_sp2fc"] = fc
        out[f"f{nf}_sp2fcdmu"] = r.sp2fcdmu(sp[-1])
        out[f"f{nf}_fc2sp"] = r.fc2sp(fc)
        pu, pv = r.dv2uv(sp[nl:2 * nl], sp[:nl])
        out[f"f{nf}_dv2uv_u0"], out[f"f{nf}_dv2uv_v0"] = pu[0], pv[0]
        f = [golden_input((nl, r.nlat, r.nlon), 2.1 + 0.5 * k) for k in range(6)]
        d, t, z = r.mktend(*f)
        out[f"f{nf}_mktend_d0"], out[f"f{nf}_mktend_z0"] = d[0], z[0]
    # back to the default (no filter) for later users of this library
    g = np.zeros(r.ntru + 1)
    r.lib.ref_t21_set_filter(c_int(0), c_int(0), c_int(0), c_double(8.0), c_int(8),
                             c_double(15.0), _p(g), _p(g.copy()))
    path = os.path.join(HERE, "reference_filters_t21.npz")
    np.savez_compressed(path, **out)
    print(path, os.path.getsize(path), "bytes")


# vertical set-up of the dynamics pins: a t0 profile from 220 K at the top
# to 288 K at the bottom on the neqsig = 0 levels of the T21 L10 model
def plasim_vertical(ora, nlev):
    return ora.vertical(nlev, t0=np.linspace(220.0, 288.0, nlev))


PLASIM_DELT = 0.05
# (rdbrv, adv, psurf, ww) of the Earth set-up (src/p_earth.f90:41-58,
# src/plasim.f90:1317-1318)
PLASIM_SCAL = (287.0 / 461.5, 1870.0 / (287.0 / 0.286) - 1.0, 101100.0, 7.292e-5)


def plasim_inputs(nlev, nhor, ntru):
    """deterministic inputs of the semi-implicit and calcgp pins"""
    nsp = (ntru + 1) * (ntru + 2)
    si = {k: golden_input(shape, 0.37 * i + 0.1) for i, (k, shape) in enumerate(
        (("sop", (nsp,)), ("spm", (nsp,)), ("atm", (nlev, nsp)), ("adm", (nlev, nsp)),
         ("stt", (nlev, nsp)), ("sdt", (nlev, nsp)), ("spt", (nsp,))))}
    gi = lambda shape, salt: golden_input(shape, salt)  # noqa: E731
    f = {k: 0.1 * gi((nlev, nhor), 0.5 + 0.3 * i) for i, k in
         enumerate(("gu", "gv", "gt", "gd", "gz"))}
    f["gq"] = 1e-3 * np.abs(gi((nlev, nhor), 2.9))
    f["dq"] = 1e-3 * gi((nlev, nhor), 3.3)
    f["gp"] = np.exp(0.05 * gi((nhor,), 3.7))
    f["gpj"] = 0.01 * gi((nhor,), 4.1)
    f["gpm"] = 0.01 * gi((nhor,), 4.5)
    f["rcsq"] = 1.0 / (1.0 - 0.98 * gi((nhor,), 4.9) ** 2)
    return si, f


CALCGP_KEEP = 16  # every 16th grid point of the calcgp outputs is stored


def plasim_golden():
    from oracle import Oracle, RefPlasim

    r = RefPlasim()
    ora = Oracle()
    v = plasim_vertical(ora, r.nlev)
    out = {}
    rv = r.initsi(v["sigma"], v["sigmah"], v["dsigma"], v["rdsig"], v["t0"], v["akap"])
    for k in ("g", "tau", "c", "t01s2", "tkp"):
        out["initsi_" + k] = rv[k]
    bm1 = r.makebm(PLASIM_DELT, v["t0"], v["dsigma"], rv["g"], rv["tau"])
    out["bm1"] = bm1
    a = golden_input((12, 12), 0.77) + 4.0 * np.eye(12)
    out["minvers_12"] = r.minvers(a)
    si, f = plasim_inputs(r.nlev, r.nhor, r.ntru)
    stt, sdt, spt = r.si_apply(PLASIM_DELT, v["t0"], v["dsigma"], rv["g"], rv["tau"], bm1,
                               si["sop"], si["spm"], si["atm"], si["adm"], si["stt"],
                               si["sdt"], si["spt"])
    out["si_stt"], out["si_sdt"], out["si_spt"] = stt, sdt, spt
    for nq in (1, 0):
        o = r.calcgp(f, {**v, **rv}, PLASIM_SCAL, nq)
        for k, x in o.items():
            out[f"calcgp{nq}_{k}"] = x[..., ::CALCGP_KEEP]
    path = os.path.join(HERE, "reference_plasim_t21.npz")
    np.savez_compressed(path, **out)
    print(path, os.path.getsize(path), "bytes")


# parameter sets of the long-wave fixture: namelist defaults, then grey
# clouds without pressure broadening
LWR_PARAMS = [{}, {"clgray": 0.7, "npbroaden": 0, "tpofmt": 0.8, "th2oc": 0.03}]
LWR_KEEP = 96  # columns stored per parameter set


def lwr_columns(ncol, nlev):
    """synthetic columns covering every branch of lwr: clear and cloudy
    (random cloud cover and liquid water), ozone, land and ocean"""
    sys.path.insert(0, ROOT)
    from exoplasim_amd.columns import make_columns

    c = make_columns(ncol, nlev, seed=77)
    rng = np.random.default_rng(78)
    half = ncol // 2
    c["dcc"][half:] = rng.uniform(0.0, 1.0, (ncol - half, nlev)) * (rng.uniform(size=(ncol - half, nlev)) < 0.6)
    c["dql"][half:] = rng.uniform(0.0, 3e-4, (ncol - half, nlev)) * (c["dcc"][half:] > 0)
    c["dqo3"][::3] = rng.uniform(0.0, 1e-5, (len(c["dqo3"][::3]), nlev))
    c["dls"][1::2] = 1.0
    return c


def lwr_golden():
    from oracle import RefRad

    r = RefRad()
    cols = lwr_columns(r.nhor, r.nlev)
    keep = np.linspace(0, r.nhor - 1, LWR_KEEP).round().astype(int)
    out = {"keep": keep}
    for k, v in cols.items():
        out["in_" + k] = v[keep]
    for i, prm in enumerate(LWR_PARAMS):
        fu, fd = r.lwr(cols, **prm)
        out[f"ftu_{i}"], out[f"ftd_{i}"] = fu[keep], fd[keep]
    path = os.path.join(HERE, "reference_lwr.npz")
    np.savez_compressed(path, **out)
    print(path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    if sys.argv[1:] == ["lwr"]:
        lwr_golden()
    elif sys.argv[1:] == ["plasim"]:
        plasim_golden()
    elif sys.argv[1:] == ["pyfft"]:
        pyfft_golden()
    elif sys.argv[1:] == ["gridpointa"]:
        gridpointa_golden()
    else:
        main()
