"""Pins of the oracle functions that the round-trip goldens do not reach --
CPU only: the physics filters of legini and their application in the
transforms (src/legmod.f90:123-155), the semi-implicit matrices and solve
(initsi, makebm, minvers, spectrala sections 1-3; src/plasim.f90:1683-1888,
2817-2876) and the grid-point tendencies calcgp (src/plasim.f90:2587-2750).

(a) golden vectors written by tests/golden/make_golden.py from the
reference's own plasim.f90 and legmod.f90, compiled as they lie
(oracle/Makefile, oracle/ref_glue_*.f90); (b) the compiled reference itself
when oracle/_ref/ is built in this tree (bitwise where the oracle follows
the reference's association order); (c) properties the mathematics fixes.
"""
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
from make_golden import (CALCGP_KEEP, FILTER_PARAMS, PLASIM_DELT, PLASIM_SCAL,  # noqa: E402
                         golden_input, plasim_inputs, plasim_vertical, spectral_input)

GOLD_F = os.path.join(HERE, "golden", "reference_filters_t21.npz")
GOLD_P = os.path.join(HERE, "golden", "reference_plasim_t21.npz")
NTRU, NLAT, NLON, NLEV, NHOR = 21, 32, 64, 10, 2048


def close(a, b, rtol):
    a, b = np.asarray(a), np.asarray(b)
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300) < rtol


def filters(oracle, nf):
    return oracle.filters(NTRU, nf, 1, 1, FILTER_PARAMS["kappa"], FILTER_PARAMS["nexp"],
                          FILTER_PARAMS["n0"])


# ---------------------------------------------------------------------------
# physics filters
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("nf", [1, 2, 3, 4])
def test_golden_filters(oracle, nf):
    gold = np.load(GOLD_F)
    g, s = filters(oracle, nf)
    assert close(g, gold[f"f{nf}_skgpsp"], 1e-15) and close(s, gold[f"f{nf}_skspgp"], 1e-15)
    tab = oracle.legini(NTRU, NLAT)
    sp = spectral_input(NTRU, 3 * NLEV + 1, 0.3)
    # sp -> gp transforms take skspgp, gp -> sp transforms skgpsp
    fc = oracle.sp2fc(NTRU, NLAT, NLON, tab["qi"], s, sp[0])
    assert close(fc, gold[f"f{nf}_sp2fc"], 1e-15)
    assert close(oracle.sp2fc(NTRU, NLAT, NLON, tab["qj"], s, sp[-1]),
                 gold[f"f{nf}_sp2fcdmu"], 1e-15)
    assert close(oracle.fc2sp(NTRU, NLAT, NLON, tab["qc"], g, fc), gold[f"f{nf}_fc2sp"], 1e-14)
    pu, pv = oracle.dv2uv(NTRU, NLAT, NLON, NLEV, tab["qu"], tab["qv"], s, 1.63299310207,
                          sp[NLEV:2 * NLEV], sp[:NLEV])
    assert close(pu[0], gold[f"f{nf}_dv2uv_u0"], 1e-15)
    assert close(pv[0], gold[f"f{nf}_dv2uv_v0"], 1e-15)
    f = [golden_input((NLEV, NLAT, NLON), 2.1 + 0.5 * k) for k in range(6)]
    d, t, z = oracle.mktend(NTRU, NLAT, NLON, NLEV, tab["qc"], tab["qe"], tab["qm"],
                            tab["qq"], g, *f)
    assert close(d[0], gold[f"f{nf}_mktend_d0"], 1e-14)
    assert close(z[0], gold[f"f{nf}_mktend_z0"], 1e-14)


def test_filters_closed_forms(oracle):
    # the four filters written out with the reference's 1-based n
    # (filter value of total wavenumber n0 = 0..NTRU uses n = n0 + 1)
    n = np.arange(1, NTRU + 2, dtype=np.float64)
    want = {1: 1.0 - n / (NTRU + 1),
            2: np.exp(-8.0 * (n / NTRU) ** 8),
            3: np.exp(-(n * (n + 1) / (15.0 * 16.0)) ** 2),
            4: (1.0 - n / (NTRU + 1)) ** 2}
    for nf, w in want.items():
        g, s = filters(oracle, nf)
        assert close(g, w, 1e-14) and close(s, w, 1e-14), nf
        # one direction only
        g1, s1 = oracle.filters(NTRU, nf, 1, 0)
        assert np.all(s1 == 1.0) and close(g1, w, 1e-14)
    g0, s0 = oracle.filters(NTRU, 0, 1, 1)
    assert np.all(g0 == 1.0) and np.all(s0 == 1.0)


def test_filter_application_is_per_wavenumber(oracle):
    # a filter multiplies each spectral coefficient of total wavenumber n by
    # its factor: sp2fc with the filter equals sp2fc of the scaled input
    tab = oracle.legini(NTRU, NLAT)
    sp = spectral_input(NTRU, 1, 0.6)[0]
    ni = np.concatenate([[nn, nn] for m in range(NTRU + 1) for nn in range(m, NTRU + 1)])
    _, s = filters(oracle, 3)
    a = oracle.sp2fc(NTRU, NLAT, NLON, tab["qi"], s, sp)
    b = oracle.sp2fc(NTRU, NLAT, NLON, tab["qi"], np.ones(NTRU + 1), sp * s[ni])
    assert close(a, b, 1e-14)


def test_live_reference_filters(oracle):
    from oracle import RefT21, ref_available

    if not ref_available():
        pytest.skip("oracle/_ref not built (needs the reference tree)")
    from ctypes import c_double, c_int

    from oracle import _p

    r = RefT21()
    tab = oracle.legini(NTRU, NLAT, r.sid, r.gwd)
    rng = np.random.default_rng(31)
    sp = rng.standard_normal(2 * r.ncsp)
    try:
        for nf in (1, 2, 3, 4):
            g = np.zeros(NTRU + 1)
            s = np.zeros(NTRU + 1)
            r.lib.ref_t21_set_filter(c_int(nf), c_int(1), c_int(1), c_double(8.0), c_int(8),
                                     c_double(15.0), _p(g), _p(s))
            og, os_ = filters(oracle, nf)
            assert close(og, g, 1e-15) and close(os_, s, 1e-15)
            fc = oracle.sp2fc(NTRU, NLAT, NLON, tab["qi"], s, sp)
            assert close(fc, r.sp2fc(sp), 1e-15)
            assert close(oracle.fc2sp(NTRU, NLAT, NLON, tab["qc"], g, fc), r.fc2sp(fc), 1e-15)
    finally:
        z = np.zeros(NTRU + 1)
        r.lib.ref_t21_set_filter(c_int(0), c_int(0), c_int(0), c_double(8.0), c_int(8),
                                 c_double(15.0), _p(z), _p(z.copy()))


# ---------------------------------------------------------------------------
# semi-implicit matrices and solve, calcgp
# ---------------------------------------------------------------------------
@pytest.fixture(scope="module")
def plasim_gold():
    return np.load(GOLD_P)


@pytest.fixture(scope="module")
def vert(oracle):
    return plasim_vertical(oracle, NLEV)


def test_golden_initsi(vert, plasim_gold):
    for k in ("g", "tau", "c", "t01s2", "tkp"):
        assert close(vert[k], plasim_gold["initsi_" + k], 1e-15), k


def test_golden_makebm_minvers(oracle, vert, plasim_gold):
    bm1 = oracle.makebm(NLEV, NTRU, PLASIM_DELT, vert["t0"], vert["dsigma"], vert["g"],
                        vert["tau"])
    assert close(bm1, plasim_gold["bm1"], 1e-14)
    a = golden_input((12, 12), 0.77) + 4.0 * np.eye(12)
    assert close(oracle.minvers(a), plasim_gold["minvers_12"], 1e-15)


def test_golden_si_apply(oracle, vert, plasim_gold):
    si, _ = plasim_inputs(NLEV, NHOR, NTRU)
    bm1 = oracle.makebm(NLEV, NTRU, PLASIM_DELT, vert["t0"], vert["dsigma"], vert["g"],
                        vert["tau"])
    stt, sdt, spt = oracle.si_apply(NTRU, NLEV, PLASIM_DELT, vert["t0"], vert["dsigma"],
                                    vert["g"], vert["tau"], bm1, si["sop"], si["spm"],
                                    si["atm"], si["adm"], si["stt"], si["sdt"], si["spt"])
    assert close(stt, plasim_gold["si_stt"], 1e-14)
    assert close(sdt, plasim_gold["si_sdt"], 1e-14)
    assert close(spt, plasim_gold["si_spt"], 1e-14)


@pytest.mark.parametrize("nq", [1, 0])
def test_golden_calcgp(oracle, vert, plasim_gold, nq):
    _, f = plasim_inputs(NLEV, NHOR, NTRU)
    out = oracle.calcgp(f, vert, PLASIM_SCAL, nq)
    for k, x in out.items():
        assert close(x[..., ::CALCGP_KEEP], plasim_gold[f"calcgp{nq}_{k}"], 1e-14), k


def test_calcgp_properties(oracle, vert):
    """What the equations fix without the reference: with no divergence,
    no wind and no ln ps gradient there is no vertical motion and every
    tendency vanishes; the geopotential term is g^T (T_v - T) * 2 / rcsq;
    the humidity tendency is linear in the divergence and conserves the
    column total."""
    rng = np.random.default_rng(4)
    nh = 64
    z = lambda: np.zeros((NLEV, nh))  # noqa: E731
    f = {"gu": z(), "gv": z(), "gd": z(), "gz": rng.standard_normal((NLEV, nh)),
         "gt": rng.standard_normal((NLEV, nh)), "gq": 1e-3 * rng.random((NLEV, nh)),
         "dq": z(), "gp": np.ones(nh), "gpj": np.zeros(nh), "gpm": np.zeros(nh),
         "rcsq": 1.0 + rng.random(nh)}
    out = oracle.calcgp(f, vert, PLASIM_SCAL, 1)
    for k in ("gtn", "gqn", "guz", "gvz", "dw"):
        assert np.all(out[k] == 0.0), k
    assert np.all(out["gvp"] == 0.0)
    rdbrv = PLASIM_SCAL[0]
    zq = np.maximum(f["gq"] / f["gp"], 0.0)
    tv1 = (f["gt"] + vert["t0"][:, None]) * (1.0 + (1.0 / rdbrv - 1.0) * zq) - vert["t0"][:, None]
    gphi = (vert["g"].T @ (tv1 - f["gt"])) * 2.0 / f["rcsq"]
    assert close(out["gphi"], gphi, 1e-13)
    # gqn is bilinear in (zsdotp, gq): doubling the divergence doubles it
    f2 = dict(f, gd=rng.standard_normal((NLEV, nh)))
    a = oracle.calcgp(f2, vert, PLASIM_SCAL, 1)["gqn"]
    b = oracle.calcgp(dict(f2, gd=2.0 * f2["gd"]), vert, PLASIM_SCAL, 1)["gqn"]
    assert close(b, 2.0 * a, 1e-13)
    # the vertical advection of humidity is in flux form: its column
    # integral, sum_j gqn_j / rdsig_j, telescopes to zero
    col = (a / vert["rdsig"][:, None]).sum(0)
    assert np.abs(col).max() < 1e-12 * np.abs(a / vert["rdsig"][:, None]).max()


def test_initsi_properties(vert):
    # g is lower triangular with g(1,1) = 1; c(j, k) = g(k, j) dsigma(j) /
    # dsigma(k); tkp = kappa t0; t01s2 are the t0 differences
    g, c = vert["g"], vert["c"]
    assert np.all(np.triu(g, 1) == 0.0) and g[0, 0] == 1.0
    ds = vert["dsigma"]
    assert close(c, g.T * ds[:, None] / ds[None, :], 1e-15)
    assert close(vert["tkp"], 0.286 * vert["t0"], 1e-15)
    assert close(vert["t01s2"][:-1], np.diff(vert["t0"]), 1e-15) and vert["t01s2"][-1] == 0.0


def test_live_reference_plasim(oracle, vert):
    from oracle import RefPlasim, ref_available

    if not ref_available("libref_plasim.so"):
        pytest.skip("oracle/_ref not built (needs the reference tree)")
    r = RefPlasim()
    rv = r.initsi(vert["sigma"], vert["sigmah"], vert["dsigma"], vert["rdsig"], vert["t0"])
    for k in ("g", "tau", "c", "t01s2", "tkp"):
        assert np.array_equal(rv[k], vert[k]), k
    rng = np.random.default_rng(12)
    delt = 0.03
    b_o = oracle.makebm(NLEV, NTRU, delt, vert["t0"], vert["dsigma"], vert["g"], vert["tau"])
    b_r = r.makebm(delt, vert["t0"], vert["dsigma"], vert["g"], vert["tau"])
    assert np.array_equal(b_o, b_r)
    a = rng.standard_normal((9, 9)) + 3 * np.eye(9)
    assert np.array_equal(oracle.minvers(a), r.minvers(a))
    nsp = r.nspp
    arr = lambda *s: rng.standard_normal(s)  # noqa: E731
    args = (arr(nsp), arr(nsp), arr(NLEV, nsp), arr(NLEV, nsp), arr(NLEV, nsp),
            arr(NLEV, nsp), arr(nsp))
    o = oracle.si_apply(NTRU, NLEV, delt, vert["t0"], vert["dsigma"], vert["g"], vert["tau"],
                        b_o, *args)
    rr = r.si_apply(delt, vert["t0"], vert["dsigma"], vert["g"], vert["tau"], b_o, *args)
    for x, y in zip(o, rr):
        assert np.array_equal(x, y)
    f = {k: 0.1 * rng.standard_normal((NLEV, NHOR)) for k in ("gu", "gv", "gt", "gd", "gz")}
    f["gq"] = 1e-3 * rng.standard_normal((NLEV, NHOR))
    f["dq"] = 1e-3 * rng.standard_normal((NLEV, NHOR))
    f["gp"] = np.exp(0.05 * rng.standard_normal(NHOR))
    f["gpj"], f["gpm"] = 0.01 * rng.standard_normal(NHOR), 0.01 * rng.standard_normal(NHOR)
    f["rcsq"] = 1.0 / (1.0 - rng.uniform(0.0, 0.99, NHOR))
    for nq in (1, 0):
        x, y = oracle.calcgp(f, vert, PLASIM_SCAL, nq), r.calcgp(f, vert, PLASIM_SCAL, nq)
        for k in x:
            assert np.array_equal(x[k], y[k]), (nq, k)
