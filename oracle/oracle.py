"""Python access to the oracle -- TEST INFRASTRUCTURE ONLY.

Loads ``liboracle.so`` (the FP64 C restatement of the reference's spectral
transform core, oracle/exo_oracle.c) and, when they were built in this tree,
the reference's own Fortran compiled into ``oracle/_ref/``.  Only tests/,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of bench.py import
this module; the product package never does.

Array conventions are the reference's (Fortran column-major, first index
fastest); see exo_oracle.h.
"""
from __future__ import annotations

import ctypes
import math
import os
from ctypes import POINTER, byref, c_double, c_int, c_long

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF_DIR = os.path.join(HERE, "_ref")

_dp = POINTER(c_double)

# order of the eight legmod tables (src/legmod.f90:13-20 as stored by legini)
TABLES = ("qi", "qj", "qc", "qe", "qm", "qq", "qu", "qv")
# plavor as the model holds it (src/plasimmod.f90:118,227: plavor = EZ)
PLAVOR = 1.63299310207
# kappa = R / cp of the Earth set-up (src/p_earth.f90:41-58)
AKAP_EARTH = 0.286


def _libm():
    """the C library's single-precision exp, which the reference's
    default-real Fortran calls"""
    L = ctypes.CDLL("libm.so.6")
    L.expf.restype = ctypes.c_float
    L.expf.argtypes = [ctypes.c_float]
    return L


def _p(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


def dims(ntru: int):
    nlat = 3 * ntru // 2 + 1  # NTRU = (NLON-1)/3 with NLON = 2 NLAT
    nlat = {21: 32, 42: 64, 63: 96, 85: 128, 127: 192, 170: 256}.get(ntru, nlat)
    nlon = 2 * nlat
    ncsp = (ntru + 1) * (ntru + 2) // 2
    return nlat, nlon, ncsp


class Oracle:
    """ctypes wrapper of liboracle.so."""

    def __init__(self, path: str | None = None):
        path = path or os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} not built (run `make -C oracle`)")
        self.lib = L = ctypes.CDLL(path)
        L.ora_ql.restype = c_double
        L.ora_ql.argtypes = [c_int, c_double]
        L.ora_qld.restype = c_double
        L.ora_qld.argtypes = [c_int, c_double]
        L.ora_roundtrip_worksize.restype = c_long
        for name in ("ora_gp2fc", "ora_fc2gp", "ora_roundtrip", "ora_fft_supported",
                     "ora_minvers"):
            getattr(L, name).restype = c_int

    # -- grid and tables -------------------------------------------------
    def inigau(self, nlat: int):
        sid = np.zeros(nlat)
        gwd = np.zeros(nlat)
        self.lib.ora_inigau(c_int(nlat), _p(sid), _p(gwd))
        return sid, gwd

    def legini(self, ntru: int, nlat: int, sid=None, gwd=None) -> dict:
        if sid is None:
            sid, gwd = self.inigau(nlat)
        ncsp = (ntru + 1) * (ntru + 2) // 2
        tab = np.zeros((8, nlat, ncsp))
        self.lib.ora_legini(c_int(ntru), c_int(nlat), _p(np.ascontiguousarray(sid)),
                            _p(np.ascontiguousarray(gwd)), _p(tab))
        return {k: tab[i] for i, k in enumerate(TABLES)} | {"_all": tab}

    def filters(self, ntru, nfilter=0, ngptfilter=0, nspvfilter=0,
                filterkappa=8.0, nfilterexp=8, landhoskn0=15.0):
        g = np.zeros(ntru + 1)
        s = np.zeros(ntru + 1)
        self.lib.ora_filters(c_int(ntru), c_int(nfilter), c_int(ngptfilter),
                             c_int(nspvfilter), c_double(filterkappa),
                             c_int(nfilterexp), c_double(landhoskn0), _p(g), _p(s))
        return g, s

    # -- transforms (single field-level unless noted) --------------------
    def sp2fc(self, ntru, nlat, nlon, q, sk, sp):
        fc = np.zeros((nlat, nlon))
        self.lib.ora_sp2fc(c_int(ntru), c_int(nlat), c_int(nlon),
                           _p(np.ascontiguousarray(q)), _p(sk),
                           _p(np.ascontiguousarray(sp)), _p(fc))
        return fc

    def fc2sp(self, ntru, nlat, nlon, qc, sk, fc):
        ncsp = (ntru + 1) * (ntru + 2) // 2
        sp = np.zeros(2 * ncsp)
        self.lib.ora_fc2sp(c_int(ntru), c_int(nlat), c_int(nlon),
                           _p(np.ascontiguousarray(qc)), _p(sk),
                           _p(np.ascontiguousarray(fc)), _p(sp))
        return sp

    def dv2uv(self, ntru, nlat, nlon, nlev, qu, qv, sk, plavor, pd, pz):
        pu = np.zeros((nlev, nlat, nlon))
        pv = np.zeros((nlev, nlat, nlon))
        self.lib.ora_dv2uv(c_int(ntru), c_int(nlat), c_int(nlon), c_int(nlev),
                           _p(np.ascontiguousarray(qu)), _p(np.ascontiguousarray(qv)),
                           _p(sk), c_double(plavor), _p(np.ascontiguousarray(pd)),
                           _p(np.ascontiguousarray(pz)), _p(pu), _p(pv))
        return pu, pv

    def uv2dv(self, ntru, nlat, nlon, nlev, qe, qm, sk, pu, pv):
        ncsp = (ntru + 1) * (ntru + 2) // 2
        pd = np.zeros((nlev, 2 * ncsp))
        pz = np.zeros((nlev, 2 * ncsp))
        self.lib.ora_uv2dv(c_int(ntru), c_int(nlat), c_int(nlon), c_int(nlev),
                           _p(np.ascontiguousarray(qe)), _p(np.ascontiguousarray(qm)),
                           _p(sk), _p(np.ascontiguousarray(pu)),
                           _p(np.ascontiguousarray(pv)), _p(pd), _p(pz))
        return pd, pz

    def qtend(self, ntru, nlat, nlon, nlev, qc, qe, qm, sk, qn, uq, vq):
        ncsp = (ntru + 1) * (ntru + 2) // 2
        q = np.zeros((nlev, 2 * ncsp))
        c = np.ascontiguousarray
        self.lib.ora_qtend(c_int(ntru), c_int(nlat), c_int(nlon), c_int(nlev),
                           _p(c(qc)), _p(c(qe)), _p(c(qm)), _p(sk), _p(c(qn)),
                           _p(c(uq)), _p(c(vq)), _p(q))
        return q

    def mktend(self, ntru, nlat, nlon, nlev, qc, qe, qm, qq, sk, tn, fu, fv, ke,
               ut, vt):
        ncsp = (ntru + 1) * (ntru + 2) // 2
        d = np.zeros((nlev, 2 * ncsp))
        t = np.zeros((nlev, 2 * ncsp))
        z = np.zeros((nlev, 2 * ncsp))
        c = np.ascontiguousarray
        self.lib.ora_mktend(c_int(ntru), c_int(nlat), c_int(nlon), c_int(nlev),
                            _p(c(qc)), _p(c(qe)), _p(c(qm)), _p(c(qq)), _p(sk),
                            _p(c(tn)), _p(c(fu)), _p(c(fv)), _p(c(ke)), _p(c(ut)),
                            _p(c(vt)), _p(d), _p(t), _p(z))
        return d, t, z

    def fc2sp_unfolded(self, ntru, nlon, qc, sk, fc):
        """fc2sp's multi-process branch (src/legmod.f90:191-202) over the
        latitudes of qc (nlpp, ncsp) and fc (nlpp, nlon): partial sums"""
        ncsp = (ntru + 1) * (ntru + 2) // 2
        qc = np.ascontiguousarray(qc, dtype=np.float64)
        fc = np.ascontiguousarray(fc, dtype=np.float64)
        sp = np.zeros(2 * ncsp)
        self.lib.ora_fc2sp_unfolded(c_int(ntru), c_int(qc.shape[0]), c_int(nlon), _p(qc),
                                    _p(np.ascontiguousarray(sk, dtype=np.float64)), _p(fc),
                                    _p(sp))
        return sp

    def gp2fc(self, a: np.ndarray, dft: bool = False) -> np.ndarray:
        a = np.array(a, dtype=np.float64, order="C")
        n = a.shape[-1]
        lot = a.size // n
        if dft:
            self.lib.ora_gp2fc_dft(_p(a), c_int(n), c_int(lot))
        elif self.lib.ora_gp2fc(_p(a), c_int(n), c_int(lot)) != 0:
            raise ValueError(f"fftmod does not support n={n}")
        return a

    def fc2gp(self, a: np.ndarray, dft: bool = False) -> np.ndarray:
        a = np.array(a, dtype=np.float64, order="C")
        n = a.shape[-1]
        lot = a.size // n
        if dft:
            self.lib.ora_fc2gp_dft(_p(a), c_int(n), c_int(lot))
        elif self.lib.ora_fc2gp(_p(a), c_int(n), c_int(lot)) != 0:
            raise ValueError(f"fftmod does not support n={n}")
        return a

    def roundtrip(self, ntru, nlev, sp_in, tables=None, sk=None, use_dft=None,
                  plavor=PLAVOR):
        """One member round trip; returns (gp, sp_out)."""
        nlat, nlon, ncsp = dims(ntru)
        if tables is None:
            tables = self.legini(ntru, nlat)["_all"]
        if sk is None:
            sk = (np.ones(ntru + 1), np.ones(ntru + 1))
        if use_dft is None:
            use_dft = not bool(self.lib.ora_fft_supported(c_int(nlon)))
        nfl = 3 * nlev + 1
        ngp = nfl + 1 + 2 * nlev
        sp_in = np.ascontiguousarray(sp_in, dtype=np.float64).reshape(nfl, 2 * ncsp)
        gp = np.zeros((ngp, nlat, nlon))
        sp_out = np.zeros((nfl, 2 * ncsp))
        work = np.zeros(self.lib.ora_roundtrip_worksize(ntru, nlat, nlon, nlev))
        rc = self.lib.ora_roundtrip(c_int(ntru), c_int(nlat), c_int(nlon), c_int(nlev),
                                    _p(np.ascontiguousarray(tables)), _p(sk[1]),
                                    _p(sk[0]), c_double(plavor), c_int(int(use_dft)),
                                    _p(sp_in), _p(gp), _p(sp_out), _p(work))
        if rc != 0:
            raise ValueError("oracle round trip failed")
        return gp, sp_out

    def minvers(self, a: np.ndarray) -> np.ndarray:
        b = np.asfortranarray(a, dtype=np.float64).copy(order="F")
        flat = np.ascontiguousarray(b.T)  # column-major buffer
        self.lib.ora_minvers(_p(flat), c_int(a.shape[0]))
        return flat.T.copy()

    def si_apply(self, ntru, nlev, delt, t0, dsigma, g, tau, bm1, sop, spm, atm, adm, stt,
                 sdt, spt):
        """spectrala's semi-implicit step; g, tau (nlev, nlev) math order;
        bm1 (ntru, nlev, nlev) as returned by makebm; spectral arrays
        (nlev, nsp) (level-major).  Returns (stt, sdt, spt) updated copies."""
        nsp = (ntru + 1) * (ntru + 2)
        ni = np.concatenate([[n, n] for m in range(ntru + 1) for n in range(m, ntru + 1)])
        ni = np.ascontiguousarray(ni, dtype=np.int32)
        c = lambda a: np.ascontiguousarray(a, dtype=np.float64)  # noqa: E731
        gcm, tcm = c(np.asarray(g).T), c(np.asarray(tau).T)
        bcm = c(np.transpose(bm1, (0, 2, 1)))
        stt, sdt, spt = c(stt).copy(), c(sdt).copy(), c(spt).copy()
        self.lib.ora_si_apply(c_int(nlev), c_int(nsp), ni.ctypes.data_as(POINTER(c_int)),
                              c_double(delt), _p(c(t0)), _p(c(dsigma)), _p(gcm), _p(tcm),
                              _p(bcm), _p(c(sop)), _p(c(spm)), _p(c(atm)), _p(c(adm)),
                              _p(stt), _p(sdt), _p(spt))
        return stt, sdt, spt

    def makebm(self, nlev, ntru, delt, t0, dsigma, g, tau):
        """g, tau given as (nlev, nlev) numpy arrays in math (row, col) order."""
        gcm = np.ascontiguousarray(np.asarray(g, dtype=np.float64).T)
        tcm = np.ascontiguousarray(np.asarray(tau, dtype=np.float64).T)
        bm1 = np.zeros((ntru, nlev, nlev))
        self.lib.ora_makebm(c_int(nlev), c_int(ntru), c_double(delt),
                            _p(np.ascontiguousarray(t0, dtype=np.float64)),
                            _p(np.ascontiguousarray(dsigma, dtype=np.float64)),
                            _p(gcm), _p(tcm), _p(bm1))
        return np.transpose(bm1, (0, 2, 1)).copy()  # bm1[jn][row][col]

    # -- grid-point dynamics (exo_oracle_dyn.c) ---------------------------
    def initsi(self, sigmah, dsigma, rdsig, t0, akap=AKAP_EARTH):
        """initsi's matrices; g, tau, c returned in math (row, col) order."""
        n = len(t0)
        c = lambda a: np.ascontiguousarray(a, dtype=np.float64)  # noqa: E731
        g, tau, cm = (np.zeros((n, n)) for _ in range(3))
        t01s2, tkp = np.zeros(n), np.zeros(n)
        self.lib.ora_initsi(c_int(n), _p(c(sigmah)), _p(c(dsigma)), _p(c(rdsig)), _p(c(t0)),
                            c_double(akap), _p(g), _p(tau), _p(cm), _p(t01s2), _p(tkp))
        return {"g": g.T.copy(), "tau": tau.T.copy(), "c": cm.T.copy(), "t01s2": t01s2,
                "tkp": tkp}

    def vertical(self, nlev, t0=None, akap=AKAP_EARTH):
        """The vertical set-up of the dynamics: neqsig = 0 sigma levels,
        rdsig = 0.5 / dsigma (src/plasim.f90:1553-1567), a t0 profile
        (default 250 K, src/plasimmod.f90:593) and initsi's matrices."""
        s, sh, ds = sigma_levels(self.lib, nlev)
        t0 = np.full(nlev, 250.0) if t0 is None else np.asarray(t0, dtype=np.float64)
        v = {"sigma": s, "sigmah": sh, "dsigma": ds, "rdsig": 0.5 / ds, "t0": t0,
             "akap": akap}
        v.update(self.initsi(sh, ds, v["rdsig"], t0, akap))
        return v

    def calcgp(self, f, v, scal, nqspec=1):
        """calcgp on grid fields f (gu gv gt gd gz gq dq: (nlev, nhor); gp,
        gpj, gpm, rcsq: (nhor,)), vertical set-up v (vertical()) and scal =
        (rdbrv, adv, psurf, ww).  Returns gtn gqn guz gvz gphi dw (nlev, nhor)
        and gvp (nhor,)."""
        nlev, nhor = f["gu"].shape
        c = lambda a: np.ascontiguousarray(a, dtype=np.float64)  # noqa: E731
        out = {k: np.zeros((nlev, nhor)) for k in ("gtn", "gqn", "guz", "gvz", "gphi", "dw")}
        out["gvp"] = np.zeros(nhor)
        sc = c([v["akap"], *scal])
        self.lib.ora_calcgp(c_int(nlev), c_long(nhor), *[_p(c(f[k])) for k in
                            ("gu", "gv", "gt", "gd", "gz", "gq", "gp", "gpj", "gpm", "dq",
                             "rcsq")],
                            *[_p(c(v[k])) for k in ("sigmah", "dsigma", "rdsig", "t0",
                                                     "t01s2", "tkp")],
                            _p(c(np.asarray(v["c"]).T)), _p(c(np.asarray(v["g"]).T)), _p(sc),
                            c_int(nqspec), *[_p(out[k]) for k in
                                             ("gtn", "gqn", "guz", "gvz", "gvp", "gphi", "dw")])
        return out

    def gridpointa(self, ntru, nlev, state, v, scal, nqspec=1, gq=None, dq=None, sk=None,
                   plavor=PLAVOR):
        """The transform group of gridpointa (src/plasim.f90:2402-2486, its
        diagnostics left out) for one member, composed of the oracle's
        restatements in the reference's order: invlega (src/legmod.f90:
        635-656), gpmt (:2425-2430), fc2gp, gp = exp(gp), calcgp, the products
        (:2436-2446), gp2fc, fc2sp(gvpp), mktend (fu <- gvz, fv <- guz) and
        qtend.  state: (4*nlev+1, nrsp) = sz sd st sq sp; with nqspec = 0 the
        grid gq and dq (nlev, nhor) are the caller's.  Returns the
        tendencies in the same layout (sqt zero when nqspec = 0) and the grid
        fields (dict)."""
        nlat, nlon, ncsp = dims(ntru)
        n, nh = nlev, nlat * nlon
        sk = np.ones(ntru + 1) if sk is None else sk
        t = self.legini(ntru, nlat)
        sz, sd, st, sq, sp = (state[:n], state[n:2 * n], state[2 * n:3 * n],
                              state[3 * n:4 * n], state[4 * n])
        fc = lambda x, q="qi": self.sp2fc(ntru, nlat, nlon, t[q], sk, x)  # noqa: E731
        lev = lambda a: np.stack([fc(a[j]) for j in range(n)])  # noqa: E731
        dft = nlon == 192  # T63 runs fft991 in the reference build (pkg/compile.sh:103-108)
        grid = lambda f: self.fc2gp(f, dft).reshape(f.shape[0] if f.ndim == 3 else 1, nh)  # noqa
        pu, pv = self.dv2uv(ntru, nlat, nlon, n, t["qu"], t["qv"], sk, plavor, sd, sz)
        f = {"gd": grid(lev(sd)), "gt": grid(lev(st)), "gz": grid(lev(sz)),
             "gu": grid(pu), "gv": grid(pv)}
        f["gq"] = grid(lev(sq)) if nqspec == 1 else np.asarray(gq, dtype=np.float64)
        f["dq"] = np.zeros((n, nh)) if dq is None else np.asarray(dq, dtype=np.float64)
        fcp = fc(sp)
        m = np.arange(nlon // 2)
        fm = np.empty_like(fcp)
        fm[:, 0::2] = -fcp[:, 1::2] * m
        fm[:, 1::2] = fcp[:, 0::2] * m
        f["gpm"] = grid(fm)[0]
        f["gpj"] = grid(fc(sp, "qj"))[0]
        f["gp"] = np.exp(grid(fcp)[0])
        sid, _ = self.inigau(nlat)
        f["rcsq"] = np.repeat(1.0 / (1.0 - sid * sid), nlon)
        out = self.calcgp(f, v, scal, nqspec)
        f.update(out)
        f["gut"], f["gvt"] = f["gu"] * f["gt"], f["gv"] * f["gt"]
        f["gke"] = f["gu"] * f["gu"] + f["gv"] * f["gv"] + out["gphi"]
        if nqspec == 1:
            f["guq"], f["gvq"] = f["gu"] * f["gq"], f["gv"] * f["gq"]
        four = lambda a: self.gp2fc(a.reshape(-1, nlat, nlon), dft)  # noqa: E731
        F = {k: four(f[k]) for k in ("gtn", "gqn", "gut", "gvt", "guz", "gvz", "gke")}
        tend = np.zeros((4 * n + 1, 2 * ncsp))
        tend[4 * n] = self.fc2sp(ntru, nlat, nlon, t["qc"], sk, four(out["gvp"])[0])
        d, tt, z = self.mktend(ntru, nlat, nlon, n, t["qc"], t["qe"], t["qm"], t["qq"], sk,
                               F["gtn"], F["gvz"], F["guz"], F["gke"], F["gut"], F["gvt"])
        tend[:n], tend[n:2 * n], tend[2 * n:3 * n] = z, d, tt
        if nqspec == 1:
            tend[3 * n:4 * n] = self.qtend(ntru, nlat, nlon, n, t["qc"], t["qe"], t["qm"], sk,
                                           F["gqn"], four(f["guq"]), four(f["gvq"]))
        return tend, f


# ---------------------------------------------------------------------------
# the multi-process pieces around the transforms
# ---------------------------------------------------------------------------
def writesp(pf: np.ndarray, ntru: int, pscale: float, poff: float, real=np.float64):
    """writesp's values (src/outmod.f90:141-169) of one spectral field pf
    (NRSP): zf = real(4)(pf * spnorm * pscale), zf(1) = zf(1) + poff, with
    spnorm = +-1/sqrt(2) alternating with m (src/plasim.f90:1638-1651), all
    in the model's real kind `real`"""
    norm = real(1.0) / np.sqrt(real(2.0))
    spn = []
    for m in range(ntru + 1):
        spn += [norm] * (2 * (ntru + 1 - m))
        norm = -norm
    spn = np.array(spn, dtype=real)
    zf = ((np.asarray(pf, dtype=real) * spn) * real(pscale)).astype(np.float32)
    zf[0] = np.float32(real(zf[0]) + real(poff))
    return zf


def sumsc(psf: np.ndarray, nspp: int) -> np.ndarray:
    """mpsumsc (src/mpimod.f90:279-291): MPI_Reduce_scatter with counts
    nscatsp = NSPP (src/plasimmod.f90:306) over the processes' partial sums
    psf (nproc, nfield, NESP); process r receives elements r NSPP .. of the
    sum, here summed in process order."""
    nproc = psf.shape[0]
    out = np.zeros((nproc, psf.shape[1], nspp))
    for r in range(nproc):
        acc = psf[0, :, r * nspp:(r + 1) * nspp].copy()
        for q in range(1, nproc):
            acc = acc + psf[q, :, r * nspp:(r + 1) * nspp]
        out[r] = acc
    return out


def mrtrunc(sphi: np.ndarray, nthi: int, ntlo: int) -> np.ndarray:
    """mrtrunc (src/mpimod_multi.f90:117-131) of one field sphi(2, *)"""
    splo = np.zeros(2 * (ntlo + 1) * (ntlo + 2) // 2)
    jl = jh = 0
    for jm in range(ntlo + 1):
        jn = ntlo - jm
        splo[2 * jl:2 * (jl + jn + 1)] = sphi[2 * jh:2 * (jh + jn + 1)]
        jl += ntlo - jm + 1
        jh += nthi - jm + 1
    return splo


def mrexpand(splo: np.ndarray, ntlo: int, sphi: np.ndarray, nthi: int) -> np.ndarray:
    """mrexpand (src/mpimod_multi.f90:137-153) into a copy of sphi: the
    reference clears sphi(:, 1:nthi), the first nthi complex values, then
    copies the low truncation in"""
    out = np.array(sphi, dtype=np.float64, copy=True)
    out[:2 * nthi] = 0.0
    jl = jh = 0
    for jm in range(ntlo + 1):
        jn = ntlo - jm
        out[2 * jh:2 * (jh + jn + 1)] = splo[2 * jl:2 * (jl + jn + 1)]
        jl += ntlo - jm + 1
        jh += nthi - jm + 1
    return out


# ---------------------------------------------------------------------------
# the reference's own Fortran, compiled into oracle/_ref (optional)
# ---------------------------------------------------------------------------

def _ref_path(name: str) -> str:
    return os.path.join(REF_DIR, name)


def ref_available(name: str = "libref_t21.so") -> bool:
    return os.path.exists(_ref_path(name))


class RefT21:
    """The reference legmod/gaussmod/fftmod at its default resolution module
    (T21 L10, src/resmod_def.f90), FP64."""

    def __init__(self):
        self.lib = L = ctypes.CDLL(_ref_path("libref_t21.so"))
        d = (c_int * 6)()
        L.ref_t21_dims(d)
        self.nlat, self.nlon, self.ntru, self.nlev, self.ncsp, self.nesp = list(d)
        self.sid = np.zeros(self.nlat)
        self.gwd = np.zeros(self.nlat)
        L.ref_t21_init(_p(self.sid), _p(self.gwd))
        tab = np.zeros((8, self.nlat, self.ncsp))
        L.ref_t21_tables(_p(tab))
        self.tables = tab
        pl = c_double()
        L.ref_t21_plavor(byref(pl))
        self.plavor = pl.value

    def _fc(self, nlev=1):
        return np.zeros((nlev, self.nlat, self.nlon))

    def sp2fc(self, sp):
        fc = self._fc()[0]
        self.lib.sp2fc_(_p(np.ascontiguousarray(sp)), _p(fc))
        return fc

    def sp2fcdmu(self, sp):
        fc = self._fc()[0]
        self.lib.sp2fcdmu_(_p(np.ascontiguousarray(sp)), _p(fc))
        return fc

    def fc2sp(self, fc):
        sp = np.zeros(self.nesp)
        self.lib.fc2sp_(_p(np.ascontiguousarray(fc)), _p(sp))
        return sp

    def dv2uv(self, pd, pz):
        pu, pv = self._fc(self.nlev), self._fc(self.nlev)
        pz = np.array(pz, dtype=np.float64)
        self.lib.dv2uv_(_p(np.ascontiguousarray(pd)), _p(pz), _p(pu), _p(pv))
        return pu, pv

    def uv2dv(self, pu, pv):
        pd = np.zeros((self.nlev, self.nesp))
        pz = np.zeros((self.nlev, self.nesp))
        self.lib.uv2dv_(_p(np.ascontiguousarray(pu)), _p(np.ascontiguousarray(pv)),
                        _p(pd), _p(pz))
        return pd, pz

    def qtend(self, qn, uq, vq):
        q = np.zeros((self.nlev, self.nesp))
        c = np.ascontiguousarray
        self.lib.qtend_(_p(q), _p(c(qn)), _p(c(uq)), _p(c(vq)))
        return q

    def mktend(self, tn, fu, fv, ke, ut, vt):
        d, t, z = (np.zeros((self.nlev, self.nesp)) for _ in range(3))
        c = np.ascontiguousarray
        self.lib.mktend_(_p(d), _p(t), _p(z), _p(c(tn)), _p(c(fu)), _p(c(fv)),
                         _p(c(ke)), _p(c(ut)), _p(c(vt)))
        return d, t, z

    def gp2fc(self, a):
        a = np.array(a, dtype=np.float64, order="C")
        n = a.shape[-1]
        self.lib.gp2fc_(_p(a), byref(c_int(n)), byref(c_int(a.size // n)))
        return a

    def fc2gp(self, a):
        a = np.array(a, dtype=np.float64, order="C")
        n = a.shape[-1]
        self.lib.fc2gp_(_p(a), byref(c_int(n)), byref(c_int(a.size // n)))
        return a


class RefPyfft:
    """The reference's resolution-parameterised transform copies
    (exoplasim/pyfft.f90): kind "r8" is the -fdefault-real-8 build (every
    literal in double precision), kind "r4" the default-real build the
    post-processor's f2py step makes (exoplasim/__init__.py:84-87).  The
    wrappers below take and return C-ordered numpy arrays: spectral
    (nlev, nrsp), grid (nlev, nlat, nlon), Fourier (nlev, nlat, nlon)."""

    def __init__(self, kind: str = "r8"):
        name = {"r8": "libref_pyfft.so", "r4": "libref_pyfft_r4.so"}[kind]
        self.lib = ctypes.CDLL(_ref_path(name))

    @staticmethod
    def _i(*v):
        return [byref(c_int(int(x))) for x in v]

    def sp2gp(self, ntru, nlat, sp, nfs=0):
        sp = np.ascontiguousarray(sp, dtype=np.float64)
        nlev = sp.shape[0]
        gp = np.zeros((nlev, nlat, 2 * nlat))
        self.lib.sp2gp_(_p(sp), _p(gp), *self._i(nlat, 2 * nlat, ntru, nlev, nfs))
        return gp

    def gp2sp(self, ntru, gp, nfs=0):
        gp = np.ascontiguousarray(gp, dtype=np.float64)
        nlev, nlat, nlon = gp.shape
        sp = np.zeros((nlev, (ntru + 1) * (ntru + 2)))
        self.lib.gp2sp_(_p(gp), _p(sp), *self._i(nlat, nlon, ntru, nlev, nfs))
        return sp

    def sp3fc(self, ntru, nlat, sp, nfs=0):
        sp = np.ascontiguousarray(sp, dtype=np.float64)
        nlev = sp.shape[0]
        fc = np.zeros((nlev, nlat, 2 * nlat))
        self.lib.sp3fc_(_p(sp), _p(fc), *self._i(nlat, 2 * nlat, ntru, nlev, nfs))
        return fc

    def gp3fc(self, gp):
        gp = np.ascontiguousarray(gp, dtype=np.float64)
        nlev, nlat, nlon = gp.shape
        fc = np.zeros_like(gp)
        self.lib.gp3fc_(_p(gp), _p(fc), *self._i(nlon, nlat, nlev))
        return fc

    def fc3gp(self, fc):
        fc = np.ascontiguousarray(fc, dtype=np.float64)
        nlev, nlat, nlon = fc.shape
        gp = np.zeros_like(fc)
        self.lib.fc3gp_(_p(fc), _p(gp), *self._i(nlon, nlat, nlev))
        return gp

    def spvgp(self, ntru, sd, sz, rdcostheta, nfs=0):
        sd = np.ascontiguousarray(sd, dtype=np.float64)
        sz = np.ascontiguousarray(sz, dtype=np.float64)
        rd = np.ascontiguousarray(rdcostheta, dtype=np.float64)
        nlev, nlat = sd.shape[0], rd.size
        gu = np.zeros((nlev, nlat, 2 * nlat))
        gv = np.zeros_like(gu)
        self.lib.spvgp_(_p(sd), _p(sz), _p(rd), _p(gu), _p(gv),
                        *self._i(nlat, 2 * nlat, ntru, nlev, nfs))
        return gu, gv

    def gpvsp(self, ntru, gu, gv, costhetadr, nfs=0):
        gu = np.ascontiguousarray(gu, dtype=np.float64)
        gv = np.ascontiguousarray(gv, dtype=np.float64)
        cd = np.ascontiguousarray(costhetadr, dtype=np.float64)
        nlev, nlat, nlon = gu.shape
        sd = np.zeros((nlev, (ntru + 1) * (ntru + 2)))
        sz = np.zeros_like(sd)
        self.lib.gpvsp_(_p(gu), _p(gv), _p(cd), _p(sd), _p(sz),
                        *self._i(nlat, nlon, ntru, nlev, nfs))
        return sd, sz

    def legini(self, ntru, nlat):
        nlon = 2 * nlat
        ncsp = (ntru + 1) * (ntru + 2) // 2
        tab = np.zeros((8, nlat, ncsp))
        sfilt = np.zeros(ntru + 1)
        args = [byref(c_int(nlat)), byref(c_int(nlon)), byref(c_int(ntru)),
                byref(c_int(1))]
        bufs = [np.zeros((nlat, ncsp)) for _ in range(8)]
        args += [_p(b) for b in bufs]
        self.lib.legini_(*args, _p(sfilt), byref(c_int(0)))
        for i in range(8):
            tab[i] = bufs[i]
        return tab

    def sfilt(self, ntru, nlat, nfs):
        """the physics filter legini returns (pyfft.f90:150-153)"""
        ncsp = (ntru + 1) * (ntru + 2) // 2
        bufs = [np.zeros((nlat, ncsp)) for _ in range(8)]
        sf = np.zeros(ntru + 1)
        self.lib.legini_(byref(c_int(nlat)), byref(c_int(2 * nlat)), byref(c_int(ntru)),
                         byref(c_int(1)), *[_p(b) for b in bufs], _p(sf), byref(c_int(nfs)))
        return sf

    def sp2fc(self, ntru, nlat, sp):
        fc = np.zeros((nlat, 2 * nlat))
        self.lib.sp2fc_(_p(np.ascontiguousarray(sp)), _p(fc), byref(c_int(nlat)),
                        byref(c_int(2 * nlat)), byref(c_int(ntru)), byref(c_int(1)),
                        byref(c_int(0)))
        return fc

    def fc2sp(self, ntru, nlat, fc):
        ncsp = (ntru + 1) * (ntru + 2) // 2
        sp = np.zeros(2 * ncsp)
        self.lib.fc2sp_(_p(np.ascontiguousarray(fc)), _p(sp), byref(c_int(nlat)),
                        byref(c_int(2 * nlat)), byref(c_int(ntru)), byref(c_int(1)),
                        byref(c_int(0)))
        return sp


class PyfftOracle:
    """The post-processor's transform wrappers (exoplasim/pyfft.f90:
    sp2gp :971-1005, spvgp :1012-1069, gpvsp :1076-1134, gp2sp :1141-1174,
    sp3fc :344-370, fc3gp :905-930, gp3fc :938-963) restated over the
    oracle's legmod/fftmod transforms: pyfft's legini has the same tables as
    legmod (pyfft.f90:73-146), dv2uv has no planetary vorticity
    (pyfft.f90:527-545), one filter serves both directions, and fc2gp ends
    with a multiplication by sqrt 2 (pyfft.f90:891).

    Rows of 192 (T63) run the FFT991 definition (the post-processor's
    pyfft991.f90 copy; pyfft.f90's own FFT stops for that length,
    pyfft.f90:791-792).

    kind "r4" evaluates the default-real pieces as the post-processor's
    default-real build does: the sqrt 2 literal and the filter formula
    (pyfft.f90:150-153) in single precision.  kind "r8" evaluates them in
    double precision, as the -fdefault-real-8 build of the same file does.
    Arrays are C-ordered as RefPyfft's."""

    def __init__(self, oracle: "Oracle", kind: str = "r4"):
        self.o = oracle
        self.kind = kind
        self.sqrt2 = (float(np.float32(1.4142135623730951)) if kind == "r4"
                      else 1.4142135623730951)
        self._tabs = {}

    def sfilt(self, ntru, nfs):
        """pyfft.f90:150-153, value of n = 1 .. NTRU+1 at index n-1.  The
        power x**8 is evaluated as seven successive products and exp by the C
        library, as the reference compiled here (amdflang) evaluates them;
        in the default-real build a compiler that expands the power by
        squaring instead moves the filter tail by up to 3e-6 relative."""
        if nfs == 0:
            return np.ones(ntru + 1)
        if self.kind == "r8":
            x = np.arange(1, ntru + 2, dtype=np.float64) / ntru
            p = x.copy()
            for _ in range(7):
                p = p * x
            return np.array([math.exp(-8.0 * v) for v in p])
        x = np.arange(1, ntru + 2, dtype=np.float32) / np.float32(ntru)
        p = x.copy()
        for _ in range(7):
            p = p * x
        expf = _libm().expf
        return np.array([expf(float(np.float32(-8.0) * v)) for v in p], dtype=np.float64)

    def _fc2gp(self, a):
        return self.o.fc2gp(a, dft=np.shape(a)[-1] == 192)

    def _gp2fc(self, a):
        return self.o.gp2fc(a, dft=np.shape(a)[-1] == 192)

    def _t(self, ntru, nlat):
        if (ntru, nlat) not in self._tabs:
            self._tabs[(ntru, nlat)] = self.o.legini(ntru, nlat)
        return self._tabs[(ntru, nlat)]

    def sp2gp(self, ntru, nlat, sp, nfs=0):
        sp = np.atleast_2d(sp)
        t, f = self._t(ntru, nlat), self.sfilt(ntru, nfs)
        fc = np.stack([self.o.sp2fc(ntru, nlat, 2 * nlat, t["qi"], f, x) for x in sp])
        return self._fc2gp(fc) * self.sqrt2

    def gp2sp(self, ntru, gp, nfs=0):
        gp = np.asarray(gp, dtype=np.float64)
        nlat = gp.shape[1]
        t, f = self._t(ntru, nlat), self.sfilt(ntru, nfs)
        fc = self._gp2fc(gp / self.sqrt2)
        return np.stack([self.o.fc2sp(ntru, nlat, 2 * nlat, t["qc"], f, x) for x in fc])

    def sp3fc(self, ntru, nlat, sp, nfs=0):
        sp = np.atleast_2d(sp)
        t, f = self._t(ntru, nlat), self.sfilt(ntru, nfs)
        return np.stack([self.o.sp2fc(ntru, nlat, 2 * nlat, t["qi"], f, x) for x in sp])

    def gp3fc(self, gp):
        return self._gp2fc(gp)

    def fc3gp(self, fc):
        return self._fc2gp(fc) * self.sqrt2

    def spvgp(self, ntru, sd, sz, rdcostheta, nfs=0):
        sd, sz = np.atleast_2d(sd), np.atleast_2d(sz)
        rd = np.asarray(rdcostheta, dtype=np.float64)
        nlat = rd.size
        t, f = self._t(ntru, nlat), self.sfilt(ntru, nfs)
        pu, pv = self.o.dv2uv(ntru, nlat, 2 * nlat, sd.shape[0], t["qu"], t["qv"], f, 0.0,
                              sd, sz)
        gu = (self._fc2gp(pu) * self.sqrt2) * rd[None, :, None]
        gv = (self._fc2gp(pv) * self.sqrt2) * rd[None, :, None]
        return gu, gv

    def gpvsp(self, ntru, gu, gv, costhetadr, nfs=0):
        cd = np.asarray(costhetadr, dtype=np.float64)
        nlat = cd.size
        t, f = self._t(ntru, nlat), self.sfilt(ntru, nfs)
        fu = self._gp2fc((np.asarray(gu) * cd[None, :, None]) / self.sqrt2)
        fv = self._gp2fc((np.asarray(gv) * cd[None, :, None]) / self.sqrt2)
        return self.o.uv2dv(ntru, nlat, 2 * nlat, fu.shape[0], t["qe"], t["qm"], f, fu, fv)


class RefPlasim:
    """The reference's plasim.f90 at its default resolution module (T21 L10,
    src/resmod_def.f90), FP64: initsi, makebm, minvers, the semi-implicit
    part of spectrala and calcgp, through oracle/ref_glue_plasim.f90.
    Loaded with lazy binding: plasim.f90 references model subroutines this
    library does not link, none of which these calls reach."""

    def __init__(self):
        self.lib = ctypes.CDLL(_ref_path("libref_plasim.so"), mode=os.RTLD_LAZY | os.RTLD_LOCAL)
        d = (c_int * 8)()
        self.lib.ref_plasim_dims(d)
        (self.nlat, self.nlon, self.ntru, self.nlev, self.ncsp, self.nesp, self.nhor,
         self.nspp) = list(d)

    @staticmethod
    def _c(a):
        return np.ascontiguousarray(a, dtype=np.float64)

    def initsi(self, sigma, sigmah, dsigma, rdsig, t0, akap=AKAP_EARTH):
        n = self.nlev
        g, tau, cm = (np.zeros((n, n)) for _ in range(3))
        t01s2, tkp = np.zeros(n), np.zeros(n)
        c = self._c
        self.lib.ref_plasim_initsi(_p(c(sigma)), _p(c(sigmah)), _p(c(dsigma)), _p(c(rdsig)),
                                   _p(c(t0)), byref(c_double(akap)), _p(g), _p(tau), _p(cm),
                                   _p(t01s2), _p(tkp))
        return {"g": g.T.copy(), "tau": tau.T.copy(), "c": cm.T.copy(), "t01s2": t01s2,
                "tkp": tkp}

    def makebm(self, delt, t0, dsigma, g, tau):
        """bm1[jn][row][col] as Oracle.makebm; g, tau in math order."""
        n, nt = self.nlev, self.ntru
        bm1 = np.zeros((nt, n, n))
        c = self._c
        self.lib.ref_plasim_makebm(byref(c_double(delt)), _p(c(t0)), _p(c(dsigma)),
                                   _p(c(np.asarray(g).T)), _p(c(np.asarray(tau).T)), _p(bm1))
        return np.transpose(bm1, (0, 2, 1)).copy()

    def minvers(self, a):
        flat = np.ascontiguousarray(np.asarray(a, dtype=np.float64).T)
        self.lib.minvers_(_p(flat), byref(c_int(a.shape[0])))
        return flat.T.copy()

    def si_apply(self, delt, t0, dsigma, g, tau, bm1, sop, spm, atm, adm, stt, sdt, spt):
        """spectrala sections 1-3 (arguments as Oracle.si_apply, spectral
        arrays (nlev, nsp)); returns (stt, sdt, spt)."""
        c = self._c
        stt, sdt, spt = c(stt).copy(), c(sdt).copy(), c(spt).copy()
        self.lib.ref_plasim_si(byref(c_double(delt)), _p(c(t0)), _p(c(dsigma)),
                               _p(c(np.asarray(g).T)), _p(c(np.asarray(tau).T)),
                               _p(c(np.transpose(bm1, (0, 2, 1)))), _p(c(sop)), _p(c(spm)),
                               _p(c(atm)), _p(c(adm)), _p(stt), _p(sdt), _p(spt))
        return stt, sdt, spt

    def calcgp(self, f, v, scal, nqspec=1):
        """calcgp with the arguments of Oracle.calcgp at (NLEV, NHOR); dq
        has NLEV levels here and is padded to the model's NLEP."""
        n, nh = self.nlev, self.nhor
        c = self._c
        out = {k: np.zeros((n, nh)) for k in ("gtn", "gqn", "guz", "gvz", "gphi", "dw")}
        out["gvp"] = np.zeros(nh)
        dq = np.zeros((n + 1, nh))
        dq[:n] = f["dq"]
        self.lib.ref_plasim_calcgp(*[_p(c(f[k])) for k in ("gu", "gv", "gt", "gd", "gz", "gq",
                                                             "gp", "gpj", "gpm")],
                                   _p(dq), _p(c(f["rcsq"])),
                                   *[_p(c(v[k])) for k in ("sigmah", "dsigma", "rdsig", "t0")],
                                   byref(c_double(v["akap"])), _p(c(scal)), c_int(nqspec),
                                   *[_p(out[k]) for k in ("gtn", "gqn", "guz", "gvz", "gvp",
                                                          "gphi", "dw")])
        return out


class RefGridpointa:
    """The reference's gridpointa (src/plasim.f90:2360-2585) with legmod,
    gaussmod and fftmod at the default resolution module (T21 L10), FP64,
    through oracle/ref_glue_gridpointa.f90; lazy binding as RefPlasim."""

    def __init__(self):
        self.lib = ctypes.CDLL(_ref_path("libref_gpa.so"), mode=os.RTLD_LAZY | os.RTLD_LOCAL)
        self.ntru, self.nlat, self.nlon, self.nlev = 21, 32, 64, 10

    def run(self, state, v, scal, nqspec=1, gq=None, dq=None):
        """state (4*nlev+1, nrsp) = sz sd st sq sp; returns the tendencies in
        the same layout and the grid fields gz gd gt gu gv dw (nlev, nhor),
        gp gpj (nhor,)"""
        n, nh = self.nlev, self.nlat * self.nlon
        c = lambda a: np.ascontiguousarray(a, dtype=np.float64)  # noqa: E731
        tend = np.zeros_like(c(state))
        grid = np.zeros((6 * n + 2, nh))
        gq = np.zeros((n, nh)) if gq is None else gq
        dq = np.zeros((n, nh)) if dq is None else dq
        self.lib.ref_gridpointa(_p(c(state)), *[_p(c(v[k])) for k in ("sigmah", "dsigma",
                                                                       "rdsig", "t0")],
                                byref(c_double(v["akap"])), _p(c(scal)), c_int(nqspec),
                                _p(c(gq)), _p(c(dq)), _p(tend), _p(grid))
        names = ("gz", "gd", "gt", "gu", "gv", "dw")
        g = {k: grid[i * n:(i + 1) * n] for i, k in enumerate(names)}
        g["gp"], g["gpj"] = grid[6 * n], grid[6 * n + 1]
        return tend, g


class RefFft991:
    """The reference's FFT991 module (src/fft991mod.f90), FP64."""

    def __init__(self):
        self.lib = ctypes.CDLL(_ref_path("libref_fft991.so"))

    def gp2fc(self, a):
        a = np.array(a, dtype=np.float64, order="C")
        n = a.shape[-1]
        self.lib.gp2fc_(_p(a), byref(c_int(n)), byref(c_int(a.size // n)))
        return a

    def fc2gp(self, a):
        a = np.array(a, dtype=np.float64, order="C")
        n = a.shape[-1]
        self.lib.fc2gp_(_p(a), byref(c_int(n)), byref(c_int(a.size // n)))
        return a


# ---------------------------------------------------------------------------
# long-wave radiation column (src/radmod.f90:2122-2456, 1100-1108)
# ---------------------------------------------------------------------------
# physical constants of the Earth setup (src/p_earth.f90:41-58,
# src/plasim.f90:1317-1318, src/plasimmod.f90:122)
GA = 9.80665
GASCON = 287.0
AKAP = 0.286
ACPD = GASCON / AKAP
ADV = 1870.0 / ACPD - 1.0
# radmod namelist defaults (src/radmod.f90:37-43,71)
LWR_DEFAULTS = dict(th2oc=0.024, tpofmt=1.0, acllwr=0.1, clgray=-1.0, npbroaden=1)


def sigma_levels(oracle_lib, nlev):
    s = np.zeros(nlev)
    sh = np.zeros(nlev)
    ds = np.zeros(nlev)
    oracle_lib.ora_sigma_levels(c_int(nlev), _p(s), _p(sh), _p(ds))
    return s, sh, ds


def lwr(ora: "Oracle", cols: dict, nlev: int, **params):
    """Run the oracle lwr on columns.  cols: dt (ncol, nlev+1), dq, dqo3, dqco2,
    dcc, dql (ncol, nlev), dp, dls (ncol).  Returns dftu, dftd (ncol, nlev+1),
    heating (ncol, nlev) in K/s and OLR (ncol)."""
    p = dict(LWR_DEFAULTS)
    p.update(params)
    s, sh, ds = sigma_levels(ora.lib, nlev)
    ncol = cols["dp"].shape[0]
    c = {k: np.ascontiguousarray(v, dtype=np.float64) for k, v in cols.items()}
    dftu = np.zeros((ncol, nlev + 1))
    dftd = np.zeros((ncol, nlev + 1))
    ora.lib.ora_lwr(c_int(nlev), c_int(ncol), _p(s), _p(sh), _p(ds), c_double(GA),
                    c_double(p["th2oc"]), c_double(p["tpofmt"]), c_double(p["acllwr"]),
                    c_double(p["clgray"]), c_int(int(p["npbroaden"])), _p(c["dp"]), _p(c["dt"]),
                    _p(c["dq"]), _p(c["dqo3"]), _p(c["dqco2"]), _p(c["dcc"]), _p(c["dql"]),
                    _p(c["dls"]), _p(dftu), _p(dftd))
    heat = np.zeros((ncol, nlev))
    ora.lib.ora_lwr_heating(c_int(nlev), c_int(ncol), _p(ds), c_double(GA), c_double(ACPD),
                            c_double(ADV), _p(c["dp"]), _p(c["dq"]), _p(dftu), _p(dftd),
                            _p(heat))
    return dftu, dftd, heat, -dftu[:, 0]


class RefRad:
    """The reference's lwr at its default resolution module (NHOR=2048
    columns, NLEV=10), FP64; loaded with lazy binding because radmod's other
    routines reference model subroutines that this library does not link."""

    def __init__(self, name: str = "libref_rad_t21.so"):
        self.lib = ctypes.CDLL(_ref_path(name), mode=os.RTLD_LAZY | os.RTLD_LOCAL)
        d = (c_int * 2)()
        self.lib.ref_rad_dims(d)
        self.nhor, self.nlev = list(d)

    def lwr(self, cols: dict, **params):
        p = dict(LWR_DEFAULTS)
        p.update(params)
        nl, nh = self.nlev, self.nhor
        s, sh, ds = sigma_levels(Oracle().lib, nl)
        f = {k: np.asfortranarray(np.asarray(v, dtype=np.float64)) for k, v in cols.items()}
        # (ncol, lev) arrays are passed as Fortran (NHOR, lev): transpose to C order
        t = {k: np.ascontiguousarray(v.T) if v.ndim == 2 else np.ascontiguousarray(v)
             for k, v in f.items()}
        ftu = np.zeros((nl + 1, nh))
        ftd = np.zeros((nl + 1, nh))
        dbl = lambda x: byref(c_double(x))  # noqa: E731
        self.lib.ref_rad_lwr(_p(s), _p(sh), _p(ds), dbl(GA), dbl(p["th2oc"]), dbl(p["tpofmt"]),
                             dbl(p["acllwr"]), dbl(p["clgray"]), c_int(int(p["npbroaden"])),
                             _p(t["dp"]), _p(t["dt"]), _p(t["dq"]), _p(t["dqo3"]),
                             _p(t["dqco2"]), _p(t["dcc"]), _p(t["dql"]), _p(t["dls"]),
                             _p(ftu), _p(ftd))
        return ftu.T.copy(), ftd.T.copy()
