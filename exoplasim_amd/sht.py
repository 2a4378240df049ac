"""Host-side mirror of the reference's spectral-transform interface on top of
libexosht.so.

Two layers, as in include/exo_sht.h:

* legmod / fftmod / gaussmod-named functions (``inigau``, ``legini``,
  ``sp2fc``, ``sp2fcdmu``, ``fc2sp``, ``dv2uv``, ``gp2fc``, ``fc2gp``) taking
  numpy arrays in the reference layout (exoplasim/plasim/src/legmod.f90,
  fftmod.f90, gaussmod.f90); float64 arrays use the FP64 path, float32 the
  FP32 path.  ``legini`` must be called first, as in the model
  (src/plasim.f90:183,411).
* ``Plan``: a device-resident batched transform of one truncation working on
  torch tensors that already live on the GPU; the caller's current torch
  stream is used.

Everything runs on the GPU through the HIP library; there is no CPU path.
"""
from __future__ import annotations

import ctypes
from ctypes import POINTER, byref, c_double, c_float, c_int, c_void_p

import numpy as np

from ._lib import ExoError, check, lib

PLAVOR = 1.63299310207  # src/plasimmod.f90:118,227 (plavor = EZ)
MODES = {"f64": 0, "f32": 1, "f32mfma": 2, "bf16x3": 3}
TABLES = ("qi", "qj", "qc", "qe", "qm", "qq", "qu", "qv")
NLAT_OF = {21: 32, 42: 64, 63: 96, 127: 192}

_dp = POINTER(c_double)
_fp = POINTER(c_float)


def _cd(a: np.ndarray):
    return a.ctypes.data_as(_dp if a.dtype == np.float64 else _fp)


def _opt(a):
    if a is None:
        return None
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a


def mode_id(mode) -> int:
    if isinstance(mode, int):
        return mode
    return MODES[mode]


def dims(ntru: int):
    nlat = NLAT_OF[ntru]
    return nlat, 2 * nlat, (ntru + 1) * (ntru + 2) // 2


# ---------------------------------------------------------------------------
# drop-in layer (host arrays)
# ---------------------------------------------------------------------------
_state = {"ntru": None, "nlat": None, "nlev": None}


def inigau(klat: int):
    """Gaussian abscissas (north->south) and weights (src/gaussmod.f90:5)."""
    pz0 = np.zeros(klat)
    pzw = np.zeros(klat)
    check(lib().exo_inigau(klat, _cd(pz0), _cd(pzw)), "inigau")
    return pz0, pzw


def legini(ntru: int, nlat: int, nlev: int, sid=None, gwd=None, skgpsp=None, skspgp=None,
           plavor: float = PLAVOR) -> None:
    """Build the Legendre tables on the GPU (src/legmod.f90:30)."""
    keep = [_opt(x) for x in (sid, gwd, skgpsp, skspgp)]
    ptrs = [None if x is None else _cd(x) for x in keep]
    check(lib().exo_legini(ntru, nlat, nlev, *ptrs, plavor), "legini")
    _state.update(ntru=ntru, nlat=nlat, nlev=nlev)


def legini_tables() -> np.ndarray:
    nlat, _, ncsp = dims(_need())
    t = np.zeros((8, nlat, ncsp))
    check(lib().exo_legini_tables(_cd(t)), "legini_tables")
    return t


def _need() -> int:
    if _state["ntru"] is None:
        raise ExoError("legini has not been called")
    return _state["ntru"]


def _suffix(a: np.ndarray) -> str:
    if a.dtype == np.float64:
        return "f64"
    if a.dtype == np.float32:
        return "f32"
    raise TypeError("float64 or float32 arrays expected")


def _call(name, *arrs, extra=()):
    fn = getattr(lib(), f"exo_{name}_{_suffix(arrs[0])}")
    check(fn(*[_cd(a) for a in arrs], *extra), name)


def sp2fc(sp: np.ndarray) -> np.ndarray:
    """sp(2,NCSP[,nfl]) -> fc(2,NLON/2,NLAT[,nfl]) (src/legmod.f90:230)."""
    ntru = _need()
    nlat, nlon, ncsp = dims(ntru)
    sp = np.ascontiguousarray(sp)
    nfl = sp.size // (2 * ncsp)
    fc = np.zeros((nfl, nlat, nlon), dtype=sp.dtype)
    _call("sp2fc", sp, fc, extra=(nfl,))
    return fc


def sp2fcdmu(sp: np.ndarray) -> np.ndarray:
    """Fourier coefficients of (1-mu^2) d/dmu (src/legmod.f90:262)."""
    ntru = _need()
    nlat, nlon, ncsp = dims(ntru)
    sp = np.ascontiguousarray(sp)
    nfl = sp.size // (2 * ncsp)
    fc = np.zeros((nfl, nlat, nlon), dtype=sp.dtype)
    _call("sp2fcdmu", sp, fc, extra=(nfl,))
    return fc


def fc2sp(fc: np.ndarray) -> np.ndarray:
    """fc(2,NLON/2,NLAT[,nfl]) -> sp(2,NCSP[,nfl]) (src/legmod.f90:178)."""
    ntru = _need()
    nlat, nlon, ncsp = dims(ntru)
    fc = np.ascontiguousarray(fc)
    nfl = fc.size // (nlat * nlon)
    sp = np.zeros((nfl, 2 * ncsp), dtype=fc.dtype)
    _call("fc2sp", fc, sp, extra=(nfl,))
    return sp


def dv2uv(pd: np.ndarray, pz: np.ndarray):
    """(divergence, absolute vorticity) -> (u cos, v cos) Fourier
    coefficients over NLEV levels (src/legmod.f90:312)."""
    ntru = _need()
    nlat, nlon, ncsp = dims(ntru)
    nlev = _state["nlev"]
    pd = np.ascontiguousarray(pd)
    pz = np.ascontiguousarray(pz, dtype=pd.dtype)
    for name, a in (("pd", pd), ("pz", pz)):
        if a.size != nlev * 2 * ncsp:
            raise ExoError(f"dv2uv: {name} holds {a.size} values, NLEV*NRSP = {nlev * 2 * ncsp}")
    pu = np.zeros((nlev, nlat, nlon), dtype=pd.dtype)
    pv = np.zeros((nlev, nlat, nlon), dtype=pd.dtype)
    _call("dv2uv", pd, pz, pu, pv)
    return pu, pv


def _levels(*arrs):
    ntru = _need()
    nlat, nlon, ncsp = dims(ntru)
    nlev = _state["nlev"]
    dt = np.asarray(arrs[0]).dtype
    return [np.ascontiguousarray(a, dtype=dt).reshape(nlev, nlat, nlon) for a in arrs], nlev, ncsp, dt


def uv2dv(pu: np.ndarray, pv: np.ndarray):
    """(u cos, v cos) Fourier coefficients -> (divergence, vorticity)
    spectral coefficients over NLEV levels (src/legmod.f90:356)."""
    (pu, pv), nlev, ncsp, dt = _levels(pu, pv)
    pd = np.zeros((nlev, 2 * ncsp), dtype=dt)
    pz = np.zeros((nlev, 2 * ncsp), dtype=dt)
    _call("uv2dv", pu, pv, pd, pz)
    return pd, pz


def qtend(qn, uq, vq):
    """Tracer tendency q (src/legmod.f90:433)."""
    (qn, uq, vq), nlev, ncsp, dt = _levels(qn, uq, vq)
    q = np.zeros((nlev, 2 * ncsp), dtype=dt)
    _call("qtend", q, qn, uq, vq)
    return q


def mktend(tn, fu, fv, ke, ut, vt):
    """Spectral tendencies (d, t, z) of the adiabatic terms
    (src/legmod.f90:505)."""
    ins, nlev, ncsp, dt = _levels(tn, fu, fv, ke, ut, vt)
    d, t, z = (np.zeros((nlev, 2 * ncsp), dtype=dt) for _ in range(3))
    _call("mktend", d, t, z, *ins)
    return d, t, z


def gp2fc(a: np.ndarray) -> np.ndarray:
    """Forward real FFT of rows, returns the transformed copy
    (src/fftmod.f90:29)."""
    a = np.array(a, order="C")
    n = a.shape[-1]
    _call("gp2fc", a, extra=(n, a.size // n))
    return a


def fc2gp(a: np.ndarray) -> np.ndarray:
    """Inverse real FFT of rows, returns the transformed copy
    (src/fftmod.f90:64)."""
    a = np.array(a, order="C")
    n = a.shape[-1]
    _call("fc2gp", a, extra=(n, a.size // n))
    return a


# ---------------------------------------------------------------------------
# device-resident batched layer
# ---------------------------------------------------------------------------
def _ptr(t) -> int:
    if t is None:
        return None
    if not t.is_cuda:
        raise ExoError("device tensors expected")
    if not t.is_contiguous():
        raise ExoError("contiguous tensors expected")
    return t.data_ptr()


def _stream():
    import torch

    return torch.cuda.current_stream().cuda_stream


def _fields(what, t, dtype, tail, nmem, *runs):
    """Check one buffer of a batched call before its address reaches a
    kernel: element type, rank, trailing shape, at least nmem members, and
    every run (first, count) of field-levels inside shape[1]."""
    if t.dtype != dtype:
        raise ExoError(f"{what}: {t.dtype} buffer, the mode computes on {dtype}")
    if t.dim() != 2 + len(tail) or tuple(t.shape[2:]) != tuple(tail):
        raise ExoError(f"{what}: shape {tuple(t.shape)}, expected (nmem, nfield, "
                       f"{', '.join(str(x) for x in tail)})")
    if nmem < 0 or nmem > t.shape[0]:
        raise ExoError(f"{what}: {nmem} members requested, buffer holds {t.shape[0]}")
    for first, count in runs:
        if first < 0 or count < 0 or first + count > t.shape[1]:
            raise ExoError(f"{what}: field-levels [{first}, {first + count}) outside "
                           f"the buffer's {t.shape[1]}")


class Plan:
    """Transforms of one truncation with FP64 tables built on the GPU."""

    def __init__(self, ntru: int, sid=None, gwd=None, skgpsp=None, skspgp=None,
                 plavor: float = PLAVOR):
        keep = [_opt(x) for x in (sid, gwd, skgpsp, skspgp)]
        ptrs = [None if x is None else _cd(x) for x in keep]
        h = c_void_p()
        check(lib().exo_sht_plan_create(byref(h), ntru, *ptrs, plavor), "plan_create")
        self._h = h
        d = (c_int * 4)()
        check(lib().exo_sht_plan_dims(h, d), "plan_dims")
        self.ntru, self.nlat, self.nlon, self.ncsp = list(d)
        self.nrsp = 2 * self.ncsp

    def close(self):
        if getattr(self, "_h", None):
            lib().exo_sht_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def grid(self):
        sid = np.zeros(self.nlat)
        gwd = np.zeros(self.nlat)
        check(lib().exo_sht_plan_grid(self._h, _cd(sid), _cd(gwd)), "plan_grid")
        return sid, gwd

    def tables(self) -> np.ndarray:
        t = np.zeros((8, self.nlat, self.ncsp))
        check(lib().exo_sht_plan_tables(self._h, _cd(t)), "plan_tables")
        return t

    @staticmethod
    def dtype_of(mode):
        import torch

        return torch.float64 if mode_id(mode) == 0 else torch.float32

    def inverse(self, mode, sp, gp, nfield, src0=0, dst0=0, table="qi", nmem=None):
        """sp2fc (table qi) or sp2fcdmu (qj) + fc2gp of field-levels
        [src0, src0+nfield) of every member into grid fields [dst0, ...).
        sp: (nmem, nfl, nrsp), gp: (nmem, ngp, nlat, nlon)."""
        nmem = sp.shape[0] if nmem is None else nmem
        if table not in ("qi", "qj"):
            raise ExoError(f"inverse: table {table!r}, expected 'qi' or 'qj'")
        dt = self.dtype_of(mode)
        _fields("inverse sp", sp, dt, (self.nrsp,), nmem, (src0, nfield))
        _fields("inverse gp", gp, dt, (self.nlat, self.nlon), nmem, (dst0, nfield))
        check(lib().exo_sht_inverse(self._h, mode_id(mode), TABLES.index(table), nmem,
                                    _ptr(sp), sp[0].numel(), src0, _ptr(gp), gp[0].numel(),
                                    dst0, nfield, _stream()), "inverse")

    def inverse_uv(self, mode, sp, gp, nlev, vor0, div0, u0, v0, nmem=None):
        nmem = sp.shape[0] if nmem is None else nmem
        dt = self.dtype_of(mode)
        _fields("inverse_uv sp", sp, dt, (self.nrsp,), nmem, (vor0, nlev), (div0, nlev))
        _fields("inverse_uv gp", gp, dt, (self.nlat, self.nlon), nmem, (u0, nlev), (v0, nlev))
        check(lib().exo_sht_inverse_uv(self._h, mode_id(mode), nmem, _ptr(sp), sp[0].numel(),
                                       vor0, div0, _ptr(gp), gp[0].numel(), u0, v0, nlev,
                                       _stream()), "inverse_uv")

    def direct(self, mode, gp, sp, nfield, src0=0, dst0=0, nmem=None):
        """gp2fc + fc2sp of grid fields [src0, src0+nfield) into spectral
        field-levels [dst0, ...)."""
        nmem = gp.shape[0] if nmem is None else nmem
        dt = self.dtype_of(mode)
        _fields("direct gp", gp, dt, (self.nlat, self.nlon), nmem, (src0, nfield))
        _fields("direct sp", sp, dt, (self.nrsp,), nmem, (dst0, nfield))
        check(lib().exo_sht_direct(self._h, mode_id(mode), nmem, _ptr(gp), gp[0].numel(),
                                   src0, _ptr(sp), sp[0].numel(), dst0, nfield, _stream()),
              "direct")

    def _roundtrip_buffers(self, what, mode, nlev, sp_in, gp, sp_out=None):
        """(nmem, 3*nlev+1, nrsp) spectral and (nmem, 5*nlev+2, nlat, nlon)
        grid buffers of one round trip, all with the same nmem."""
        if nlev < 1:
            raise ExoError(f"{what}: nlev = {nlev}")
        dt = self.dtype_of(mode)
        nfl = 3 * nlev + 1
        nmem = sp_in.shape[0] if sp_in.dim() else -1
        bufs = [("sp_in", sp_in, (self.nrsp,), nfl), ("gp", gp, (self.nlat, self.nlon), nfl + 1 + 2 * nlev)]
        if sp_out is not None:
            bufs.append(("sp_out", sp_out, (self.nrsp,), nfl))
        for name, t, tail, nf in bufs:
            _fields(f"{what} {name}", t, dt, tail, nmem, (0, nf))
            if t.shape[0] != nmem or t.shape[1] != nf:
                raise ExoError(f"{what} {name}: shape {tuple(t.shape)}, expected "
                               f"({nmem}, {nf}, {', '.join(str(x) for x in tail)})")
        return nmem

    def roundtrip(self, mode, sp_in, gp, sp_out, nlev):
        """One round trip of every member (see exo_sht_roundtrip)."""
        nmem = self._roundtrip_buffers("roundtrip", mode, nlev, sp_in, gp, sp_out)
        check(lib().exo_sht_roundtrip(self._h, mode_id(mode), nmem, nlev,
                                      _ptr(sp_in), _ptr(gp), _ptr(sp_out), _stream()),
              "roundtrip")

    def roundtrip_noalloc(self, mode, sp_in, gp, sp_out, nlev):
        """The same round trip with no device allocation: the scratch is the
        library's fixed per-GPU array (exo_sht_roundtrip_noalloc)."""
        nmem = self._roundtrip_buffers("roundtrip_noalloc", mode, nlev, sp_in, gp, sp_out)
        check(lib().exo_sht_roundtrip_noalloc(self._h, mode_id(mode), nmem, nlev,
                                              _ptr(sp_in), _ptr(gp), _ptr(sp_out), _stream()),
              "roundtrip_noalloc")

    def inverse_all(self, mode, sp_in, gp, nlev):
        """The inverse half of a round trip in one launch (exo_sht_inverse_all)."""
        nmem = self._roundtrip_buffers("inverse_all", mode, nlev, sp_in, gp)
        check(lib().exo_sht_inverse_all(self._h, mode_id(mode), nmem, nlev,
                                        _ptr(sp_in), _ptr(gp), _stream()), "inverse_all")

    # ---- Fourier-space forms (no fused FFT): Fourier arrays have the grid's
    # shape (nmem, nfield, nlat, nlon) and hold fc(2, NLON/2, NLAT) per field
    def sp2fc(self, mode, sp, fc, nfield, src0=0, dst0=0, table="qi", nmem=None):
        """sp2fc (qi) or sp2fcdmu (qj) of field-levels [src0, src0+nfield)
        into Fourier fields [dst0, ...) (exo_sht_sp2fc)."""
        nmem = sp.shape[0] if nmem is None else nmem
        if table not in ("qi", "qj"):
            raise ExoError(f"sp2fc: table {table!r}, expected 'qi' or 'qj'")
        dt = self.dtype_of(mode)
        _fields("sp2fc sp", sp, dt, (self.nrsp,), nmem, (src0, nfield))
        _fields("sp2fc fc", fc, dt, (self.nlat, self.nlon), nmem, (dst0, nfield))
        check(lib().exo_sht_sp2fc(self._h, mode_id(mode), TABLES.index(table), nmem, _ptr(sp),
                                  sp[0].numel(), src0, _ptr(fc), fc[0].numel(), dst0, nfield,
                                  _stream()), "sp2fc")

    def dv2uv(self, mode, sp, fc, nlev, vor0, div0, u0, v0, nmem=None):
        nmem = sp.shape[0] if nmem is None else nmem
        dt = self.dtype_of(mode)
        _fields("dv2uv sp", sp, dt, (self.nrsp,), nmem, (vor0, nlev), (div0, nlev))
        _fields("dv2uv fc", fc, dt, (self.nlat, self.nlon), nmem, (u0, nlev), (v0, nlev))
        check(lib().exo_sht_dv2uv(self._h, mode_id(mode), nmem, _ptr(sp), sp[0].numel(), vor0,
                                  div0, _ptr(fc), fc[0].numel(), u0, v0, nlev, _stream()),
              "dv2uv")

    def fc2sp(self, mode, fc, sp, nfield, src0=0, dst0=0, nmem=None):
        nmem = fc.shape[0] if nmem is None else nmem
        dt = self.dtype_of(mode)
        _fields("fc2sp fc", fc, dt, (self.nlat, self.nlon), nmem, (src0, nfield))
        _fields("fc2sp sp", sp, dt, (self.nrsp,), nmem, (dst0, nfield))
        check(lib().exo_sht_fc2sp(self._h, mode_id(mode), nmem, _ptr(fc), fc[0].numel(), src0,
                                  _ptr(sp), sp[0].numel(), dst0, nfield, _stream()), "fc2sp")

    def fft(self, mode, a, direction):
        """In-place gp2fc (direction -1) or fc2gp (+1) of every row of a
        (any shape whose last dimension is nlon)."""
        if a.dtype != self.dtype_of(mode) or a.dim() < 1 or a.shape[-1] != self.nlon:
            raise ExoError(f"fft: {a.dtype} {tuple(a.shape)}, expected rows of {self.nlon} "
                           f"{self.dtype_of(mode)}")
        check(lib().exo_sht_fft(self._h, mode_id(mode), int(direction), _ptr(a),
                                a.numel() // self.nlon, _stream()), "fft")

    def _multi(self, what, mode, nlev, fourier, spectral):
        dt = self.dtype_of(mode)
        nmem = fourier[0].shape[0]
        for t in fourier:
            _fields(f"{what} Fourier input", t, dt, (self.nlat, self.nlon), nmem, (0, nlev))
            if t.shape != fourier[0].shape:
                raise ExoError(f"{what}: Fourier inputs of different shapes")
        for t in spectral:
            _fields(f"{what} spectral output", t, dt, (self.nrsp,), nmem, (0, nlev))
            if t.shape != spectral[0].shape:
                raise ExoError(f"{what}: spectral outputs of different shapes")
        return nmem, fourier[0][0].numel(), spectral[0][0].numel()

    def uv2dv(self, mode, pu, pv, pd, pz, nlev):
        """uv2dv of Fourier u, v (nmem, >=nlev, nlat, nlon) into spectral D,
        zeta (nmem, >=nlev, nrsp) (exo_sht_uv2dv)."""
        nmem, fs, ss = self._multi("uv2dv", mode, nlev, (pu, pv), (pd, pz))
        check(lib().exo_sht_uv2dv(self._h, mode_id(mode), nmem, nlev, _ptr(pu), _ptr(pv), fs,
                                  _ptr(pd), _ptr(pz), ss, _stream()), "uv2dv")

    def qtend(self, mode, q, qn, uq, vq, nlev):
        nmem, fs, ss = self._multi("qtend", mode, nlev, (qn, uq, vq), (q,))
        check(lib().exo_sht_qtend(self._h, mode_id(mode), nmem, nlev, _ptr(q), ss, _ptr(qn),
                                  _ptr(uq), _ptr(vq), fs, _stream()), "qtend")

    def mktend(self, mode, d, t, z, tn, fu, fv, ke, ut, vt, nlev):
        nmem, fs, ss = self._multi("mktend", mode, nlev, (tn, fu, fv, ke, ut, vt), (d, t, z))
        check(lib().exo_sht_mktend(self._h, mode_id(mode), nmem, nlev, _ptr(d), _ptr(t),
                                   _ptr(z), ss, _ptr(tn), _ptr(fu), _ptr(fv), _ptr(ke),
                                   _ptr(ut), _ptr(vt), fs, _stream()), "mktend")

    def set_tiles(self, mode, inv_var=0, uv_var=0, fwd_var=0):
        check(lib().exo_sht_set_tiles(self._h, mode_id(mode), inv_var, uv_var, fwd_var),
              "set_tiles")

    def tiles(self, mode):
        v = (c_int * 3)()
        check(lib().exo_sht_get_tiles(self._h, mode_id(mode), v), "get_tiles")
        return tuple(v)

    def autotune(self, mode, nmem: int, nlev: int):
        """Time the tile variants of the three fused kernels and keep the
        fastest (returns (inverse, inverse_uv, direct) variants)."""
        v = (c_int * 3)()
        check(lib().exo_sht_autotune(self._h, mode_id(mode), nmem, nlev, _stream(), v),
              "autotune")
        return tuple(v)

    def alloc_roundtrip(self, nmem: int, nlev: int, mode):
        """Zeroed device buffers (sp_in, gp, sp_out) of a round trip."""
        import torch

        dt = self.dtype_of(mode)
        nfl = 3 * nlev + 1
        ngp = nfl + 1 + 2 * nlev
        dev = torch.device("cuda")
        sp_in = torch.zeros((nmem, nfl, self.nrsp), dtype=dt, device=dev)
        gp = torch.zeros((nmem, ngp, self.nlat, self.nlon), dtype=dt, device=dev)
        sp_out = torch.zeros((nmem, nfl, self.nrsp), dtype=dt, device=dev)
        return sp_in, gp, sp_out


class _EnsGroup(ctypes.Structure):
    _fields_ = [("ntru", c_int), ("nlev", c_int), ("nmem", c_int), ("sp_in", c_void_p),
                ("gp", c_void_p), ("sp_out", c_void_p)]


class Ensemble:
    """Round trips of an ensemble of members with different truncations.

    groups: list of (ntru, nlev, nmem) with ntru in 21, 42, 63.  The device
    buffers of each group are allocated here in the layout of
    Plan.roundtrip (``self.bufs[k] = (sp_in, gp, sp_out)``); all round trips
    run as two persistent kernels captured in one hipGraph, or as one
    batched round trip per group in that graph (set_schedule; autotune picks
    the faster) (include/exo_sht.h, exo_sht_ens_*).  mode: "f32", "f32mfma"
    or "bf16x3".
    """

    def __init__(self, groups, mode="bf16x3"):
        import torch

        self.mode = mode
        self.groups = [tuple(int(x) for x in g) for g in groups]
        self.bufs = []
        arr = (_EnsGroup * len(self.groups))()
        dev = torch.device("cuda")
        for k, (ntru, nlev, nmem) in enumerate(self.groups):
            nlat, nlon, ncsp = dims(ntru)
            nfl = 3 * nlev + 1
            ngp = nfl + 1 + 2 * nlev
            b = (torch.zeros((nmem, nfl, 2 * ncsp), dtype=torch.float32, device=dev),
                 torch.zeros((nmem, ngp, nlat, nlon), dtype=torch.float32, device=dev),
                 torch.zeros((nmem, nfl, 2 * ncsp), dtype=torch.float32, device=dev))
            self.bufs.append(b)
            arr[k] = _EnsGroup(ntru, nlev, nmem, _ptr(b[0]), _ptr(b[1]), _ptr(b[2]))
        torch.cuda.synchronize()
        h = c_void_p()
        check(lib().exo_sht_ens_create(byref(h), mode_id(mode), len(self.groups), arr),
              "ens_create")
        self._h = h

    @property
    def members(self) -> int:
        return sum(g[2] for g in self.groups)

    def close(self):
        if getattr(self, "_h", None):
            lib().exo_sht_ens_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def run(self):
        """One launch of the captured graph (all round trips)."""
        check(lib().exo_sht_ens_run(self._h, _stream()), "ens_run")

    def run_eager(self):
        check(lib().exo_sht_ens_run_eager(self._h, _stream()), "ens_run_eager")

    def run_groups(self):
        check(lib().exo_sht_ens_run_groups(self._h, _stream()), "ens_run_groups")

    def run_members(self):
        check(lib().exo_sht_ens_run_members(self._h, _stream()), "ens_run_members")

    def tiles(self):
        v = (c_int * (2 * len(self.groups)))()
        check(lib().exo_sht_ens_get_tiles(self._h, v), "ens_get_tiles")
        return tuple(v)

    def set_tiles(self, vars):
        v = (c_int * (2 * len(self.groups)))(*vars)
        check(lib().exo_sht_ens_set_tiles(self._h, v), "ens_set_tiles")

    def autotune(self):
        """Choose the tile variants, then the faster schedule (persistent
        kernels or batched round trips) for the graph; returns the variants."""
        v = (c_int * (2 * len(self.groups)))()
        check(lib().exo_sht_ens_autotune(self._h, _stream(), v), "ens_autotune")
        return tuple(v)

    PERSISTENT, BATCHED = 0, 1

    def schedule(self) -> str:
        """What the graph holds: "persistent" (the two persistent kernels) or
        "batched" (one batched round trip per group)."""
        v = c_int()
        check(lib().exo_sht_ens_get_schedule(self._h, byref(v)), "ens_get_schedule")
        return "batched" if v.value == self.BATCHED else "persistent"

    def set_schedule(self, sched: str):
        """Capture the graph again with "persistent" or "batched" (the latter
        after one direct pass, which writes the outputs)."""
        code = {"persistent": self.PERSISTENT, "batched": self.BATCHED}.get(sched)
        if code is None:
            raise ValueError(f"schedule {sched!r}: 'persistent' or 'batched'")
        import torch

        torch.cuda.synchronize()
        check(lib().exo_sht_ens_set_schedule(self._h, code), "ens_set_schedule")
