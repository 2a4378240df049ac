"""Long-wave radiation on the GPU: the reference column and its learned emulator.

* ``LwrColumn``: the reference's long-wave column (src/radmod.f90:2122-2456,
  subroutine lwr) with its heating rate (src/radmod.f90:1100-1108) as a HIP
  kernel, one thread per column (include/exo_rad.h: exo_lwr).  It is the
  emulator's training-data generator, its fallback and the reference for
  its heating-rate error.
* ``LwEmulator``: a 43 -> 256 -> 256 -> 256 -> 21 GELU MLP whose forward and
  backward passes run on this package's own BF16 MFMA kernels
  (exo_mlp_forward / exo_mlp_backward_dx / exo_mlp_backward_dw), registered
  as torch operators (exoplasim_amd::lw_mlp_forward / lw_mlp_backward) with
  autograd; FP32 master weights, Adam.  Inputs per column: 20-level
  temperature, 20-level specific humidity, surface temperature, surface
  pressure, CO2; outputs: 20-level heating rates (K/day) and OLR (W/m2).

Column arrays follow the reference layout field(NHOR, lev) (column index
fastest): torch tensors of shape (nlev, ncol).
"""
from __future__ import annotations

import ctypes
import math
from ctypes import POINTER, c_double, c_int, c_void_p

import numpy as np

from ._lib import ExoError, lib

GA = 9.80665                 # src/p_earth.f90:58
GASCON = 287.0               # src/p_earth.f90:43
AKAP = 0.286                 # src/p_earth.f90:41
ACPD = GASCON / AKAP         # src/plasim.f90:1317
ADV = 1870.0 / ACPD - 1.0    # src/plasim.f90:1318, ACPV = 1870 (src/plasimmod.f90:122)
LWR_DEFAULTS = dict(th2oc=0.024, tpofmt=1.0, acllwr=0.1, clgray=-1.0, npbroaden=1)


class _Params(ctypes.Structure):
    _fields_ = [("nlev", c_int), ("sigma", c_void_p), ("sigmah", c_void_p), ("dsigma", c_void_p),
                ("ga", c_double), ("th2oc", c_double), ("tpofmt", c_double),
                ("acllwr", c_double), ("clgray", c_double), ("npbroaden", c_int),
                ("acpd", c_double), ("adv", c_double)]


class _Cols(ctypes.Structure):
    _fields_ = [(k, c_void_p) for k in ("dp", "dt", "dq", "dqo3", "dqco2", "dcc", "dql", "dls")]


class _Out(ctypes.Structure):
    _fields_ = [(k, c_void_p) for k in ("dftu", "dftd", "heat", "olr")]


class _Net(ctypes.Structure):
    _fields_ = [("nlev", c_int), ("w", c_void_p * 4), ("b", c_void_p * 4), ("ymean", c_void_p),
                ("ystd", c_void_p)]


def _proto():
    L = lib()
    if not getattr(L, "_rad_proto", False):
        L.exo_lwr.restype = c_int
        L.exo_lwr.argtypes = [POINTER(_Params), c_int, c_int, POINTER(_Cols), POINTER(_Out),
                              c_void_p]
        vp, fp = c_void_p, c_void_p
        L.exo_mlp_forward.restype = c_int
        L.exo_mlp_forward.argtypes = [vp, vp, fp, c_int, c_int, c_int, c_int, c_int, fp, vp, fp,
                                      vp]
        L.exo_mlp_backward_dx.restype = c_int
        L.exo_mlp_backward_dx.argtypes = [vp, vp, fp, c_int, c_int, c_int, c_int, vp, fp, vp]
        L.exo_mlp_backward_dw.restype = c_int
        L.exo_mlp_backward_dw.argtypes = [vp, vp, fp, c_int, c_int, c_int, c_int, fp, fp, fp, vp]
        L.exo_mlp_cast_weights.restype = c_int
        L.exo_mlp_cast_weights.argtypes = [fp, c_int, c_int, c_int, c_int, vp, vp, vp]
        L.exo_mlp_cast_rows.restype = c_int
        L.exo_mlp_cast_rows.argtypes = [fp, c_int, c_int, c_int, vp, vp]
        L.exo_lw_emulate.restype = c_int
        L.exo_lw_emulate.argtypes = [POINTER(_Net), c_int, fp, fp, fp, fp, fp, fp, vp]
        L.exo_lw_emulate_guarded.restype = c_int
        L.exo_lw_emulate_guarded.argtypes = [POINTER(_Net), fp, fp, c_int, fp, fp, fp, fp, fp,
                                             fp, vp, vp]
        L._rad_proto = True
    return L


def _chk(rc, what):
    if rc != 0:
        raise ExoError(f"{what} failed ({rc})")


def _stream():
    import torch

    return torch.cuda.current_stream().cuda_stream


def sigma_levels(nlev: int):
    """sigma, sigmah, dsigma of the neqsig = 0 setup (src/plasim.f90:1553-1567)."""
    zsk = np.arange(1, nlev + 1) / nlev
    sigmah = 0.75 * zsk + 1.75 * zsk ** 3 - 1.5 * zsk ** 4
    dsigma = np.diff(np.concatenate([[0.0], sigmah]))
    sigma = np.concatenate([[0.5 * sigmah[0]], 0.5 * (sigmah[:-1] + sigmah[1:])])
    return sigma, sigmah, dsigma


class LwrColumn:
    """GPU long-wave column with the reference's physics."""

    def __init__(self, nlev: int = 20, fp64: bool = False, **params):
        import torch

        p = dict(LWR_DEFAULTS)
        p.update(params)
        self.nlev, self.fp64 = nlev, fp64
        self.dtype = torch.float64 if fp64 else torch.float32
        s, sh, ds = sigma_levels(nlev)
        dev = torch.device("cuda")
        self._lev = [torch.tensor(x, dtype=torch.float64, device=dev) for x in (s, sh, ds)]
        self._p = _Params(nlev, self._lev[0].data_ptr(), self._lev[1].data_ptr(),
                          self._lev[2].data_ptr(), GA, p["th2oc"], p["tpofmt"], p["acllwr"],
                          p["clgray"], int(p["npbroaden"]), ACPD, ADV)

    def __call__(self, cols: dict, fluxes: bool = False):
        """cols: torch tensors (nlev[+1], ncol) dt, dq, dqo3, dqco2, dcc, dql and
        (ncol) dp, dls on the GPU.  Returns heating (nlev, ncol) in K/day and
        OLR (ncol) [and dftu, dftd]."""
        import torch

        c = {k: v.to(self.dtype).contiguous() for k, v in cols.items()}
        ncol = c["dp"].shape[0]
        heat = torch.empty((self.nlev, ncol), dtype=self.dtype, device=c["dp"].device)
        olr = torch.empty(ncol, dtype=self.dtype, device=c["dp"].device)
        fu = fd = None
        if fluxes:
            fu = torch.empty((self.nlev + 1, ncol), dtype=self.dtype, device=c["dp"].device)
            fd = torch.empty_like(fu)
        ci = _Cols(*[c[k].data_ptr() for k in ("dp", "dt", "dq", "dqo3", "dqco2", "dcc", "dql",
                                                  "dls")])
        co = _Out(fu.data_ptr() if fu is not None else None,
                  fd.data_ptr() if fd is not None else None, heat.data_ptr(), olr.data_ptr())
        _chk(_proto().exo_lwr(ctypes.byref(self._p), int(self.fp64), ncol, ctypes.byref(ci),
                              ctypes.byref(co), _stream()), "exo_lwr")
        return (heat, olr, fu, fd) if fluxes else (heat, olr)


def columns_to_device(cols: dict, dtype=None):
    """numpy columns [col][lev] (exoplasim_amd.columns.make_columns) -> torch
    tensors (lev, col) on the GPU."""
    import torch

    dtype = dtype or torch.float64
    out = {}
    for k, v in cols.items():
        a = np.asarray(v)
        t = torch.from_numpy(np.ascontiguousarray(a.T if a.ndim == 2 else a))
        out[k] = t.to(device="cuda", dtype=dtype)
    return out


# ---------------------------------------------------------------------------
# emulator
# ---------------------------------------------------------------------------
NIN, NHID, NOUT = 43, 256, 21
KIN, NOUTP = 64, 32  # padded sizes of the input and output layers
LAYERS = [(NIN, NHID), (NHID, NHID), (NHID, NHID), (NHID, NOUT)]
PADDED = [(KIN, NHID), (NHID, NHID), (NHID, NHID), (NHID, NOUTP)]


def features(cols: dict, nlev: int = 20):
    """(ncol, 43) fp32 network inputs from device columns (lev, col):
    T (K) scaled, ln q scaled, surface T, ln ps, ln CO2."""
    import torch

    t = cols["dt"].float()
    f = [(t[:nlev] - 250.0) / 50.0,
         (torch.log(cols["dq"].float().clamp_min(1e-12)) + 10.0) / 3.5,
         ((t[nlev] - 265.0) / 40.0)[None, :],
         (torch.log(cols["dp"].float() / 1e5) / 1.3)[None, :],
         ((torch.log(cols["dqco2"][0].float()) - 5.75) / 3.3)[None, :]]
    return torch.cat(f, dim=0).t().contiguous()


class _MLP:
    """Buffers and raw kernel calls of one batch size."""

    def __init__(self, M: int, device):
        import torch

        bf = torch.bfloat16
        self.M = M
        self.x0 = torch.empty((M, KIN), dtype=bf, device=device)
        self.z = [torch.empty((M, NHID), dtype=torch.float32, device=device) for _ in range(3)]
        self.h = [torch.empty((M, NHID), dtype=bf, device=device) for _ in range(3)]
        self.y = torch.empty((M, NOUTP), dtype=torch.float32, device=device)
        self.dz = [torch.empty((M, NHID), dtype=bf, device=device) for _ in range(3)]
        self.dz32 = [torch.empty((M, NHID), dtype=torch.float32, device=device) for _ in range(3)]
        self.dy = torch.empty((M, NOUTP), dtype=bf, device=device)
        self.nsplit = max(1, min(64, M // 256))
        self.part = torch.empty(self.nsplit * NHID * NHID, dtype=torch.float32, device=device)


def _ptr(t):
    return t.data_ptr()


def _register_ops():
    """The network's forward and backward passes as torch operators
    (torch.library custom ops ``exoplasim_amd::lw_mlp_forward`` and
    ``exoplasim_amd::lw_mlp_backward``) on this package's HIP kernels, with
    autograd and shape functions registered: the returned function
    ``(x, W1, b1, ..., W4, b4) -> y`` is differentiable in the parameters."""
    import torch

    def _weights(params, dev):
        L, s = _proto(), _stream()
        wb, wtb = [], []
        for li, ((nin, nout), (K, N)) in enumerate(zip(LAYERS, PADDED)):
            w = params[li].contiguous()
            W = torch.empty((N, K), dtype=torch.bfloat16, device=dev)
            Wt = torch.empty((K, N), dtype=torch.bfloat16, device=dev)
            _chk(L.exo_mlp_cast_weights(_ptr(w), nout, nin, N, K, _ptr(W), _ptr(Wt), s),
                 "cast_weights")
            wb.append(W)
            wtb.append(Wt)
        return wb, wtb

    t8 = "(" + ", ".join(["Tensor"] * 8) + ")"

    @torch.library.custom_op("exoplasim_amd::lw_mlp_forward", mutates_args=(),
                             schema="(Tensor x, Tensor w1, Tensor b1, Tensor w2, Tensor b2, "
                                    "Tensor w3, Tensor b3, Tensor w4, Tensor b4) -> " + t8)
    def lw_mlp_forward(x, w1, b1, w2, b2, w3, b3, w4, b4):
        """y (M, 21) and the activations the backward pass reads: the bf16
        inputs x0, the pre-activations z1..z3 (fp32) and the bf16 hidden
        layers h1..h3"""
        L, s = _proto(), _stream()
        bias = (b1, b2, b3, b4)
        M = x.shape[0]
        buf = _MLP(M, x.device)
        wb, _ = _weights((w1, w2, w3, w4), x.device)
        xf = x.float().contiguous()
        _chk(L.exo_mlp_cast_rows(_ptr(xf), M, NIN, KIN, _ptr(buf.x0), s), "cast_rows")
        inp = buf.x0
        for li in range(3):
            K, N = PADDED[li]
            _chk(L.exo_mlp_forward(_ptr(inp), _ptr(wb[li]), _ptr(bias[li]), M, K, N,
                                   LAYERS[li][1], 1, _ptr(buf.z[li]), _ptr(buf.h[li]), None,
                                   s), "mlp_forward")
            inp = buf.h[li]
        _chk(L.exo_mlp_forward(_ptr(inp), _ptr(wb[3]), _ptr(b4), M, NHID, NOUTP, NOUT, 0, None,
                               None, _ptr(buf.y), s), "mlp_forward")
        return (buf.y[:, :NOUT].clone(), buf.x0, *buf.z, *buf.h)

    @lw_mlp_forward.register_fake
    def _(x, w1, b1, w2, b2, w3, b3, w4, b4):
        M = x.shape[0]
        f32, bf = torch.float32, torch.bfloat16
        return (x.new_empty((M, NOUT), dtype=f32), x.new_empty((M, KIN), dtype=bf),
                *[x.new_empty((M, NHID), dtype=f32) for _ in range(3)],
                *[x.new_empty((M, NHID), dtype=bf) for _ in range(3)])

    @torch.library.custom_op("exoplasim_amd::lw_mlp_backward", mutates_args=(),
                             schema="(Tensor gy, Tensor x0, Tensor z1, Tensor z2, Tensor z3, "
                                    "Tensor h1, Tensor h2, Tensor h3, Tensor w1, Tensor w2, "
                                    "Tensor w3, Tensor w4) -> " + t8)
    def lw_mlp_backward(gy, x0, z1, z2, z3, h1, h2, h3, w1, w2, w3, w4):
        """gradients (dW1, db1, ..., dW4, db4) for the output gradient gy"""
        L, s = _proto(), _stream()
        M = gy.shape[0]
        dev = gy.device
        buf = _MLP(M, dev)
        z, h = (z1, z2, z3), (h1, h2, h3)
        _, wtb = _weights((w1, w2, w3, w4), dev)
        gy32 = torch.zeros((M, NOUTP), dtype=torch.float32, device=dev)
        gy32[:, :NOUT] = gy
        _chk(L.exo_mlp_cast_rows(_ptr(gy32), M, NOUTP, NOUTP, _ptr(buf.dy), s), "cast_rows")
        grads = [None] * 8
        dz, dz32 = buf.dy, gy32
        for li in (3, 2, 1, 0):
            K, N = PADDED[li]
            nin, nout = LAYERS[li]
            xin = h[li - 1] if li > 0 else x0
            dW = torch.empty((N, K), dtype=torch.float32, device=dev)
            db = torch.empty(N, dtype=torch.float32, device=dev)
            _chk(L.exo_mlp_backward_dw(_ptr(dz), _ptr(xin), _ptr(dz32), M, N, K, buf.nsplit,
                                       _ptr(buf.part), _ptr(dW), _ptr(db), s),
                 "mlp_backward_dw")
            grads[2 * li] = dW[:nout, :nin].contiguous()
            grads[2 * li + 1] = db[:nout].contiguous()
            if li > 0:
                Kp = PADDED[li - 1][1]
                _chk(L.exo_mlp_backward_dx(_ptr(dz), _ptr(wtb[li]), _ptr(z[li - 1]), M, N, Kp,
                                           LAYERS[li - 1][1], _ptr(buf.dz[li - 1]),
                                           _ptr(buf.dz32[li - 1]), s), "mlp_backward_dx")
                dz, dz32 = buf.dz[li - 1], buf.dz32[li - 1]
        return tuple(grads)

    @lw_mlp_backward.register_fake
    def _(gy, x0, z1, z2, z3, h1, h2, h3, w1, w2, w3, w4):
        out = []
        for w in (w1, w2, w3, w4):
            out += [torch.empty_like(w), w.new_empty((w.shape[0],))]
        return tuple(out)

    def setup_context(ctx, inputs, output):
        ctx.save_for_backward(*output[1:], inputs[1], inputs[3], inputs[5], inputs[7])

    def backward(ctx, gy, *unused):
        x0, z1, z2, z3, h1, h2, h3, w1, w2, w3, w4 = ctx.saved_tensors
        g = torch.ops.exoplasim_amd.lw_mlp_backward(gy.contiguous(), x0, z1, z2, z3, h1, h2, h3,
                                                    w1, w2, w3, w4)
        return (None, *g)

    torch.library.register_autograd("exoplasim_amd::lw_mlp_forward", backward,
                                    setup_context=setup_context)

    def apply(x, *params):
        return torch.ops.exoplasim_amd.lw_mlp_forward(x, *params)[0]

    return apply


_OP = None


def emulator_op():
    """the network as a differentiable function of (features, W1, b1, ...,
    W4, b4), through the registered torch operators"""
    global _OP
    if _OP is None:
        _OP = _register_ops()
    return _OP


class LwEmulator:
    """The learned long-wave column with FP32 master weights on the GPU."""

    def __init__(self, nlev: int = 20, seed: int = 0):
        import torch

        assert nlev == 20, "the network is sized for 20 levels"
        self.nlev = nlev
        g = torch.Generator(device="cpu").manual_seed(seed)
        self.params = []
        for nin, nout in LAYERS:
            w = torch.randn((nout, nin), generator=g) * math.sqrt(2.0 / nin)
            self.params += [w.cuda().requires_grad_(), torch.zeros(nout, device="cuda",
                                                                   requires_grad=True)]
        self.ymean = torch.zeros(NOUT, device="cuda")
        self.ystd = torch.ones(NOUT, device="cuda")
        # the training envelope in feature space (set by fit)
        self.flo = torch.full((NIN,), -float("inf"), device="cuda")
        self.fhi = torch.full((NIN,), float("inf"), device="cuda")
        self._net = None  # bf16 inference weights, rebuilt after training

    def targets(self, heat, olr, cols):
        """normalised targets: heating * ps/1e5 (removes the 1/ps of the
        heating-rate formula) and OLR."""
        import torch

        ps = cols["dp"].float() / 1e5
        return torch.cat([(heat.float() * ps[None, :]).t(), olr.float()[:, None]], dim=1)

    def raw(self, x):
        return emulator_op()(x, *self.params)

    def _freeze(self):
        """bf16 padded copies of the weights for the fused inference kernel"""
        import torch

        L = _proto()
        s = _stream()
        ws = []
        for (nin, nout), (K, N), w in zip(LAYERS, PADDED, self.params[0::2]):
            W = torch.empty((N, K), dtype=torch.bfloat16, device="cuda")
            Wt = torch.empty((K, N), dtype=torch.bfloat16, device="cuda")
            _chk(L.exo_mlp_cast_weights(_ptr(w.detach().contiguous()), nout, nin, N, K, _ptr(W),
                                        _ptr(Wt), s), "cast_weights")
            ws.append(W)
        bs = [b.detach().float().contiguous() for b in self.params[1::2]]
        ym, ys = self.ymean.float().contiguous(), self.ystd.float().contiguous()
        net = _Net(self.nlev, (c_void_p * 4)(*[_ptr(w) for w in ws]),
                   (c_void_p * 4)(*[_ptr(b) for b in bs]), _ptr(ym), _ptr(ys))
        self._net = (net, ws, bs, ym, ys)  # keeps the buffers alive
        self._net_ref = ctypes.byref(net)
        self._emulate = L.exo_lw_emulate

    def predict(self, cols, out=None):
        """heating (nlev, ncol) K/day and OLR (ncol) of device columns, in
        one kernel (exo_lw_emulate).  out=(heat, olr) reuses fp32 buffers."""
        import torch

        if self._net is None:
            self._freeze()

        def f32(t):
            return t if (t.dtype == torch.float32 and t.is_contiguous()) else t.float().contiguous()

        dt, dq, dp, co2 = f32(cols["dt"]), f32(cols["dq"]), f32(cols["dp"]), f32(cols["dqco2"])
        ncol = dp.shape[0]
        if out is None:
            heat = torch.empty((self.nlev, ncol), dtype=torch.float32, device=dp.device)
            olr = torch.empty(ncol, dtype=torch.float32, device=dp.device)
        else:
            heat, olr = out
        _chk(self._emulate(self._net_ref, ncol, dt.data_ptr(), dq.data_ptr(), dp.data_ptr(),
                           co2.data_ptr(), heat.data_ptr(), olr.data_ptr(), _stream()),
             "exo_lw_emulate")
        return heat, olr

    def predict_guarded(self, cols, column: "LwrColumn"):
        """heating and OLR like predict, with every column whose inputs lie
        outside the training envelope computed by the plain long-wave column
        instead (the reference's own lwr, src/radmod.f90:2122, and heating
        rate, :1107-1108).  Returns (heat, olr, number of plain columns)."""
        import torch

        if self._net is None:
            self._freeze()

        def f32(t):
            return t if (t.dtype == torch.float32 and t.is_contiguous()) else t.float().contiguous()

        dt, dq, dp, co2 = f32(cols["dt"]), f32(cols["dq"]), f32(cols["dp"]), f32(cols["dqco2"])
        ncol = dp.shape[0]
        heat = torch.empty((self.nlev, ncol), dtype=torch.float32, device=dp.device)
        olr = torch.empty(ncol, dtype=torch.float32, device=dp.device)
        outside = torch.empty(ncol, dtype=torch.int32, device=dp.device)
        flo, fhi = self.flo.float().contiguous(), self.fhi.float().contiguous()
        _chk(_proto().exo_lw_emulate_guarded(self._net_ref, flo.data_ptr(), fhi.data_ptr(), ncol,
                                             dt.data_ptr(), dq.data_ptr(), dp.data_ptr(),
                                             co2.data_ptr(), heat.data_ptr(), olr.data_ptr(),
                                             outside.data_ptr(), _stream()),
             "exo_lw_emulate_guarded")
        idx = torch.nonzero(outside, as_tuple=True)[0]
        n = int(idx.numel())
        if n:
            sub = {k: (v[..., idx] if v.dim() > 1 else v[idx]).contiguous()
                   for k, v in cols.items()}
            h, o = column(sub)
            heat[:, idx] = h.float()
            olr[idx] = o.float()
        return heat, olr, n

    def predict_layers(self, cols):
        """the same through the per-layer training kernels (features in torch)"""
        import torch

        with torch.no_grad():
            y = self.raw(features(cols, self.nlev)) * self.ystd + self.ymean
        ps = cols["dp"].float() / 1e5
        heat = (y[:, :self.nlev] / ps[:, None]).t().contiguous()
        return heat, y[:, self.nlev].contiguous()

    def fit(self, x, y, epochs=8, batch=8192, lr=2e-3, seed=1, log=None):
        """Adam on the normalised targets; x (n, 43), y (n, 21) device tensors."""
        import torch

        self._net = None
        # the envelope of the training inputs, widened by 2% of each range
        lo, hi = x.min(dim=0).values.float(), x.max(dim=0).values.float()
        pad = 0.02 * (hi - lo)
        self.flo, self.fhi = (lo - pad).contiguous(), (hi + pad).contiguous()
        self.ymean = y.mean(dim=0)
        self.ystd = y.std(dim=0).clamp_min(1e-6)
        yn = (y - self.ymean) / self.ystd
        n = x.shape[0]
        opt = torch.optim.Adam(self.params, lr=lr)
        steps = epochs * (n // batch)
        sched = torch.optim.lr_scheduler.OneCycleLR(opt, max_lr=lr, total_steps=steps,
                                                    pct_start=0.1)
        g = torch.Generator(device="cuda").manual_seed(seed)
        it = 0
        for ep in range(epochs):
            perm = torch.randperm(n, device="cuda", generator=g)
            tot = 0.0
            for b in range(n // batch):
                idx = perm[b * batch:(b + 1) * batch]
                out = self.raw(x[idx])
                loss = torch.mean((out - yn[idx]) ** 2)
                opt.zero_grad(set_to_none=True)
                loss.backward()
                opt.step()
                sched.step()
                tot += float(loss.detach()) if b % 50 == 0 else 0.0
                it += 1
            if log:
                log(f"epoch {ep}: loss {float(loss.detach()):.4e}")
        return float(loss.detach())
