"""Pins of the emulator network oracle (oracle/mlp_oracle.py) on the CPU.

The oracle's calculus is checked with its rounding switched off against
torch's float64 autograd (an independent implementation) and against
central finite differences; its bfloat16 rounding against torch's own
float32 -> bfloat16 conversion (round to nearest even).
"""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(__file__)), "oracle"))
import mlp_oracle as MO  # noqa: E402

SIZES = [(43, 256), (256, 256), (256, 256), (256, 21)]


def _params(rng, sizes=SIZES):
    ps = []
    for fi, fo in sizes:
        ps.append(rng.standard_normal((fo, fi)) * np.sqrt(2.0 / fi))
        ps.append(rng.standard_normal(fo) * 0.1)
    return ps


def test_bf16_matches_torch_rne():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.standard_normal(20000) * 10.0 ** rng.integers(-20, 20, 20000),
                        # ties: low 16 bits exactly 0x8000, both parities of bit 16
                        (np.arange(1, 2001, dtype=np.uint32) << 16 | 0x8000).view(np.float32)])
    x = x.astype(np.float32)
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float64).numpy()
    got = MO.bf16(x)
    assert np.array_equal(got, ref)


def test_gelu_vs_torch():
    torch = pytest.importorskip("torch")
    z = np.linspace(-8, 8, 2001)
    t = torch.from_numpy(z).requires_grad_(True)
    y = torch.nn.functional.gelu(t)  # erf form
    y.sum().backward()
    assert np.abs(MO.gelu(z) - y.detach().numpy()).max() < 1e-14
    assert np.abs(MO.gelu_d(z) - t.grad.numpy()).max() < 1e-14


def test_backward_vs_torch_autograd():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(3)
    ps = _params(rng)
    x = rng.standard_normal((64, 43))
    gy = rng.standard_normal((64, 21))
    y, saved = MO.forward(x, ps, rnd=MO.exact)
    g = MO.backward(gy, ps, saved, rnd=MO.exact)
    tp = [torch.from_numpy(p).requires_grad_(True) for p in ps]
    h = torch.from_numpy(x)
    for li in range(4):
        h = h @ tp[2 * li].T + tp[2 * li + 1]
        if li < 3:
            h = torch.nn.functional.gelu(h)
    assert np.abs(h.detach().numpy() - y).max() < 1e-12
    (h * torch.from_numpy(gy)).sum().backward()
    for i in range(8):
        ref = tp[i].grad.numpy()
        assert np.linalg.norm(g[i] - ref) <= 1e-12 * np.linalg.norm(ref), i


def test_backward_vs_finite_differences():
    rng = np.random.default_rng(5)
    sizes = [(5, 7), (7, 6), (6, 6), (6, 3)]
    ps = _params(rng, sizes)
    x = rng.standard_normal((4, 5))
    gy = rng.standard_normal((4, 3))
    _, saved = MO.forward(x, ps, rnd=MO.exact)
    g = MO.backward(gy, ps, saved, rnd=MO.exact)

    def loss(pp):
        return float((MO.forward(x, pp, rnd=MO.exact)[0] * gy).sum())

    h = 1e-6
    for i, p in enumerate(ps):
        for idx in [tuple(rng.integers(0, s) for s in p.shape) for _ in range(4)]:
            pp = [q.copy() for q in ps]
            pp[i][idx] += h
            lp = loss(pp)
            pp[i][idx] -= 2 * h
            lm = loss(pp)
            fd = (lp - lm) / (2 * h)
            assert abs(fd - g[i][idx]) < 1e-7 * max(1.0, abs(fd)), (i, idx, fd, g[i][idx])
