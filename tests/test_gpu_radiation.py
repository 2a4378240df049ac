"""Long-wave radiation on the GPU: the reference column kernel against the
oracle, the emulator's forward/backward kernels against the FP64 CPU oracle,
and what the trained emulator gains over the plain column."""
import os
import sys
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _oracle_cols(cols):
    return cols


def test_lwr_column_vs_oracle(oracle, gpu):
    import torch

    import oracle as O
    from exoplasim_amd.columns import make_columns
    from exoplasim_amd.radiation import LwrColumn, columns_to_device

    nlev, ncol = 20, 4096
    print("lwr column", flush=True)
    cols = make_columns(ncol, nlev, 11)
    rng = np.random.default_rng(3)
    cols["dcc"] = rng.uniform(0, 0.8, (ncol, nlev))
    cols["dql"] = rng.uniform(0, 2e-4, (ncol, nlev))
    cols["dqo3"] = rng.uniform(0, 1e-5, (ncol, nlev))
    cols["dls"] = (np.arange(ncol) % 3 == 0).astype(float)
    fu, fd, heat, olr = O.lwr(oracle, cols, nlev)
    heat_kday = heat * 86400.0
    dcols = columns_to_device(cols)
    # fp64: 1e-10, the reference's own -O0/-O2 spread where the effective
    # radiation cancels near ztau0 = 1 - 1e-6 (see test_oracle.LWR_TOL).
    # fp32: 5e-3.  In isothermal layers (profiles clipped at 150 K) the sign
    # test that picks the effective-radiation formula (src/radmod.f90:2380)
    # is decided by rounding, and the two formulas differ; an FP32 C
    # restatement of the column differs from the FP64 oracle by 2.5e-3 on
    # these columns, and the reference's own single-precision build
    # (oracle/_ref/libref_rad_t21_r4.so) from its FP64 build by 4.7e-2.
    for fp64, tol in ((True, 1e-10), (False, 5e-3)):
        col = LwrColumn(nlev, fp64=fp64)
        h, o, u, d = col(dcols, fluxes=True)
        torch.cuda.synchronize()
        h = h.double().cpu().numpy().T
        o = o.double().cpu().numpy()
        u = u.double().cpu().numpy().T
        e_u = np.abs(u - fu).max() / np.abs(fu).max()
        e_h = np.linalg.norm(h - heat_kday) / np.linalg.norm(heat_kday)
        e_o = np.abs(o - olr).max() / np.abs(olr).max()
        print(f"fp64={fp64}: flux {e_u:.2e} heating {e_h:.2e} olr {e_o:.2e}", flush=True)
        assert e_u < tol and e_h < (10 * tol if fp64 else tol) and e_o < tol


def test_emulator_gradients_vs_oracle(gpu):
    import torch

    import mlp_oracle as MO
    from exoplasim_amd.radiation import LwEmulator

    print("emulator gradients", flush=True)
    em = LwEmulator(seed=4)
    M = 1000
    g = torch.Generator(device="cpu").manual_seed(2)
    x = torch.randn((M, 43), generator=g).cuda()
    gy = torch.randn((M, 21), generator=g).cuda() / M
    y = em.raw(x)
    y.backward(gy)
    params = [p.detach().double().cpu().numpy() for p in em.params]
    yo, saved = MO.forward(x.double().cpu().numpy(), params)
    assert np.linalg.norm(y.detach().double().cpu().numpy() - yo) / np.linalg.norm(yo) < 1e-3
    go = MO.backward(gy.double().cpu().numpy(), params, saved)
    for i, p in enumerate(em.params):
        gg = p.grad.double().cpu().numpy()
        e = np.linalg.norm(gg - go[i]) / np.linalg.norm(go[i])
        print(f"param {i}: grad rel err {e:.2e}", flush=True)
        assert e < 5e-3, (i, e)


def test_emulator_fused_inference(gpu):
    """the one-kernel inference (exo_lw_emulate) against the per-layer
    kernels and the FP64 oracle network, on a column count that leaves a
    partial block"""
    import torch

    import mlp_oracle as MO
    from exoplasim_amd.columns import make_columns
    from exoplasim_amd.radiation import LwEmulator, columns_to_device, features

    print("emulator fused inference", flush=True)
    em = LwEmulator(seed=5)
    g = torch.Generator(device="cpu").manual_seed(9)
    em.ymean = (torch.randn(21, generator=g) * 3.0).cuda()
    em.ystd = (torch.rand(21, generator=g) + 0.5).cuda()
    with torch.no_grad():
        for b in em.params[1::2]:
            b.copy_(torch.randn(b.shape, generator=g) * 0.1)
    cols = columns_to_device(make_columns(1000, 20, 21), torch.float32)
    h1, o1 = em.predict(cols)
    h2, o2 = em.predict_layers(cols)
    torch.cuda.synchronize()
    rel = lambda a, b: float(torch.linalg.norm(a - b) / torch.linalg.norm(b))  # noqa: E731
    assert torch.isfinite(h1).all() and torch.isfinite(o1).all()
    assert rel(h1, h2) < 2e-3 and rel(o1, o2) < 2e-3, (rel(h1, h2), rel(o1, o2))
    # oracle: same features, FP64 network with the kernels' bf16 rounding
    params = [p.detach().double().cpu().numpy() for p in em.params]
    y, _ = MO.forward(features(cols, 20).double().cpu().numpy(), params)
    y = y * em.ystd.double().cpu().numpy() + em.ymean.double().cpu().numpy()
    ps = cols["dp"].double().cpu().numpy() / 1e5
    h_o = (y[:, :20] / ps[:, None]).T
    e_h = np.linalg.norm(h1.double().cpu().numpy() - h_o) / np.linalg.norm(h_o)
    e_o = np.linalg.norm(o1.double().cpu().numpy() - y[:, 20]) / np.linalg.norm(y[:, 20])
    print(f"fused vs oracle: heating {e_h:.2e}, olr {e_o:.2e}", flush=True)
    assert e_h < 2e-3 and e_o < 2e-3
    # captured in a HIP graph: every replay, with eager work in between,
    # gives the eager result
    out = (torch.empty_like(h1), torch.empty_like(o1))
    g = torch.cuda.CUDAGraph()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        em.predict(cols, out=out)
    torch.cuda.current_stream().wait_stream(s)
    with torch.cuda.graph(g):
        em.predict(cols, out=out)
    for _ in range(3):
        out[0].fill_(float("nan"))
        out[1].fill_(float("nan"))
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(out[0], h1) and torch.equal(out[1], o1)


def test_emulator_training_and_gain(gpu, oracle):
    """train on columns produced by the GPU reference column, then compare
    heating-rate RMS and throughput with the column it replaces: the plain
    column on the same GPU and the reference's column on the host"""
    import torch

    from exoplasim_amd.columns import make_columns
    from exoplasim_amd.radiation import LwEmulator, LwrColumn, columns_to_device, features

    nlev = 20
    ntrain = 2_000_000
    print("emulator training", flush=True)
    col = LwrColumn(nlev, fp64=False)
    xs, ys = [], []
    em = LwEmulator(seed=0)
    for k in range(ntrain // 100_000):
        c = columns_to_device(make_columns(100_000, nlev, 100 + k), torch.float32)
        h, o = col(c)
        xs.append(features(c, nlev))
        ys.append(em.targets(h, o, c))
    x, y = torch.cat(xs), torch.cat(ys)
    t0 = time.time()
    em.fit(x, y, epochs=6, batch=4096, lr=2e-3, log=lambda s: print(s, flush=True))
    print(f"trained in {time.time() - t0:.1f} s", flush=True)
    test = columns_to_device(make_columns(8192, nlev, 7), torch.float32)
    h_ref, o_ref = col(test)
    h_em, o_em = em.predict(test)
    rms = float(torch.sqrt(torch.mean((h_em - h_ref) ** 2)))
    rms_olr = float(torch.sqrt(torch.mean((o_em - o_ref) ** 2)))
    ref_rms = float(torch.sqrt(torch.mean(h_ref ** 2)))
    print(f"heating RMS error {rms:.3f} K/day (signal RMS {ref_rms:.2f}), OLR RMS {rms_olr:.2f}",
          flush=True)
    # timing of 8192-column steps
    def timeit(fn, n=20):
        for _ in range(3):
            fn()
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(n):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t) / n

    t_col = timeit(lambda: col(test))
    t_em = timeit(lambda: em.predict(test))
    t_em_layers = timeit(lambda: em.predict_layers(test))
    print(f"emulator per-layer kernels {8192 / t_em_layers:.3e} col/s", flush=True)
    print(f"plain column {8192 / t_col:.3e} col/s, emulator {8192 / t_em:.3e} col/s, "
          f"gain x{t_col / t_em:.2f}", flush=True)
    # the network learns the column: far better than predicting the mean,
    # and within the 0.5 K/day RMS heating-rate bar
    assert rms < 0.5 * ref_rms
    assert rms <= 0.5, rms
    # what the trained network gains over the column it stands in for: the
    # GPU plain column runs one lane per level with every CU busy at 8192
    # columns, the network is still faster ...
    gain = t_col / t_em
    assert gain >= 2.0, gain
    # ... and >= 20x the reference's column on the host (its C restatement,
    # one thread timed for about a second, scaled by the 16 host threads a
    # model process would use: an upper bound on the host rate)
    import oracle as O

    sub = {k: v[:256] for k, v in make_columns(256, nlev, 9).items()}
    n, t0 = 0, time.perf_counter()
    while time.perf_counter() - t0 < 1.0:
        O.lwr(oracle, sub, nlev)
        n += 256
    cpu_rate = 16 * n / (time.perf_counter() - t0)
    print(f"host column x16 threads {cpu_rate:.3e} col/s, gain x{8192 / t_em / cpu_rate:.1f}",
          flush=True)
    assert 8192 / t_em >= 20.0 * cpu_rate
    # columns outside the training envelope go to the plain column
    out = columns_to_device(make_columns(512, nlev, 8), torch.float32)
    out["dqco2"][:, :64] = 1e7  # 100x the largest CO2 of the training columns
    h_g, o_g, n_plain = em.predict_guarded(out, col)
    h_c, o_c = col(out)
    h_e, o_e = em.predict(out)
    torch.cuda.synchronize()
    assert n_plain >= 64
    assert torch.equal(h_g[:, :64], h_c[:, :64]) and torch.equal(o_g[:64], o_c[:64])
    inside = torch.ones(512, dtype=torch.bool, device="cuda")
    inside[:64] = False
    # the envelope holds almost every column drawn like the training set
    assert n_plain <= 64 + 512 // 20, n_plain
    same = torch.all(h_g == h_e, dim=0) & (o_g == o_e)
    assert bool(torch.all(same | ~inside | (h_g == h_c).all(dim=0)))
