"""Throughput of the T42 L20 grid-point <-> spectral round trip on one MI355X.

One step = one round trip of every member of a batch of independent
T42 L20 states (64x128 Gaussian grid, 946 complex coefficients per field
level) already resident in HBM:
    sp2fc (61 field-levels) + sp2fcdmu (ln ps) + dv2uv (20 levels), each fused
    with fc2gp  ->  102 grid fields per member;
    gp2fc fused with fc2sp on the 61 prognostic grid fields -> 61 spectral.
The headline value is the split-BF16 (three-way split, FP32 accumulate, MFMA)
path; the plain FP64 GPU path is timed beside it, and both are checked
against the FP64 oracle on two members.

The oracle (oracle/, test infrastructure) appears here in three roles only:
as the checker of the outputs the GPU already produced (the error fields,
computed on the host after every timed loop has ended), as the timed CPU
baseline (cpu_baseline, and the plain CPU column of the emulator leg), and
as the source of the dynamics leg's vertical set-up (initsi's arrays, which
the model computes once on the host before any transform).
No timed GPU region calls it, and the transforms fail without the HIP
library instead of falling back to it.

Prints one JSON line (rank 0, N=1).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import threading
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

NTRU, NLEV = 42, 20
HBM_PEAK_GBS = 8000.0

# HBM bytes per launch of the two halves at this workload (128 members,
# split-BF16), from rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE runs of this
# script, each scaled by the counter's reading for a 256 MiB device copy
# (FETCH_SIZE x 2.000, WRITE_SIZE x 1.000); the runs and the command are in
# profiles/t42_l20_hbm_traffic.md.  Keyed by (mode, call, members).
MEASURED_TRAFFIC = {("bf16x3", "inverse_all", 128): 658.9e6,
                    ("bf16x3", "direct_qc", 128): 316.4e6}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--members", type=int, default=128)
    ap.add_argument("--mode", default="bf16x3")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--settle", type=float, default=0.3,
                    help="seconds of untimed round trips before the warm-up steps")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-tune", action="store_true",
                    help="keep the default tiles instead of timing the variants")
    ap.add_argument("--no-extras", action="store_true",
                    help="only the T42 L20 headline (skip ensemble, T127, emulator)")
    ap.add_argument("--extras", default="ensemble,t127,dynamics,emulator",
                    help="comma-separated extra workloads to run beside the headline")
    ap.add_argument("--copy-probe-mb", type=int, default=0,
                    help="also run one device-to-device copy of this many MiB (the "
                         "reference kernel that scales the profiler's byte counters)")
    ap.add_argument("--emulator-columns", type=int, default=2_000_000)
    ap.add_argument("--emulator-epochs", type=int, default=6)
    return ap.parse_args()


def rel_l2(a, b):
    import numpy as np

    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


# field-levels of one round trip's grid output (per member) and of its
# returned spectra, by field
def field_slices(nlev):
    nfl = 3 * nlev + 1
    grid = {"vor": (0, nlev), "div": (nlev, 2 * nlev), "tmp": (2 * nlev, 3 * nlev),
            "lnps": (nfl - 1, nfl), "dlnps_dmu": (nfl, nfl + 1),
            "u_cos": (nfl + 1, nfl + 1 + nlev), "v_cos": (nfl + 1 + nlev, nfl + 1 + 2 * nlev)}
    spec = {k: grid[k] for k in ("vor", "div", "tmp", "lnps")}
    return grid, spec


def per_field_errors(r, refs, nlev):
    """rel L2 over all levels of a field and its largest absolute error,
    worst of the checked members"""
    import numpy as np

    grid, spec = field_slices(nlev)
    out = {}
    for kind, key, sl in (("grid", "gp", grid), ("spectral", "sp_out", spec)):
        ri = 0 if kind == "grid" else 1
        for name, (a, b) in sl.items():
            l2, mx = 0.0, 0.0
            for k in range(len(refs)):
                x, y = r[key][k][a:b], refs[k][ri][a:b]
                l2 = max(l2, rel_l2(x, y))
                mx = max(mx, float(np.abs(x - y).max()))
            out[f"{kind}.{name}"] = {"rel_l2": l2, "max_abs": mx}
    return out


def time_loop(fn, steps, warmup):
    import torch

    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(steps):
        fn()
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / 1e3  # seconds


def settle(fn, seconds):
    """untimed calls of fn for at least `seconds` of wall time, so that the
    GPU clocks reach their sustained level before a timed loop"""
    import torch

    t0 = time.perf_counter()
    while time.perf_counter() - t0 < seconds:
        for _ in range(10):
            fn()
        torch.cuda.synchronize()


def cpu_baseline(seconds: float, sp_member):
    """The oracle's FP64 round trip (a C restatement of the reference loops)
    on host threads, one member per call, for about `seconds`."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    from oracle import Oracle

    ora = Oracle()
    nlat, nlon = 64, 128
    tables = ora.legini(NTRU, nlat)["_all"]
    cores = min(16, os.cpu_count() or 1)
    done = [0] * cores
    stop = time.perf_counter() + seconds

    def work(i):
        while time.perf_counter() < stop:
            ora.roundtrip(NTRU, NLEV, sp_member, tables=tables)
            done[i] += 1

    ts = [threading.Thread(target=work, args=(i,)) for i in range(cores)]
    t0 = time.perf_counter()
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    el = time.perf_counter() - t0
    return {"value": sum(done) / el, "unit": "round trips/s", "cores": cores, "kind": "port",
            "sample": f"{sum(done)} T42 L20 FP64 round trips of one seeded member on "
                      f"{cores} host threads in {el:.1f} s (oracle/exo_oracle.c, "
                      "gcc -O2)"}


# ---------------------------------------------------------------------------
# extra workloads reported beside the headline
# ---------------------------------------------------------------------------
ENS_GROUPS = [(21, 10, 48), (42, 20, 32), (63, 20, 16)]


def bench_ensemble(args, ora):
    """96 members of mixed truncation: one graph launch (the schedule the
    autotuner keeps; both schedules are timed as graphs: the two persistent
    kernels, one batched round trip per truncation) vs the batched round
    trips launched directly vs one round trip per member."""
    import numpy as np
    import torch

    from exoplasim_amd import sht
    from exoplasim_amd.synth import ensemble

    ens = sht.Ensemble(ENS_GROUPS, args.mode)
    host = []
    for k, (ntru, nlev, nmem) in enumerate(ENS_GROUPS):
        sp = ensemble(ntru, nlev, nmem, cfg=3)
        ens.bufs[k][0].copy_(torch.from_numpy(sp).float())
        host.append(sp)
    tiles = ens.autotune()
    sched = ens.schedule()
    steps = max(5, args.steps // 2)
    t_graph = time_loop(ens.run, steps, 3)
    other = "batched" if sched == "persistent" else "persistent"
    ens.set_schedule(other)
    t_other = time_loop(ens.run, steps, 3)
    ens.set_schedule(sched)
    t_pers, t_bgraph = (t_graph, t_other) if sched == "persistent" else (t_other, t_graph)
    t_groups = time_loop(ens.run_groups, steps, 3)
    t_members = time_loop(ens.run_members, max(3, steps // 5), 1)
    nm = ens.members
    ens.run()
    torch.cuda.synchronize()
    err = 0.0
    for k, (ntru, nlev, nmem) in enumerate(ENS_GROUPS):
        gk = ens.bufs[k][1].double().cpu().numpy()
        sk = ens.bufs[k][2].double().cpu().numpy()
        for j in range(nmem):  # every member of every group
            g, s_ = ora.roundtrip(ntru, nlev, host[k][j])
            err = max(err, rel_l2(gk[j], g), rel_l2(sk[j], s_))
    ens.close()
    r = lambda t, n: round(nm * n / t, 1)  # noqa: E731
    return {"groups": [{"ntru": a, "nlev": b, "members": c} for a, b, c in ENS_GROUPS],
            "mode": args.mode, "tiles": tiles, "unit": "member round trips/s",
            "graph": r(t_graph, steps), "graph_schedule": sched,
            "graph_persistent": r(t_pers, steps),
            "graph_batched": r(t_bgraph, steps),
            "batched_per_truncation": r(t_groups, steps),
            "loop_of_single_members": r(t_members, max(3, steps // 5)),
            "speedup_vs_member_loop": round(t_members / max(3, steps // 5) / (t_graph / steps), 2),
            "ms_per_ensemble_graph": round(t_graph / steps * 1e3, 4),
            "max_rel_l2_vs_fp64_oracle": err}


def _mfma_flops_inverse(ntru, nfield, nmem, fb, nterm):
    """MFMA flops executed by a bf16x3 inverse launch: per block (8 pairs x
    fb fields) and m, ceil((T+1-m)/32) 16x16x32 tiles, 6 products each"""
    nm = ntru + 1
    tiles = sum((nm - m + 31) // 32 for m in range(nm))
    groups = (nfield + fb - 1) // fb
    nblk = nmem * groups * ({21: 32, 42: 64, 63: 96, 127: 192}[ntru] // 2 // 8)
    return nblk * tiles * 6 * nterm * 2 * 16 * 16 * 32


def bench_t127(args, ora):
    """T127 L40 split-BF16 round trips: HBM GB/s, MFMA rate, error."""
    import numpy as np
    import torch

    from exoplasim_amd import sht
    from exoplasim_amd.synth import ensemble

    ntru, nlev, nmem = 127, 40, 8
    plan = sht.Plan(ntru)
    mode = args.mode
    tiles = plan.autotune(mode, nmem, nlev)
    sp = ensemble(ntru, nlev, nmem, cfg=4, slope=2.0, wave1=10.0)
    sp_in, gp, sp_out = plan.alloc_roundtrip(nmem, nlev, mode)
    sp_in.copy_(torch.from_numpy(sp).to(sp_in.dtype))
    nfl = 3 * nlev + 1
    steps = max(5, args.steps // 2)
    t = time_loop(lambda: plan.roundtrip(mode, sp_in, gp, sp_out, nlev), steps, 2) / steps
    es = sp_in.element_size()
    nbytes = nmem * (nfl + 1 + 2 * nlev + nfl) * (plan.nrsp + plan.nlat * plan.nlon) * es
    t_qi = time_loop(lambda: plan.inverse(mode, sp_in, gp, nfl, 0, 0, "qi"), steps, 2) / steps
    # fields per block of the qi launch: 4 at T127, halved by tile variants 1
    # and 3 (variant 2 keeps the fields and doubles the threads)
    fb = 2 if tiles[0] in (1, 3) else 4
    fl_qi = _mfma_flops_inverse(ntru, nfl, nmem, fb, 1)
    torch.cuda.synchronize()
    plan.roundtrip(mode, sp_in, gp, sp_out, nlev)
    g, s_ = ora.roundtrip(ntru, nlev, sp[0])
    torch.cuda.synchronize()
    eg = rel_l2(gp[0].double().cpu().numpy(), g)
    es_ = rel_l2(sp_out[0].double().cpu().numpy(), s_)
    useful_qi = 2 * nmem * nfl * 2 * (plan.nrsp // 2) * (plan.nlat // 2)
    plan.close()
    return {"members": nmem, "mode": mode, "tiles": tiles,
            "rt_per_s": round(nmem / t, 1), "ms_per_step": round(t * 1e3, 4),
            "roundtrip_GBs": round(nbytes / t / 1e9, 1),
            "inverse_qi_ms": round(t_qi * 1e3, 4),
            "inverse_qi_mfma_TFLOPs": round(fl_qi / t_qi / 1e12, 1),
            "inverse_qi_mfma_frac_of_bf16_peak": round(fl_qi / t_qi / 2.5e15, 4),
            # the folded algorithm's own work: one multiply-add per (coefficient,
            # component, latitude pair) of every field; the executed MFMA flops
            # above also count the three-way split, the parity zeros and
            # empty columns
            "inverse_qi_useful_TFLOPs": round(useful_qi / t_qi / 1e12, 2),
            "grid_rel_l2_vs_fp64_oracle": eg, "spectral_rel_l2_vs_fp64_oracle": es_}


def _mfma_flops_dirx(ntru, nlev, nmem, nslot):
    """FP32 MFMA flops executed by one k_dirx_mf launch: per block (member,
    8 levels, wavenumber range) and latitude-pair step of 4, one
    16x16x4 product per (m, parity, non-empty row tile, table)"""
    npair = {21: 32, 42: 64, 63: 96, 127: 192}[ntru] // 2
    tiles = sum(((ntru - m - p) // 2 + 16) // 16 for m in range(ntru + 1) for p in (0, 1)
                if ntru - m - p >= 0)
    return nmem * ((nlev + 7) // 8) * tiles * nslot * (npair // 4) * 2 * 16 * 16 * 4


def bench_dynamics(args, ora):
    """The device-resident transform group of one dynamics step
    (exo_gridpointa: sp2fc, sp2fcdmu, dv2uv, d/dlambda, fc2gp, calcgp,
    gp2fc, fc2sp, mktend, qtend) at T42 L20 on 32 members, and its mktend
    alone: the MFMA direct sums over three tables."""
    import numpy as np
    import torch

    from exoplasim_amd import sht
    from exoplasim_amd.dynamics import Dynamics
    from exoplasim_amd.synth import spectral_fields

    ntru, nlev, nmem, mode = 42, 20, 32, args.mode
    ct = (6371220.0 * 7.292e-5) ** 2 / 287.0  # temperatures in units of ct
    scal = (287.0 / 461.5, 1870.0 / (287.0 / 0.286) - 1.0, 101100.0, 7.292e-5)
    v = ora.vertical(nlev, t0=np.linspace(220.0, 288.0, nlev) / ct)
    plan = sht.Plan(ntru)
    dyn = Dynamics.from_vertical(v, *scal)
    nhor = plan.nlat * plan.nlon
    nfl = 4 * nlev + 1
    st0 = 0.01 * np.stack([spectral_fields(ntru, nfl, 5000 + k, 1.5, 0.0) for k in range(nmem)])
    st0[:, :nlev, 2] += 1.63299310207  # planetary vorticity at Re(m=0, n=1)
    st0[:, 3 * nlev:4 * nlev] *= 0.1    # humidity
    st, work, tend = dyn.alloc_gridpointa(plan, nmem, mode)
    st.copy_(torch.from_numpy(st0).to(st.dtype))
    steps = max(5, args.steps // 2)
    t = time_loop(lambda: dyn.gridpointa(plan, mode, st, work, tend, nqspec=1), steps, 2) / steps
    # mktend alone on Fourier inputs of the same shape
    dt = plan.dtype_of(mode)
    ins = [torch.randn((nmem, nlev, plan.nlat, plan.nlon), device="cuda", dtype=dt)
           for _ in range(6)]
    outs = [torch.empty((nmem, nlev, plan.nrsp), device="cuda", dtype=dt) for _ in range(3)]
    t_mk = time_loop(lambda: plan.mktend(mode, *outs, *ins, nlev), steps, 2) / steps
    es = ins[0].element_size()
    mk_bytes = nmem * nlev * (6 * nhor + 3 * plan.nrsp) * es
    # useful flops: 16 real table terms (D 6, T 6, Z 4) per coefficient and pair
    ncsp = plan.nrsp // 2
    mk_useful = nmem * nlev * 16 * ncsp * (plan.nlat // 2) * 2
    mk_exec = sum(_mfma_flops_dirx(ntru, nlev, nmem, ns) for ns in (3, 3, 2))
    torch.cuda.synchronize()
    got = tend[0].double().cpu().numpy()
    want, _ = ora.gridpointa(ntru, nlev, st0[0].astype(np.float32).astype(np.float64)
                             if mode != "f64" else st0[0], v, scal, 1)
    errs = {}
    for i, name in enumerate(("szt", "sdt", "stt", "sqt")):
        part = slice(i * nlev, (i + 1) * nlev)
        errs[name] = rel_l2(got[part], want[part])
    errs["spt"] = rel_l2(got[4 * nlev], want[4 * nlev])
    dyn.close()
    plan.close()
    return {"members": nmem, "mode": mode,
            "gridpointa_ms_per_step": round(t * 1e3, 4),
            "member_steps_per_s": round(nmem / t, 1),
            "tendency_rel_l2_vs_fp64_oracle_member0": errs,
            "mktend_ms": round(t_mk * 1e3, 4),
            "mktend_GBs": round(mk_bytes / t_mk / 1e9, 1),
            "mktend_useful_TFLOPs": round(mk_useful / t_mk / 1e12, 2),
            "mktend_mfma_executed_TFLOPs": round(mk_exec / t_mk / 1e12, 2),
            "mktend_mfma_frac_of_fp32_peak": round(mk_exec / t_mk / 157.3e12, 4)}


def bench_emulator(args, ora):
    """Train the long-wave emulator on columns from the GPU reference column,
    then time 8192-column steps against the plain column (GPU and CPU)."""
    import numpy as np
    import torch

    from exoplasim_amd.columns import make_columns
    from exoplasim_amd.radiation import LwEmulator, LwrColumn, columns_to_device, features

    nlev, ncol = 20, 8192
    col = LwrColumn(nlev, fp64=False)
    em = LwEmulator(seed=0)
    xs, ys = [], []
    chunk = 200_000
    t0 = time.perf_counter()
    for k in range((args.emulator_columns + chunk - 1) // chunk):
        c = columns_to_device(make_columns(chunk, nlev, 1000 + k), torch.float32)
        h, o = col(c)
        xs.append(features(c, nlev))
        ys.append(em.targets(h, o, c))
    x, y = torch.cat(xs)[: args.emulator_columns], torch.cat(ys)[: args.emulator_columns]
    del xs, ys
    torch.cuda.synchronize()
    t_data = time.perf_counter() - t0
    t0 = time.perf_counter()
    em.fit(x, y, epochs=args.emulator_epochs, batch=4096, lr=2e-3)
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t0
    del x, y
    test_np = make_columns(ncol, nlev, 7)
    test = columns_to_device(test_np, torch.float32)
    h_ref, o_ref = col(test)
    h_em, o_em = em.predict(test)
    rms = float(torch.sqrt(torch.mean((h_em - h_ref) ** 2)))
    sig = float(torch.sqrt(torch.mean(h_ref ** 2)))
    rms_olr = float(torch.sqrt(torch.mean((o_em - o_ref) ** 2)))
    # steps captured in HIP graphs (10 steps per graph), so that the launch
    # cost of the Python caller does not count; every replay is checked
    # against the eager result
    outb = (torch.empty_like(h_em), torch.empty_like(o_em))
    em.predict(test, out=outb)
    ref_em = (outb[0].clone(), outb[1].clone())
    outc = col(test)
    ref_col = (outc[0].clone(), outc[1].clone())

    def graphed(fn, reps=10):
        g = torch.cuda.CUDAGraph()
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):
            fn()  # warm-up on the capture stream
        torch.cuda.current_stream().wait_stream(s)
        with torch.cuda.graph(g):
            for _ in range(reps):
                res = fn()
        return g, res

    g_em, _ = graphed(lambda: em.predict(test, out=outb))
    g_col, out_col = graphed(lambda: col(test))
    for _ in range(3):
        g_em.replay()
        g_col.replay()
        torch.cuda.synchronize()
        assert torch.equal(outb[0], ref_em[0]) and torch.equal(outb[1], ref_em[1])
        assert torch.equal(out_col[0], ref_col[0]) and torch.equal(out_col[1], ref_col[1])
    t_col = time_loop(g_col.replay, 5, 1) / 50
    t_em = time_loop(g_em.replay, 20, 2) / 200
    t_em_eager = time_loop(lambda: em.predict(test, out=outb), 200, 10) / 200
    # CPU: the oracle column (C restatement of lwr) on host threads
    import oracle as O

    cores = min(16, os.cpu_count() or 1)
    sub = {k: v[: 256] for k, v in test_np.items()}
    done = [0] * cores
    stop = time.perf_counter() + min(args.cpu_seconds, 5.0)

    def work(i):
        while time.perf_counter() < stop:
            O.lwr(ora, sub, nlev)
            done[i] += 256

    ts = [threading.Thread(target=work, args=(i,)) for i in range(cores)]
    tc = time.perf_counter()
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    cpu_cols = sum(done) / (time.perf_counter() - tc)
    return {"train_columns": int(args.emulator_columns), "epochs": args.emulator_epochs,
            "data_s": round(t_data, 2), "train_s": round(t_train, 2),
            "heating_rms_K_per_day": round(rms, 4), "heating_signal_rms_K_per_day": round(sig, 3),
            "olr_rms_W_m2": round(rms_olr, 3),
            "columns_per_step": ncol,
            "emulator_columns_per_s": round(ncol / t_em, 1),
            "emulator_columns_per_s_eager_python": round(ncol / t_em_eager, 1),
            "timing": "10 steps per captured HIP graph, replays checked against eager results",
            "gpu_plain_column_columns_per_s": round(ncol / t_col, 1),
            "cpu_plain_column_columns_per_s": round(cpu_cols, 1), "cpu_cores": cores,
            "gain_vs_gpu_plain_column": round(t_col / t_em, 2),
            "gain_vs_cpu_plain_column": round(ncol / t_em / cpu_cols, 1)}


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); "
              "the transform of independent members shards trivially and is not "
              "measured across GPUs here.", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch

    from exoplasim_amd import sht
    from exoplasim_amd.synth import ensemble

    if args.copy_probe_mb:
        src = torch.ones(args.copy_probe_mb << 18, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    nmem = args.members
    plan = sht.Plan(NTRU)
    nfl = 3 * NLEV + 1
    ngp = nfl + 1 + 2 * NLEV
    nrsp, nhor = plan.nrsp, plan.nlat * plan.nlon
    base = ensemble(NTRU, NLEV, 4, cfg=2)  # 4 distinct members, tiled
    sp_host = np.concatenate([base] * (nmem // 4) + [base[: nmem % 4]])

    results = {}
    tiles = {}
    # plan-time tile choices of every mode first (untimed, like building the
    # tables), the headline mode last, so that its timed run follows its own
    # kernels' tuning rather than a change of workload
    modes = [args.mode, "f64"] + (["f32mfma"] if args.mode != "f32mfma" else [])
    for mode in modes[::-1]:
        tiles[mode] = (plan.tiles(mode) if args.no_tune
                       else plan.autotune(mode, nmem, NLEV))
    for mode in modes:
        sp_in, gp, sp_out = plan.alloc_roundtrip(nmem, NLEV, mode)
        sp_in.copy_(torch.from_numpy(sp_host).to(sp_in.dtype))
        es = sp_in.element_size()

        def rt():
            plan.roundtrip(mode, sp_in, gp, sp_out, NLEV)

        # the same K steps from cold, reported beside the value: the clocks
        # are still rising during the first tens of milliseconds of work
        t_cold = time_loop(rt, args.steps, args.warmup)
        settle(rt, args.settle)
        t = time_loop(rt, args.steps, args.warmup)
        # per-call timing of the two halves of a round trip
        legs = {
            # pre-split (split-BF16) + sp2fc, sp2fcdmu, dv2uv fused with fc2gp,
            # one launch: each transformed field-level read and written once
            "inverse_all": (lambda: plan.inverse_all(mode, sp_in, gp, NLEV),
                            nmem * (nfl + 1 + 2 * NLEV) * (nrsp + nhor) * es),
            "direct_qc": (lambda: plan.direct(mode, gp, sp_out, nfl, 0, 0),
                          nmem * nfl * (nhor + nrsp) * es),
        }
        kt = {}
        for name, (fn, nbytes) in legs.items():
            # median of three timed loops of K calls each
            tk = sorted(time_loop(fn, args.steps, 2) for _ in range(3))[1] / args.steps
            kt[name] = {"ms": tk * 1e3, "bytes": nbytes, "GBs": nbytes / tk / 1e9}
        plan.roundtrip(mode, sp_in, gp, sp_out, NLEV)
        torch.cuda.synchronize()
        results[mode] = {
            "time": t,
            "cold_ms_per_step": t_cold / args.steps * 1e3,
            "rt_per_s": nmem * args.steps / t,
            "ms_per_step": t / args.steps * 1e3,
            "kernels": kt,
            "gp": gp[:2].double().cpu().numpy(),
            "sp_out": sp_out[:2].double().cpu().numpy(),
            "bytes": sum(v["bytes"] for v in kt.values()),
        }
        del sp_in, gp, sp_out
        torch.cuda.empty_cache()

    # error of both GPU paths against the FP64 oracle on two members
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from oracle import Oracle

    ora = Oracle()
    errs = {}
    refs = [ora.roundtrip(NTRU, NLEV, sp_host[k]) for k in range(2)]
    for mode, r in results.items():
        eg = max(rel_l2(r["gp"][k], refs[k][0]) for k in range(2))
        es_ = max(rel_l2(r["sp_out"][k], refs[k][1]) for k in range(2))
        mx = max(float(np.abs(r["gp"][k] - refs[k][0]).max() / np.abs(refs[k][0]).max())
                 for k in range(2))
        errs[mode] = {"grid_rel_l2": eg, "spectral_rel_l2": es_, "grid_max_abs_rel": mx,
                      "per_field": per_field_errors(r, refs, NLEV)}

    head = results[args.mode]
    dom_name, dom = max(head["kernels"].items(), key=lambda kv: kv[1]["ms"])
    achieved = dom["GBs"]
    traffic = MEASURED_TRAFFIC.get((args.mode, dom_name, nmem))
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": traffic,
                "traffic_source": (f"rocprofv3 FETCH_SIZE/WRITE_SIZE of {dom_name}, "
                                   f"{nmem} members {args.mode}" if traffic else
                                   "not measured for this mode and member count"),
                "kernel": dom_name, "kernel_ms": round(dom["ms"], 4),
                "algorithmic_bytes_per_launch": dom["bytes"]}
    # MFMA work of the split-BF16 inverse call: executed (the model held to
    # the hardware counters: qi and qj per field group, dv2uv per level
    # group with two terms) against the folded algorithm's own multiply-adds
    inv_mfma = None
    if args.mode == "bf16x3":
        t_inv = head["kernels"]["inverse_all"]["ms"] / 1e3
        fb = 4 if tiles[args.mode][0] in (1, 3) else 8  # tile variants 1 and 3 halve the fields
        ex = (_mfma_flops_inverse(NTRU, nfl, nmem, fb, 1) + _mfma_flops_inverse(NTRU, 1, nmem, fb, 1)
              + _mfma_flops_inverse(NTRU, NLEV, nmem, fb // 2, 2))
        per = 2 * (nrsp // 2) * (plan.nlat // 2) * 2  # flops per field and component set
        useful = nmem * per * ((nfl + 1) + 4 * NLEV)
        inv_mfma = {"executed_TFLOPs": round(ex / t_inv / 1e12, 1),
                    "executed_frac_of_bf16_peak": round(ex / t_inv / 2.5e15, 4),
                    "useful_TFLOPs": round(useful / t_inv / 1e12, 2),
                    "note": "pack + fused inverse call, the autotuned tile variant"}
    # the metric is quoted at <= 1e-6 rel L2 against the FP64 oracle: a run
    # whose headline mode misses that is reported and fails
    tol = 1e-6 if args.mode != "f64" else 1e-12
    accurate = errs[args.mode]["grid_rel_l2"] <= tol and errs[args.mode]["spectral_rel_l2"] <= tol
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(args.cpu_seconds, sp_host[0])
    extras = {}
    if not args.no_extras:
        legs = (("ensemble", bench_ensemble), ("t127", bench_t127),
                ("dynamics", bench_dynamics), ("emulator", bench_emulator))
        for name, fn in (leg for leg in legs if leg[0] in args.extras.split(",")):
            t0 = time.perf_counter()
            try:
                extras[name] = fn(args, ora)
            except Exception as exc:  # reported, never hidden
                extras[name] = {"error": f"{type(exc).__name__}: {exc}"}
            print(f"[bench] {name} done in {time.perf_counter() - t0:.1f} s", file=sys.stderr,
                  flush=True)
    out = {
        "metric": "T42-L20 gp<->sp round trips/s, split-BF16 vs FP64 GPU at <=1e-6 rel L2",
        "value": round(head["rt_per_s"], 1),
        "unit": "round trips/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(head["ms_per_step"], 4),
        "cold_ms_per_step": round(head["cold_ms_per_step"], 4),
        "settle_s": args.settle,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16x3-mfma-fp32acc" if args.mode == "bf16x3" else args.mode,
        "data": "synthetic",
        "config": {"workload": "T42 L20 round trip (64x128 grid, 946 coeffs, 61 prognostic "
                               "field-levels + dmu + u/v)",
                   "members_per_step": nmem, "mode": args.mode,
                   "tiles": {m: dict(zip(("inverse", "inverse_uv", "direct"), v))
                             for m, v in tiles.items()}},
        "roofline": roofline,
        "inverse_all_mfma": inv_mfma,
        "cpu_baseline": cpu,
        # the FP64 path's own roofline: its longer half's algorithmic bytes
        # (FP64 interface arrays) over the event-timed call, against 8 TB/s
        "fp64_gpu": {"value": round(results["f64"]["rt_per_s"], 1),
                     "ms_per_step": round(results["f64"]["ms_per_step"], 4),
                     "roofline_frac": round(max(results["f64"]["kernels"].values(),
                                                key=lambda v: v["ms"])["GBs"] / HBM_PEAK_GBS,
                                            4)},
        "speedup_vs_fp64_gpu": round(head["rt_per_s"] / results["f64"]["rt_per_s"], 3),
        "f32mfma_gpu": ({"value": round(results["f32mfma"]["rt_per_s"], 1),
                         "ms_per_step": round(results["f32mfma"]["ms_per_step"], 4)}
                        if "f32mfma" in results else None),
        "error_vs_fp64_oracle": errs,
        "accuracy_ok": accurate,
        "kernels": {m: {k: {"ms": round(v["ms"], 4), "GBs": round(v["GBs"], 1)}
                        for k, v in r["kernels"].items()} for m, r in results.items()},
        "roundtrip_GBs": round(head["bytes"] * args.steps / head["time"] / 1e9, 1),
        **extras,
    }
    print(json.dumps(out), flush=True)
    plan.close()
    if not accurate:
        print(f"bench.py: {args.mode} misses the {tol:g} rel L2 bound against the FP64 "
              f"oracle: {errs[args.mode]}", file=sys.stderr)
        sys.exit(3)


if __name__ == "__main__":
    main()
