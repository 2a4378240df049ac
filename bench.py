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


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--members", type=int, default=128)
    ap.add_argument("--mode", default="bf16x3")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-tune", action="store_true",
                    help="keep the default tiles instead of timing the variants")
    return ap.parse_args()


def rel_l2(a, b):
    import numpy as np

    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


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

    nmem = args.members
    plan = sht.Plan(NTRU)
    nfl = 3 * NLEV + 1
    ngp = nfl + 1 + 2 * NLEV
    nrsp, nhor = plan.nrsp, plan.nlat * plan.nlon
    base = ensemble(NTRU, NLEV, 4, cfg=2)  # 4 distinct members, tiled
    sp_host = np.concatenate([base] * (nmem // 4) + [base[: nmem % 4]])

    results = {}
    tiles = {}
    for mode in (args.mode, "f64"):
        # plan-time tile choice (untimed, like building the tables)
        tiles[mode] = (plan.tiles(mode) if args.no_tune
                       else plan.autotune(mode, nmem, NLEV))
        sp_in, gp, sp_out = plan.alloc_roundtrip(nmem, NLEV, mode)
        sp_in.copy_(torch.from_numpy(sp_host).to(sp_in.dtype))
        es = sp_in.element_size()

        def rt():
            plan.roundtrip(mode, sp_in, gp, sp_out, NLEV)

        t = time_loop(rt, args.steps, args.warmup)
        # per-launch timing of the four kernels of a round trip
        legs = {
            "inverse_qi": (lambda: plan.inverse(mode, sp_in, gp, nfl, 0, 0, "qi"),
                           nmem * nfl * (nrsp + nhor) * es),
            "inverse_qj": (lambda: plan.inverse(mode, sp_in, gp, 1, nfl - 1, nfl, "qj"),
                           nmem * (nrsp + nhor) * es),
            "inverse_uv": (lambda: plan.inverse_uv(mode, sp_in, gp, NLEV, 0, NLEV, nfl + 1,
                                                   nfl + 1 + NLEV),
                           nmem * NLEV * 2 * (nrsp + nhor) * es),
            "direct_qc": (lambda: plan.direct(mode, gp, sp_out, nfl, 0, 0),
                          nmem * nfl * (nhor + nrsp) * es),
        }
        kt = {}
        for name, (fn, nbytes) in legs.items():
            tk = time_loop(fn, args.steps, 2) / args.steps
            kt[name] = {"ms": tk * 1e3, "bytes": nbytes, "GBs": nbytes / tk / 1e9}
        plan.roundtrip(mode, sp_in, gp, sp_out, NLEV)
        torch.cuda.synchronize()
        results[mode] = {
            "time": t,
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
        errs[mode] = {"grid_rel_l2": eg, "spectral_rel_l2": es_, "grid_max_abs_rel": mx}

    head = results[args.mode]
    dom_name, dom = max(head["kernels"].items(), key=lambda kv: kv[1]["ms"])
    achieved = dom["GBs"]
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                "kernel": dom_name, "kernel_ms": round(dom["ms"], 4),
                "algorithmic_bytes_per_launch": dom["bytes"]}
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(args.cpu_seconds, sp_host[0])
    out = {
        "metric": "T42-L20 gp<->sp round trips/s, split-BF16 vs FP64 GPU at <=1e-6 rel L2",
        "value": round(head["rt_per_s"], 1),
        "unit": "round trips/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(head["ms_per_step"], 4),
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
        "cpu_baseline": cpu,
        "fp64_gpu": {"value": round(results["f64"]["rt_per_s"], 1),
                     "ms_per_step": round(results["f64"]["ms_per_step"], 4)},
        "speedup_vs_fp64_gpu": round(head["rt_per_s"] / results["f64"]["rt_per_s"], 3),
        "error_vs_fp64_oracle": errs,
        "kernels": {m: {k: {"ms": round(v["ms"], 4), "GBs": round(v["GBs"], 1)}
                        for k, v in r["kernels"].items()} for m, r in results.items()},
        "roundtrip_GBs": round(head["bytes"] * args.steps / head["time"] / 1e9, 1),
    }
    print(json.dumps(out), flush=True)
    plan.close()


if __name__ == "__main__":
    main()
