"""Mixed-truncation ensembles (T21 / T42 / T63 in one hipGraph of two
persistent kernels, or of one batched round trip per group) against the FP64
oracle, the per-group batched round trip and the one-member-at-a-time loop.

Each work item of the persistent kernels is the same source as one block of
the per-truncation kernels; compiled into different kernels the FMA
contraction may differ, so the schedules agree to 1e-6 (and each with the
oracle), while replays of one kernel are bitwise identical: the graph is
replayed several times with eager work in between, and each replay is
checked.
"""
import numpy as np
import pytest

from exoplasim_amd import sht
from exoplasim_amd.synth import ensemble

pytestmark = pytest.mark.gpu

GROUPS = [(21, 10, 3), (42, 20, 2), (63, 20, 2)]


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64).ravel()
    b = np.asarray(b, dtype=np.float64).ravel()
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _fill(ens):
    import torch

    host = []
    for k, (ntru, nlev, nmem) in enumerate(ens.groups):
        sp = ensemble(ntru, nlev, nmem, cfg=3 + k)
        ens.bufs[k][0].copy_(torch.from_numpy(sp).float())
        host.append(sp)
    return host


def _poison(ens):
    for _, gp, so in ens.bufs:
        gp.fill_(float("nan"))
        so.fill_(float("nan"))


def _outputs(ens):
    import torch

    torch.cuda.synchronize()
    return [(gp.cpu().numpy().copy(), so.cpu().numpy().copy()) for _, gp, so in ens.bufs]


def _same(a, b):
    return all(np.array_equal(x[0], y[0]) and np.array_equal(x[1], y[1]) for x, y in zip(a, b))


@pytest.mark.parametrize("mode", ["bf16x3", "f32mfma", "f32"])
def test_ensemble_vs_oracle_and_loops(oracle, gpu, mode):
    print(f"ensemble {GROUPS} {mode}", flush=True)
    ens = sht.Ensemble(GROUPS, mode)
    try:
        host = _fill(ens)
        _poison(ens)
        ens.run()
        ref = _outputs(ens)
        for k, (ntru, nlev, nmem) in enumerate(ens.groups):
            for m in range(nmem):
                g_o, s_o = oracle.roundtrip(ntru, nlev, host[k][m])
                eg = rel_l2(ref[k][0][m], g_o)
                es = rel_l2(ref[k][1][m], s_o)
                assert eg < 1e-6 and es < 1e-6, (ntru, m, eg, es)
        # graph replays with eager work in between: bitwise
        for run in (ens.run, ens.run, ens.run_eager, ens.run):
            _poison(ens)
            run()
            assert _same(_outputs(ens), ref), run.__name__
        # the per-truncation kernels: batched and one member at a time agree
        # bitwise with each other and to 1e-6 with the persistent kernels
        _poison(ens)
        ens.run_groups()
        grp = _outputs(ens)
        _poison(ens)
        ens.run_members()
        assert _same(_outputs(ens), grp)
        for k in range(len(GROUPS)):
            assert rel_l2(grp[k][0], ref[k][0]) < 1e-6 and rel_l2(grp[k][1], ref[k][1]) < 1e-6
    finally:
        ens.close()


@pytest.mark.parametrize("mode", ["bf16x3", "f32mfma"])
def test_ensemble_batched_schedule(oracle, gpu, mode):
    # the graph holding one batched round trip per group: every replay (with
    # eager work between) bitwise equal to the eager batched launches, and
    # back to the persistent kernels bitwise equal to their first result
    print(f"ensemble batched schedule {GROUPS} {mode}", flush=True)
    ens = sht.Ensemble(GROUPS, mode)
    try:
        host = _fill(ens)
        assert ens.schedule() == "persistent"
        ens.run()
        pers = _outputs(ens)
        _poison(ens)
        ens.run_groups()
        grp = _outputs(ens)
        ens.set_schedule("batched")
        assert ens.schedule() == "batched"
        for run in (ens.run, ens.run_groups, ens.run, ens.run_eager, ens.run):
            _poison(ens)
            run()
            assert _same(_outputs(ens), grp), run.__name__
        for k, (ntru, nlev, nmem) in enumerate(ens.groups):
            for m in range(nmem):
                g_o, s_o = oracle.roundtrip(ntru, nlev, host[k][m])
                assert rel_l2(grp[k][0][m], g_o) < 1e-6 and rel_l2(grp[k][1][m], s_o) < 1e-6
        ens.set_schedule("persistent")
        _poison(ens)
        ens.run()
        assert _same(_outputs(ens), pers)
        with pytest.raises(ValueError):
            ens.set_schedule("fused")
    finally:
        ens.close()


def test_ensemble_tiles_and_autotune(oracle, gpu):
    print("ensemble tiles/autotune", flush=True)
    ens = sht.Ensemble(GROUPS, "bf16x3")
    try:
        host = _fill(ens)
        ens.run()
        base = _outputs(ens)
        ens.set_tiles([1] * (2 * len(GROUPS)))
        assert ens.tiles() == (1,) * (2 * len(GROUPS))
        _poison(ens)
        ens.run()
        half = _outputs(ens)
        for k in range(len(GROUPS)):
            assert rel_l2(half[k][0], base[k][0]) < 1e-6
            assert rel_l2(half[k][1], base[k][1]) < 1e-6
        # variant 2: the persistent kernels with 512-thread blocks; it holds
        # for every group of a kernel or for none
        ens.set_tiles([2] * (2 * len(GROUPS)))
        _poison(ens)
        ens.run()
        wide = _outputs(ens)
        for k in range(len(GROUPS)):
            assert rel_l2(wide[k][0], base[k][0]) < 1e-6
            assert rel_l2(wide[k][1], base[k][1]) < 1e-6
        ens.set_tiles([3] * (2 * len(GROUPS)))
        _poison(ens)
        ens.run()
        wide3 = _outputs(ens)
        for k in range(len(GROUPS)):
            assert rel_l2(wide3[k][0], base[k][0]) < 1e-6
            assert rel_l2(wide3[k][1], base[k][1]) < 1e-6
        with pytest.raises(sht.ExoError):
            ens.set_tiles([2, 0] + [0] * (2 * len(GROUPS) - 2))
        v = ens.autotune()
        assert all(x in (0, 1, 2, 3) for x in v) and ens.tiles() == v
        assert ens.schedule() in ("persistent", "batched")
        for w in (0, 1):  # 512-thread variants for all groups of a kernel or for none
            assert len({x >= 2 for x in v[w::2]}) == 1
        # groups of one truncation share a variant
        _poison(ens)
        ens.run()
        tuned = _outputs(ens)
        for k, (ntru, nlev, nmem) in enumerate(ens.groups):
            g_o, s_o = oracle.roundtrip(ntru, nlev, host[k][0])
            assert rel_l2(tuned[k][0][0], g_o) < 1e-6 and rel_l2(tuned[k][1][0], s_o) < 1e-6
        _poison(ens)
        ens.run()
        assert _same(_outputs(ens), tuned)
    finally:
        ens.close()


FULL = [(21, 10, 48), (42, 20, 32), (63, 20, 16)]


def test_ensemble_full_size_every_member(oracle, gpu):
    # the 96-member ensemble at its benchmark size: every member of every
    # group against the FP64 oracle, field by field, after an autotune
    import torch

    print(f"ensemble {FULL} bf16x3, every member", flush=True)
    ens = sht.Ensemble(FULL, "bf16x3")
    try:
        host = []
        for k, (ntru, nlev, nmem) in enumerate(FULL):
            sp = ensemble(ntru, nlev, nmem, cfg=3)
            ens.bufs[k][0].copy_(torch.from_numpy(sp).float())
            host.append(sp)
        ens.autotune()
        _poison(ens)
        ens.run()
        out = _outputs(ens)
        for k, (ntru, nlev, nmem) in enumerate(FULL):
            tab = oracle.legini(ntru, sht.dims(ntru)[0])["_all"]
            nfl = 3 * nlev + 1
            fields = {"vor": slice(0, nlev), "div": slice(nlev, 2 * nlev),
                      "tmp": slice(2 * nlev, 3 * nlev), "lnps": slice(nfl - 1, nfl)}
            grid_only = {"dmu": slice(nfl, nfl + 1), "u": slice(nfl + 1, nfl + 1 + nlev),
                         "v": slice(nfl + 1 + nlev, nfl + 1 + 2 * nlev)}
            worst = 0.0
            for m in range(nmem):
                g_o, s_o = oracle.roundtrip(ntru, nlev, host[k][m], tables=tab)
                for name, sl in {**fields, **grid_only}.items():
                    e = rel_l2(out[k][0][m][sl], g_o[sl])
                    assert e < 1e-6, (ntru, m, name, "grid", e)
                    worst = max(worst, e)
                for name, sl in fields.items():
                    e = rel_l2(out[k][1][m][sl], s_o[sl])
                    assert e < 1e-6, (ntru, m, name, "spectral", e)
                    worst = max(worst, e)
            print(f"  T{ntru}: {nmem} members, worst per-field rel L2 {worst:.2e}", flush=True)
    finally:
        ens.close()


def test_ensemble_rejects(gpu):
    with pytest.raises(sht.ExoError):
        sht.Ensemble([(127, 4, 1)], "bf16x3")
    with pytest.raises(sht.ExoError):
        sht.Ensemble([(21, 10, 1)], "f64")
