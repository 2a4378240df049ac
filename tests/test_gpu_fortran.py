"""The Fortran drop-in on the GPU: a Fortran program calls the reference's
subroutine names (inigau, sp2fc, sp2fcdmu, fc2sp, dv2uv, uv2dv, qtend,
mktend, fc2gp, gp2fc) through exoplasim_amd/fortran (ISO_C_BINDING) and
libexosht.so; its results are compared with the oracle on the same inputs.
"""
import os
import subprocess
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, os.path.join(HERE, "golden"))
from make_golden import golden_input, spectral_input  # noqa: E402

pytestmark = pytest.mark.gpu
BIN = os.path.join(ROOT, "exoplasim_amd", "fortran", "bin")


def rel_l2(a, b):
    a = np.ravel(a)
    b = np.ravel(b)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("tag,ntru,nlat,nlev,tol", [("r8", 21, 32, 10, 1e-13),
                                                    ("r4", 42, 64, 20, 2e-6)])
def test_fortran_dropin(oracle, gpu, tmp_path, tag, ntru, nlat, nlev, tol):
    exe = os.path.join(BIN, f"driver_{tag}")
    if not os.path.exists(exe):
        pytest.fail(f"{exe} not built (exoplasim_amd.build.build_fortran_driver)")
    out = tmp_path / "out.bin"
    print(f"fortran driver {tag} T{ntru} L{nlev}", flush=True)
    res = subprocess.run([exe, str(ntru), str(nlat), str(nlev), str(out)], capture_output=True,
                         text=True, timeout=120)
    assert res.returncode == 0, res.stdout + res.stderr
    nlon = 2 * nlat
    ncsp = (ntru + 1) * (ntru + 2) // 2
    nhor = nlat * nlon
    v = np.fromfile(out, dtype=np.float64)
    nfl = 3 * nlev + 1
    sizes = [("sp", 2 * ncsp * nfl), ("f", nhor * nlev * 6), ("sid", nlat), ("gwd", nlat), ("fc0", nhor), ("fcd", nhor), ("sp0", 2 * ncsp),
             ("pu", nhor * nlev), ("pv", nhor * nlev), ("pd", 2 * ncsp * nlev),
             ("pz", 2 * ncsp * nlev), ("q", 2 * ncsp * nlev), ("d", 2 * ncsp * nlev),
             ("t", 2 * ncsp * nlev), ("z", 2 * ncsp * nlev), ("g", nhor), ("g2", nhor)]
    assert v.size == sum(s for _, s in sizes)
    r, o = {}, 0
    for name, s in sizes:
        r[name] = v[o:o + s]
        o += s
    sid, gwd = oracle.inigau(nlat)
    assert np.abs(r["sid"] - sid).max() < 1e-15
    tab = oracle.legini(ntru, nlat, r["sid"], r["gwd"])
    one = np.ones(ntru + 1)
    # the inputs the program used (its sin/cos/** round differently from numpy)
    sp = r["sp"].reshape(nfl, 2 * ncsp)
    f = list(r["f"].reshape(6, nlev, nlat, nlon))
    assert np.abs(sp - spectral_input(ntru, nfl, 0.3)).max() < 1e-6
    assert np.abs(f[1] - golden_input((nlev, nlat, nlon), 1.6)).max() < 1e-6
    fc0 = oracle.sp2fc(ntru, nlat, nlon, tab["qi"], one, sp[0])
    assert rel_l2(r["fc0"], fc0) < tol
    assert rel_l2(r["fcd"], oracle.sp2fc(ntru, nlat, nlon, tab["qj"], one, sp[-1])) < tol
    assert rel_l2(r["sp0"], oracle.fc2sp(ntru, nlat, nlon, tab["qc"], one, r["fc0"])) < tol
    pu, pv = oracle.dv2uv(ntru, nlat, nlon, nlev, tab["qu"], tab["qv"], one, 1.63299310207,
                          sp[nlev:2 * nlev], sp[:nlev])
    assert rel_l2(r["pu"], pu) < tol and rel_l2(r["pv"], pv) < tol
    pd, pz = oracle.uv2dv(ntru, nlat, nlon, nlev, tab["qe"], tab["qm"], one, f[0], f[1])
    assert rel_l2(r["pd"], pd) < tol and rel_l2(r["pz"], pz) < tol
    q = oracle.qtend(ntru, nlat, nlon, nlev, tab["qc"], tab["qe"], tab["qm"], one, f[0], f[1], f[2])
    assert rel_l2(r["q"], q) < tol
    d, t, z = oracle.mktend(ntru, nlat, nlon, nlev, tab["qc"], tab["qe"], tab["qm"], tab["qq"],
                            one, *f)
    assert rel_l2(r["d"], d) < tol and rel_l2(r["t"], t) < tol and rel_l2(r["z"], z) < tol
    g = oracle.fc2gp(r["fc0"].reshape(nlat, nlon))
    assert rel_l2(r["g"], g) < tol
    assert rel_l2(r["g2"], oracle.gp2fc(r["g"].reshape(nlat, nlon))) < tol


REF_BIN = os.path.join(ROOT, "oracle", "_ref")


@pytest.mark.parametrize("tag,nf,nq,tol", [("r8", 2, 1, 1e-13), ("r8", 4, 0, 1e-13),
                                           ("r8", 0, 1, 1e-13), ("r4", 3, 1, 2e-6)])
def test_fortran_model_composites(oracle, gpu, tmp_path, tag, nf, nq, tol):
    """legini (with physics filter nf on both directions), invlega, invlegd,
    sp3fc and sp2fl called by their reference names from a program built
    against the reference's pumamod (T21 L10): every pumamod array they fill
    against the oracle on the program's own spectral state."""
    exe = os.path.join(REF_BIN, f"driver_model_{tag}")
    if not os.path.exists(exe):
        pytest.skip(f"{exe} not built (make -C oracle ref, needs the reference tree)")
    out = tmp_path / "model.bin"
    print(f"fortran model driver {tag} nfilter={nf} nqspec={nq}", flush=True)
    res = subprocess.run([exe, str(nf), str(nq), str(out)], capture_output=True, text=True,
                         timeout=120)
    assert res.returncode == 0, res.stdout + res.stderr
    ntru, nlat, nlon, nlev = 21, 32, 64, 10
    nesp, nhor = (ntru + 1) * (ntru + 2), nlat * nlon
    v = np.fromfile(out, dtype=np.float64)
    layout = [("sd", nesp * nlev), ("st", nesp * nlev), ("sz", nesp * nlev),
              ("sq", nesp * nlev), ("sp", nesp)]
    layout += [(f"a_{k}", nhor * nlev) for k in ("gd", "gt", "gz", "gq", "gu", "gv")]
    layout += [("a_gp", nhor), ("a_gpj", nhor)]
    layout += [(f"d_{k}", nhor * nlev) for k in ("gt", "gq", "gu", "gv")] + [("d_gp", nhor)]
    layout += [(f"s_{k}", nhor * nlev) for k in ("gd", "gt", "gz")] + [("fl_sq", nhor * nlev)]
    assert v.size == sum(s for _, s in layout)
    r, o = {}, 0
    for name, s in layout:
        shape = (nlev, nlat, nlon) if s == nhor * nlev else ((nlev, nesp) if s == nesp * nlev
                                                             else (-1,))
        r[name] = v[o:o + s].reshape(shape)
        o += s
    sid, gwd = oracle.inigau(nlat)
    tab = oracle.legini(ntru, nlat, sid, gwd)
    _, s = oracle.filters(ntru, nf, 1, 1)
    fc = lambda sp, t="qi": oracle.sp2fc(ntru, nlat, nlon, tab[t], s, sp)  # noqa: E731
    lev = lambda a: np.stack([fc(a[j]) for j in range(nlev)])  # noqa: E731
    plavor = 1.63299310207 if tag == "r8" else float(np.float32(1.63299310207))
    pu, pv = oracle.dv2uv(ntru, nlat, nlon, nlev, tab["qu"], tab["qv"], s, plavor, r["sd"],
                          r["sz"])
    want = {"a_gd": lev(r["sd"]), "a_gt": lev(r["st"]), "a_gz": lev(r["sz"]),
            "a_gq": lev(r["sq"]) if nq == 1 else np.zeros((nlev, nlat, nlon)),
            "a_gu": pu, "a_gv": pv, "a_gp": fc(r["sp"]).ravel(),
            "a_gpj": fc(r["sp"], "qj").ravel(),
            "d_gt": lev(r["st"]), "d_gq": lev(r["sq"]) if nq == 1 else np.zeros((nlev, nlat, nlon)),
            "d_gu": pu, "d_gv": pv, "d_gp": fc(r["sp"]).ravel(),
            "s_gd": lev(r["sd"]), "s_gt": lev(r["st"]), "s_gz": lev(r["sz"]),
            "fl_sq": lev(r["sq"])}
    for k, w in want.items():
        if not np.any(w):
            assert np.all(r[k] == 0.0), k
        else:
            assert rel_l2(r[k], w) < tol, (k, rel_l2(r[k], w))
