"""Pins of the oracle's gridpointa (Oracle.gridpointa: the transform group
of one dynamics step, src/plasim.f90:2402-2486, composed of the oracle's
restated transforms and calcgp) against the reference's own gridpointa
compiled with legmod, gaussmod and fftmod (oracle/_ref/libref_gpa.so):
golden tendencies (tests/golden/reference_gridpointa_t21.npz) and, where
oracle/_ref is built, the live library; and against what the equations fix
without the reference: an atmosphere at rest on the basic state has no
tendencies at all.
"""
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
from make_golden import (GPA_ROWS, PLASIM_SCAL, gridpointa_humidity,  # noqa: E402
                         gridpointa_state, plasim_vertical)

GOLD = os.path.join(HERE, "golden", "reference_gridpointa_t21.npz")
REF = os.path.join(os.path.dirname(HERE), "oracle", "_ref", "libref_gpa.so")
N = 10


def rel(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / np.linalg.norm(np.ravel(b)))


@pytest.fixture(scope="module")
def setup(oracle):
    v = plasim_vertical(oracle, N)
    gq, dq = gridpointa_humidity(N, 32 * 64)
    return v, gridpointa_state(21, N), gq, dq


@pytest.mark.parametrize("nq", [1, 0])
def test_gridpointa_vs_golden(oracle, setup, nq):
    v, st, gq, dq = setup
    g = np.load(GOLD)
    tend, f = oracle.gridpointa(21, N, st, v, PLASIM_SCAL, nq, gq=gq, dq=dq)
    rows = [k * N + j for k in range(4) for j in GPA_ROWS] + [4 * N]
    want = g[f"tend{nq}"]
    for i, r in enumerate(rows):
        if nq == 0 and 3 * N <= r < 4 * N:
            assert not want[i].any() and not tend[r].any()  # no qtend without nqspec
            continue
        assert rel(tend[r], want[i]) < 1e-13, (nq, r, rel(tend[r], want[i]))
    assert rel(f["gp"], g[f"gp{nq}"]) < 1e-15


@pytest.mark.skipif(not os.path.exists(REF), reason="oracle/_ref not built (needs the "
                    "reference tree)")
@pytest.mark.parametrize("nq", [1, 0])
def test_gridpointa_vs_live_reference(oracle, setup, nq):
    from oracle import RefGridpointa

    v, st, gq, dq = setup
    # another state than the golden one
    st = st.copy()
    st[N:] *= -1.7
    tr, gr = RefGridpointa().run(st, v, PLASIM_SCAL, nq, gq=gq, dq=dq)
    to, fo = oracle.gridpointa(21, N, st, v, PLASIM_SCAL, nq, gq=gq, dq=dq)
    for k in range(4 if nq else 3):
        assert rel(to[k * N:(k + 1) * N], tr[k * N:(k + 1) * N]) < 1e-13, k
    assert rel(to[4 * N], tr[4 * N]) < 1e-13
    for k, x in gr.items():
        assert rel(fo[k], x) < 1e-14, k


def test_gridpointa_rest_state(oracle, setup):
    """At rest on the basic state (absolute vorticity = plavor only, no
    divergence, T = t0, ln ps = 0, no humidity) every wind, vertical
    velocity and tendency vanishes.  A solid-body rotation on top of it (a
    relative vorticity in the (m=0, n=1) mode) is a steady state of the
    vorticity, temperature, humidity and surface-pressure equations, so only
    the divergence tendency is nonzero (kinetic energy and Coriolis force);
    taking mktend's fu and fv the wrong way round would give a vorticity
    tendency."""
    v = setup[0]
    st = np.zeros((4 * N + 1, 22 * 23))
    st[:N, 2] = 1.63299310207
    tend, f = oracle.gridpointa(21, N, st, v, PLASIM_SCAL, 1)
    assert np.abs(f["gu"]).max() < 1e-15 and np.abs(f["gv"]).max() < 1e-15
    assert np.abs(tend).max() < 1e-14
    st[:N, 2] += 0.5
    tend, f = oracle.gridpointa(21, N, st, v, PLASIM_SCAL, 1)
    assert np.abs(f["gu"]).max() > 0.1
    assert np.abs(tend[N:2 * N]).max() > 0.1
    for k in (0, 2, 3):
        assert np.abs(tend[k * N:(k + 1) * N]).max() < 1e-14, k
    assert np.abs(tend[4 * N]).max() < 1e-14
