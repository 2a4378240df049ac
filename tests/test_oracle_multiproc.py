"""Pins of the oracle's multi-process pieces: fc2sp's unfolded branch
(ora_fc2sp_unfolded, src/legmod.f90:191-202), mpsumsc (oracle.sumsc,
src/mpimod.f90:279-291) and mrtrunc / mrexpand (oracle.mrtrunc /
oracle.mrexpand, src/mpimod_multi.f90:117-153).

The reference's NPRO > 1 branch needs a resolution module with NPRO > 1,
which its build system generates and this tree does not hold, and
mpimod_multi.f90 needs MPI: neither builds here.  These functions are
pinned by what the mathematics fixes instead: the processes' partial sums
add up to the single-process folded fc2sp (itself pinned against the
compiled reference), one latitude is one product, the reduce-scatter is a
sum followed by a slice, and mrtrunc / mrexpand move coefficient (m, n)
to (m, n).
"""
import numpy as np
import pytest


def rel(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / np.linalg.norm(np.ravel(b)))


@pytest.mark.parametrize("ntru,nproc", [(21, 2), (21, 4), (42, 8)])
def test_unfolded_partials_sum_to_folded(oracle, ntru, nproc):
    from oracle import dims

    nlat, nlon, ncsp = dims(ntru)
    t = oracle.legini(ntru, nlat)
    sk = oracle.filters(ntru, 2, 1, 1)[0]  # the exponential GP->SP filter
    fc = np.random.default_rng(ntru + nproc).standard_normal((nlat, nlon))
    nlpp = nlat // nproc
    parts = [oracle.fc2sp_unfolded(ntru, nlon, t["qc"][r * nlpp:(r + 1) * nlpp], sk,
                                   fc[r * nlpp:(r + 1) * nlpp]) for r in range(nproc)]
    folded = oracle.fc2sp(ntru, nlat, nlon, t["qc"], sk, fc)
    assert rel(np.sum(parts, axis=0), folded) < 1e-14
    # one latitude: one product per coefficient
    p = oracle.fc2sp_unfolded(ntru, nlon, t["qc"][3:4], sk, fc[3:4])
    m = np.concatenate([np.full(ntru + 1 - mm, mm) for mm in range(ntru + 1)])
    n = np.concatenate([np.arange(mm, ntru + 1) for mm in range(ntru + 1)])
    want = np.empty(2 * ncsp)
    want[0::2] = t["qc"][3] * fc[3, 2 * m] * sk[n]
    want[1::2] = t["qc"][3] * fc[3, 2 * m + 1] * sk[n]
    assert np.array_equal(p, want)


def test_sumsc():
    from oracle import sumsc

    rng = np.random.default_rng(5)
    nproc, nfield, nspp = 4, 3, 17
    psf = rng.standard_normal((nproc, nfield, nproc * nspp))
    out = sumsc(psf, nspp)
    total = psf.sum(axis=0)
    for r in range(nproc):
        assert np.allclose(out[r], total[:, r * nspp:(r + 1) * nspp], rtol=1e-15, atol=1e-15)


@pytest.mark.parametrize("nthi,ntlo", [(42, 21), (63, 42), (21, 21), (127, 63)])
def test_mrtrunc_mrexpand(nthi, ntlo):
    from oracle import mrexpand, mrtrunc

    def index(nt):
        return {(mm, nn): w for w, (mm, nn) in enumerate(
            (mm, nn) for mm in range(nt + 1) for nn in range(mm, nt + 1))}

    hi, lo = index(nthi), index(ntlo)
    rng = np.random.default_rng(nthi)
    x = rng.standard_normal(2 * len(hi))
    y = mrtrunc(x, nthi, ntlo)
    for (mm, nn), w in lo.items():
        assert y[2 * w] == x[2 * hi[mm, nn]] and y[2 * w + 1] == x[2 * hi[mm, nn] + 1]
    base = rng.standard_normal(2 * len(hi))
    z = mrexpand(y, ntlo, base, nthi)
    for (mm, nn), w in hi.items():
        if (mm, nn) in lo:
            want = y[2 * lo[mm, nn]:2 * lo[mm, nn] + 2]
        elif w < nthi:
            want = [0.0, 0.0]  # the reference clears sphi(:, 1:nthi) only
        else:
            want = base[2 * w:2 * w + 2]
        assert np.array_equal(z[2 * w:2 * w + 2], want), (mm, nn)
    assert np.array_equal(mrtrunc(z, nthi, ntlo), y)


def test_writesp_closed_form():
    """writesp's normalisation is (-1)^m / sqrt(2) per coefficient of zonal
    wavenumber m, the scale, single precision and the offset on the first
    value only (src/outmod.f90:163-164)"""
    from oracle import writesp

    ntru = 21
    m = np.concatenate([np.full(ntru + 1 - mm, mm) for mm in range(ntru + 1)])
    pf = np.random.default_rng(2).standard_normal(2 * m.size)
    z = writesp(pf, ntru, 3.0, 10.0)
    want = (pf * np.repeat((-1.0) ** m, 2) / np.sqrt(2.0) * 3.0).astype(np.float32)
    want[0] = np.float32(float(want[0]) + 10.0)
    assert z.dtype == np.float32
    assert np.max(np.abs(z.astype(np.float64) - want) / np.abs(want)) < 1e-7
    assert np.array_equal(z[1:], writesp(pf, ntru, 3.0, 0.0)[1:])
