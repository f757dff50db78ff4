"""Pins the bispectrum oracle (oracle/bk.py) by brute-force enumeration of closed triangles
k1 + k2 + k3 = 0 (mod N) over the full-space shells, which equals the real-space triple sum of
the inverse-transformed shell fields by the convolution theorem."""
import itertools

import numpy as np
import pytest

from oracle import bk as Bo
from oracle import pk as P


def test_brute_force_triangles():
    N, L = 12, 2 * np.pi          # kF = 1: shells in integer units
    rng = np.random.default_rng(0)
    delta = rng.standard_normal((N, N, N))
    k1, k2 = 2.0, 3.0
    theta = np.array([0.3, 1.2, 2.5])
    b = Bo.Bk(delta, L, k1, k2, theta, MAS="CIC")
    full = np.fft.fftn(delta)
    rep = lambda i: i - N if i > N // 2 else i
    win = lambda i: P.MAS_correction(np.pi * rep(i) / N, 2)
    allk = list(itertools.product(range(N), repeat=3))
    kabs = {k: np.sqrt(sum(rep(c) ** 2 for c in k)) for k in allk}
    W = {k: win(k[0]) * win(k[1]) * win(k[2]) for k in allk}
    k_all, kmin, kmax = Bo.shell_bounds(L, N, k1, k2, theta)
    shells = [[k for k in allk if kmin[i] <= kabs[k] < kmax[i]] for i in range(len(k_all))]
    unitP = (L / N ** 2) ** 3
    unitB = (L ** 2 / N ** 3) ** 3
    for i, S in enumerate(shells):
        Pi = np.mean([abs(full[k] * W[k]) ** 2 for k in S]) * unitP
        assert b.Pk[i] == pytest.approx(Pi, rel=1e-10)
    S1, S2 = shells[0], shells[1]
    for j in range(len(theta)):
        S3 = set(shells[j + 2])
        tot, cnt = 0.0, 0
        for a in S1:
            for c in S2:
                t = tuple((-a[d] - c[d]) % N for d in range(3))
                if t in S3:
                    tot += (full[a] * W[a] * full[c] * W[c] * full[t] * W[t]).real
                    cnt += 1
        assert cnt > 0
        assert b.B[j] == pytest.approx(tot / cnt * unitB, rel=1e-9)
        Pj = b.Pk
        assert b.Q[j] == pytest.approx(b.B[j] / (Pj[0] * Pj[1] + Pj[0] * Pj[j + 2] + Pj[1] * Pj[j + 2]))
    assert np.allclose(b.k[2:], np.sqrt((k2 * np.sin(theta)) ** 2 + (k2 * np.cos(theta) + k1) ** 2))


def test_empty_shell_raises():
    N = 8
    delta = np.random.default_rng(1).standard_normal((N, N, N))
    with pytest.raises(ZeroDivisionError):
        Bo.Bk(delta, 2 * np.pi, 100.0, 2.0, np.array([0.5]))


@pytest.mark.parametrize("N", [8, 12, 16, 24])
def test_oracle_matches_plane_wave_closed_form(N):
    """The FP64 oracle and the plane-wave closed form (which the GPU tests use at every FFT size)
    agree: both restate bispectrum_library.pyx:32-196, one through FFTs, one through lattice counts."""
    from tests._bkform import TRIANGLE_WAVES, closed_form_bk, plane_wave_field, plane_wave_theta
    L = 2 * np.pi * 50.0
    kF = 2 * np.pi / L
    o = Bo.Bk(plane_wave_field(N), L, 3 * kF, 2 * kF, plane_wave_theta(), MAS="None")
    B, Pk = closed_form_bk(TRIANGLE_WAVES, N, L, o.k)
    assert np.allclose(o.Pk, Pk, rtol=1e-12, atol=1e-12 * np.abs(Pk).max())
    assert np.allclose(o.B, B, rtol=1e-12, atol=0)
    assert np.all(B != 0) and np.all(Pk > 0)
