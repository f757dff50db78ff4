"""Pins the float64 restatements of correct_MAS, Xi, FT_filter and field_smoothing (oracle/fields.py)."""
import numpy as np
import pytest

from oracle import fields as F


def test_xi_equals_bruteforce_autocorrelation():
    rng = np.random.default_rng(1)
    N, L = 8, 80.0
    d = rng.standard_normal((N, N, N))
    o = F.Xi(d, L, MAS="None", axis=2)
    # direct: xi(r) = (1/N^3) sum_x d(x) d(x+r), binned by floor(|r|) over all N^3 offsets
    off = F._offsets(N)
    sums = {}
    for a in range(N):
        for b in range(N):
            for c in range(N):
                v = np.mean(d * np.roll(d, (-a, -b, -c), axis=(0, 1, 2)))
                r2 = off[a] ** 2 + off[b] ** 2 + off[c] ** 2
                i = int(np.floor(np.sqrt(r2)))
                mu2 = 0.0 if r2 == 0 else off[c] ** 2 / r2
                s = sums.setdefault(i, [0.0, 0.0, 0, 0.0])
                s[0] += v
                s[1] += v * (3 * mu2 - 1) / 2
                s[2] += 1
                s[3] += np.sqrt(r2)
    for i in range(1, len(o.r3D) + 1):
        s = sums[i]
        assert o.Nmodes3D[i - 1] == s[2]
        assert np.isclose(o.xi[i - 1, 0], s[0] / s[2], rtol=1e-10, atol=1e-14)
        assert np.isclose(o.xi[i - 1, 1], 5 * s[1] / s[2], rtol=1e-10, atol=1e-14)
        assert np.isclose(o.r3D[i - 1], s[3] / s[2] * L / N, rtol=1e-12)


def test_correct_mas_identity_and_plane_waves():
    N, L = 16, 100.0
    rng = np.random.default_rng(2)
    d = rng.standard_normal((N, N, N))
    assert np.allclose(F.correct_MAS(d, L, MAS="None"), d, atol=1e-12)
    x = np.arange(N)
    n = 3
    W = (np.pi * n / N / np.sin(np.pi * n / N)) ** 2        # CIC window of wavenumber n
    # along z: both signs of kz are the one independent mode kz = n: fully corrected
    wz = np.cos(2 * np.pi * n * x / N)[None, None, :] * np.ones((N, N, 1))
    assert np.allclose(F.correct_MAS(wz, L, "CIC"), W * wz, atol=1e-12)
    # along x on the kz = 0 plane: +n is independent (corrected), -n is the dependent partner
    # (left as it is): the real part of the inverse is (W + 1)/2 cos
    wx = np.cos(2 * np.pi * n * x / N)[:, None, None] * np.ones((1, N, N))
    assert np.allclose(F.correct_MAS(wx, L, "CIC"), (W + 1) / 2 * wx, atol=1e-12)


@pytest.mark.parametrize("Filter", ["Top-Hat", "Gaussian"])
def test_filter_normalisation_and_values(Filter):
    N, L, R = 16, 100.0, 20.0
    f = F.filter_field(L, R, N, Filter)
    assert np.isclose(f.sum(), 1.0, rtol=1e-14)
    fk = F.FT_filter(L, R, N, Filter)
    assert np.isclose(fk[0, 0, 0].real, 1.0, rtol=1e-14)
    Rg2 = (R * N / L) ** 2
    if Filter == "Top-Hat":
        # brute-force count of the cells within R (cell units, periodic offsets)
        cnt = sum(1 for a in range(N) for b in range(N) for c in range(N)
                  if F._offsets(N)[a] ** 2 + F._offsets(N)[b] ** 2 + F._offsets(N)[c] ** 2 <= Rg2)
        assert np.isclose(f[0, 0, 0], 1.0 / cnt)
    else:
        assert np.isclose(f[1, 0, 0] / f[0, 0, 0], np.exp(-1.0 / (2 * Rg2)), rtol=1e-6)


def test_smoothing_identity_and_constant():
    N, L = 16, 100.0
    rng = np.random.default_rng(3)
    d = rng.standard_normal((N, N, N))
    point = F.FT_filter(L, 0.1, N, "Top-Hat")           # only the origin cell: identity
    assert np.allclose(F.field_smoothing(d, point), d, atol=1e-12)
    c = np.full((N, N, N), 2.5)
    assert np.allclose(F.field_smoothing(c, F.FT_filter(L, 30.0, N, "Gaussian")), c, atol=1e-12)
