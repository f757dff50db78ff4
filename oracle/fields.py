"""Test infrastructure: float64 restatements of Pk_library.correct_MAS (Pk_library.pyx:1749-1806),
Pk_library.Xi (:2035-2151) and smoothing_library.FT_filter / field_smoothing
(smoothing_library.pyx:20-114).  numpy's rfftn / irfftn stand in for FFTW (irfftn, like FFTW's
c2r, drops the imaginary parts of the DC and Nyquist terms of each last-axis row).

Pinned by tests/test_oracle_fields.py: brute-force real-space autocorrelation for Xi, the closed
form of correct_MAS on plane waves (including the reference's uncorrected dependent half-plane),
the filter's unit normalisation and brute-force cell values, identity and constant smoothing.
"""
import numpy as np

from .pk import MAS_function, frequencies


def _window(N, p):
    k = np.fft.fftfreq(N, 1.0 / N).astype(np.int64)
    k[k == -(N // 2)] = N // 2 if N % 2 == 0 else k[k == -(N // 2)]   # index N/2 is +N/2 (Pk_library.pyx:315)
    x = np.pi / N * k
    with np.errstate(invalid="ignore", divide="ignore"):
        w = np.where(x == 0, 1.0, (x / np.sin(x)) ** p)
    return w, k


def _factor(N, p):
    w, _ = _window(N, p)
    return w[:, None, None] * w[None, :, None] * w[None, None, : N // 2 + 1]


def _independent(N):
    _, k = _window(N, 0)
    middle = N // 2
    kx = k[:, None, None]
    ky = k[None, :, None]
    kz = np.arange(N // 2 + 1)[None, None, :]
    plane = (kz == 0) | ((kz == middle) & (N % 2 == 0))
    dep = plane & ((kx < 0) | (((kx == 0) | ((kx == middle) & (N % 2 == 0))) & (ky < 0)))
    return ~dep


def correct_MAS(delta, BoxSize, MAS="CIC"):
    d = np.asarray(delta, dtype=np.float64)
    N = d.shape[0]
    dk = np.fft.rfftn(d)
    f = _factor(N, MAS_function(MAS))
    dk = np.where(_independent(N), dk * f, dk)
    return np.fft.irfftn(dk, s=d.shape, axes=(0, 1, 2))


def _offsets(N):
    i = np.arange(N)
    return np.where(i > N // 2, i - N, i)


class Xi:
    def __init__(self, delta, BoxSize, MAS="CIC", axis=2):
        BoxSize = float(np.float32(BoxSize))
        d = np.asarray(delta, dtype=np.float64)
        N = d.shape[0]
        _, _, _, _, kmax = frequencies(BoxSize, N)
        dk = np.fft.rfftn(d) * _factor(N, MAS_function(MAS))
        xi_f = np.fft.irfftn(np.abs(dk) ** 2, s=d.shape, axes=(0, 1, 2))
        o = _offsets(N)
        X, Y, Z = np.meshgrid(o, o, o, indexing="ij")
        k2 = X * X + Y * Y + Z * Z
        k = np.sqrt(k2.astype(np.float64))
        idx = np.floor(k).astype(np.int64)
        par = (X, Y, Z)[axis]
        with np.errstate(invalid="ignore", divide="ignore"):
            mu2 = np.where(k == 0, 0.0, (par / k) ** 2)
        L2 = (3.0 * mu2 - 1.0) / 2.0
        L4 = (35.0 * mu2 * mu2 - 30.0 * mu2 + 3.0) / 8.0
        nb = kmax + 1
        cnt = np.bincount(idx.ravel(), minlength=nb)[:nb].astype(np.float64)
        r = np.bincount(idx.ravel(), weights=k.ravel(), minlength=nb)[:nb]
        x0 = np.bincount(idx.ravel(), weights=xi_f.ravel(), minlength=nb)[:nb]
        x2 = np.bincount(idx.ravel(), weights=(xi_f * L2).ravel(), minlength=nb)[:nb]
        x4 = np.bincount(idx.ravel(), weights=(xi_f * L4).ravel(), minlength=nb)[:nb]
        n = cnt[1:]
        self.r3D = r[1:] / n * (BoxSize * 1.0 / N)
        self.xi = np.stack([x0[1:] / n, 5.0 * x2[1:] / n, 9.0 * x4[1:] / n], axis=1) / N ** 3
        self.Nmodes3D = n


def filter_field(BoxSize, R, dims, Filter):
    """The normalised filter on the grid (smoothing_library.pyx:29-75), float64."""
    R_grid = np.float32(np.float32(R) * np.float32(dims) / np.float32(BoxSize))
    R2 = np.float32(R_grid * R_grid)
    o = _offsets(dims)
    X, Y, Z = np.meshgrid(o, o, o, indexing="ij")
    d2 = X * X + Y * Y + Z * Z
    if Filter == "Top-Hat":
        f = (d2.astype(np.float32) <= R2).astype(np.float64)
    elif Filter == "Gaussian":
        f = np.exp(-d2 / (2.0 * float(R2))).astype(np.float32).astype(np.float64)
    else:
        raise Exception("Filter %s not implemented!" % Filter)
    return f / f.sum()


def FT_filter(BoxSize, R, dims, Filter):
    return np.fft.rfftn(filter_field(BoxSize, R, dims, Filter))


def field_smoothing(field, filter_k):
    f = np.asarray(field, dtype=np.float64)
    return np.fft.irfftn(np.fft.rfftn(f) * np.asarray(filter_k), s=f.shape, axes=(0, 1, 2))
