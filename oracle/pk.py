"""FP64 oracle of Pylians' power-spectrum path (TEST INFRASTRUCTURE ONLY).

Restates in float64 arithmetic:

* ``frequencies``       Pk_library.pyx:59-64
* ``MAS_function``      Pk_library.pyx:75-81
* ``MAS_correction``    Pk_library.pyx:86-87
* ``check_number_modes`` Pk_library.pyx:90-102
* ``FFT3Dr`` / ``IFFT3Dr`` the FFTW r2c (unnormalised) / c2r (1/N^3 normalised)
  conventions of FFT3Dr_f / IFFT3Dr_f, Pk_library.pyx:120-165 (pyFFTW's
  ``FFTW.__call__`` normalises the inverse by default).  numpy's pocketfft uses the
  same conventions and the same (N, N, N//2+1) layout.
* ``class Pk``          Pk_library.pyx:266-425 (mode loop :314-381, outputs :387-425)

Deviation: the reference rounds the window factor to float32 and multiplies the
complex64 spectrum (Pk_library.pyx:275, 354-355); the oracle keeps float64.

Pinned by (tests/test_oracle_pk.py): the reference's mode-count invariant
(N^3-8)/2+8 (Pk_library.pyx:90-102) and a brute-force conjugate-pair enumeration
of the full k-grid for Nmodes/k-means in 1-D/2-D/3-D; plane-wave closed forms
(P0 = L^3/(4 Nmodes), P2 = 5 P0 and P4 = 9 P0 along the line of sight, P2 = -2.5 P0
and P4 = 3.375 P0 across it); the window closed form (x/sin x)^p; Parseval.
"""
import math
import os

import numpy as np


def frequencies(BoxSize, dims):
    """kF, kN, kmax_par, kmax_per, kmax (Pk_library.pyx:59-64, with dims//2)."""
    kF = 2.0 * np.pi / BoxSize
    middle = dims // 2
    kN = middle * kF
    kmax_par = middle
    kmax_per = int(np.sqrt(middle ** 2 + middle ** 2))
    kmax = int(np.sqrt(middle ** 2 + middle ** 2 + middle ** 2))
    return kF, kN, kmax_par, kmax_per, kmax


def MAS_function(MAS):
    """Window exponent p (Pk_library.pyx:75-81); anything unknown -> 0 (no correction)."""
    return {"NGP": 1, "CIC": 2, "TSC": 3, "PCS": 4}.get(MAS, 0)


def MAS_correction(x, MAS_index):
    """(x/sin x)^p, 1 at x == 0 (Pk_library.pyx:86-87); vectorised."""
    x = np.asarray(x, dtype=np.float64)
    with np.errstate(invalid="ignore", divide="ignore"):
        r = np.where(x == 0.0, 1.0, (x / np.sin(x)) ** MAS_index)
    return r


def indep_modes_expected(dims):
    """(N^3 - own)/2 + own, own = 8 for even N and 1 for odd (Pk_library.pyx:92-97)."""
    own = 1 if dims % 2 == 1 else 8
    return (dims ** 3 - own) // 2 + own


def check_number_modes(Nmodes, dims):
    """Raise SystemExit when not all independent modes were counted (Pk_library.pyx:90-102)."""
    if int(np.sum(Nmodes)) != indep_modes_expected(dims):
        raise SystemExit("WARNING: Not all modes counted")


def _workers():
    try:
        return max(1, min(16, len(os.sched_getaffinity(0))))
    except AttributeError:
        return 1


def FFT3Dr(a):
    """Unnormalised forward r2c, (N,N,N) -> (N,N,N//2+1), complex128 (pocketfft in float64, the
    library behind numpy.fft, run on the host's cores through scipy.fft)."""
    import scipy.fft
    return scipy.fft.rfftn(np.asarray(a, dtype=np.float64), workers=_workers())


def IFFT3Dr(a, dims=None):
    """1/N^3-normalised c2r, (N,N,N//2+1) -> (N,N,N), float64."""
    import scipy.fft
    a = np.asarray(a, dtype=np.complex128)
    n = a.shape[0] if dims is None else dims
    return scipy.fft.irfftn(a, s=(n, n, n), axes=(0, 1, 2), workers=_workers())


def mode_grid(dims):
    """Signed integer wavenumbers (kx, ky, kz) of the half spectrum, broadcastable
    to (N, N, N//2+1), and the independent-mode mask (Pk_library.pyx:314-330)."""
    middle = dims // 2
    kk = np.arange(dims)
    k = np.where(kk > middle, kk - dims, kk).astype(np.int64)
    kz = np.arange(middle + 1, dtype=np.int64)
    kx = k[:, None, None]
    ky = k[None, :, None]
    kzb = kz[None, None, :]
    even = dims % 2 == 0
    special_z = (kzb == 0) | ((kzb == middle) & even)
    special_x = (kx == 0) | ((kx == middle) & even)
    skip = special_z & ((kx < 0) | (special_x & (ky < 0)))
    return kx, ky, kzb, ~skip


def isqrt_array(n):
    """Exact floor(sqrt(n)) for a non-negative int64 array."""
    n = np.asarray(n, dtype=np.int64)
    r = np.floor(np.sqrt(n.astype(np.float64))).astype(np.int64)
    r = np.where(r * r > n, r - 1, r)
    r = np.where((r + 1) * (r + 1) <= n, r + 1, r)
    return r


def pk_sums(delta_k, dims, axis, MAS_index):
    """Raw per-bin sums of the Pk mode loop (Pk_library.pyx:314-381), float64.

    delta_k: (N,N,N//2+1) spectrum.  Returns a dict of un-normalised histograms and
    the window-corrected spectrum (only the independent modes are corrected, as
    in the reference where the `continue` precedes the correction).
    """
    middle = dims // 2
    kF, kN, kmax_par, kmax_per, kmax = frequencies(1.0, dims)
    kx, ky, kz, indep = mode_grid(dims)
    shape = (dims, dims, middle + 1)
    prefact = np.pi / dims
    corr = (MAS_correction(prefact * kx, MAS_index) * MAS_correction(prefact * ky, MAS_index)
            * MAS_correction(prefact * kz, MAS_index))
    corr = np.broadcast_to(corr, shape)
    indep = np.broadcast_to(indep, shape)
    dk = np.array(delta_k, dtype=np.complex128, copy=True)
    dk[indep] *= corr[indep]

    KX, KY, KZ = (np.broadcast_to(v, shape) for v in (kx, ky, kz))
    k2 = KX * KX + KY * KY + KZ * KZ
    kpar_s = (KX, KY, KZ)[axis]
    perp = [v for i, v in enumerate((KX, KY, KZ)) if i != axis]
    k_per = isqrt_array(perp[0] * perp[0] + perp[1] * perp[1])
    sel = indep
    k2 = k2[sel]; kpar_s = kpar_s[sel]; k_per = k_per[sel]
    d = dk[sel]
    k = np.sqrt(k2.astype(np.float64))
    k_index = isqrt_array(k2)
    with np.errstate(invalid="ignore", divide="ignore"):
        mu = np.where(k == 0, 0.0, kpar_s / k)
    mu2 = mu * mu
    k_par = np.abs(kpar_s)
    re, im = d.real, d.imag
    delta2 = re * re + im * im
    phase = np.arctan2(re, np.sqrt(delta2))

    n3 = kmax + 1
    out = {}
    m1 = k2 <= middle * middle
    out["k1D"] = np.bincount(k_par[m1], weights=k_par[m1].astype(np.float64), minlength=kmax_par + 1)
    out["Pk1D"] = np.bincount(k_par[m1], weights=delta2[m1], minlength=kmax_par + 1)
    out["Nmodes1D"] = np.bincount(k_par[m1], minlength=kmax_par + 1).astype(np.float64)
    i2 = (kmax_par + 1) * k_per + k_par
    n2 = (kmax_par + 1) * (kmax_per + 1)
    out["Pk2D"] = np.bincount(i2, weights=delta2, minlength=n2)
    out["Nmodes2D"] = np.bincount(i2, minlength=n2).astype(np.float64)
    out["k3D"] = np.bincount(k_index, weights=k, minlength=n3)
    out["P0"] = np.bincount(k_index, weights=delta2, minlength=n3)
    out["P2"] = np.bincount(k_index, weights=delta2 * (3.0 * mu2 - 1.0) / 2.0, minlength=n3)
    out["P4"] = np.bincount(k_index, weights=delta2 * (35.0 * mu2 * mu2 - 30.0 * mu2 + 3.0) / 8.0,
                            minlength=n3)
    out["Pkphase"] = np.bincount(k_index, weights=phase * phase, minlength=n3)
    out["Nmodes3D"] = np.bincount(k_index, minlength=n3).astype(np.float64)
    return out, dk


def _div(a, b):
    """Cython cdivision(False) semantics: division by zero raises ZeroDivisionError."""
    if np.any(b == 0):
        raise ZeroDivisionError("float division")
    return a / b


class Pk:
    """float64 restatement of ``Pk_library.Pk`` (Pk_library.pyx:266-425).

    Attributes: k1D, Pk1D, Nmodes1D, kpar, kper, Pk2D, Nmodes2D, k3D,
    Pk (kmax, 3) = P0/P2/P4, Pkphase, Nmodes3D [, delta_k].
    """

    def __init__(self, delta, BoxSize, axis=2, MAS="CIC", threads=1, keep_deltak=False,
                 delta_k=None, sums=None, dims=None):
        if sums is None:
            delta = np.asarray(delta)
            dims = delta.shape[0] if delta_k is None else delta_k.shape[0]
            if delta_k is None:
                delta_k = FFT3Dr(delta)
            sums, dk = pk_sums(delta_k, dims, axis, MAS_function(MAS))
            if keep_deltak:
                self.delta_k = dk
        self._finalize(sums, dims, float(BoxSize))

    def _finalize(self, s, dims, BoxSize):
        """Normalisation and outputs, Pk_library.pyx:387-425."""
        kF, kN, kmax_par, kmax_per, kmax = frequencies(BoxSize, dims)
        unit = (BoxSize / dims ** 2) ** 3

        # 1-D (Pk_library.pyx:387-395): drop the DC bin, area factor pi kmaxper^2/Nmodes/(2pi)^2
        k1D, N1D, P1D = s["k1D"][1:], s["Nmodes1D"][1:], s["Pk1D"][1:]
        P1D = P1D * unit
        k1D = _div(k1D, N1D) * kF
        kmaxper = np.sqrt(kN ** 2 - k1D ** 2)
        P1D = P1D * _div(np.pi * kmaxper ** 2, N1D) / (2.0 * np.pi) ** 2
        self.k1D, self.Pk1D, self.Nmodes1D = k1D, P1D, N1D

        # 2-D (Pk_library.pyx:397-408): keep the DC bin, bin-centre coordinates
        kp = np.arange(kmax_par + 1)
        kq = np.arange(kmax_per + 1)
        self.kpar = np.tile(0.5 * (2 * kp + 1) * kF, kmax_per + 1)
        self.kper = np.repeat(0.5 * (2 * kq + 1) * kF, kmax_par + 1)
        self.Pk2D = _div(s["Pk2D"] * unit, s["Nmodes2D"])
        self.Nmodes2D = s["Nmodes2D"]

        # 3-D (Pk_library.pyx:410-421)
        check_number_modes(s["Nmodes3D"], dims)
        N3 = s["Nmodes3D"][1:]
        self.k3D = _div(s["k3D"][1:], N3) * kF
        P = np.empty((kmax, 3))
        P[:, 0] = _div(s["P0"][1:], N3) * unit
        P[:, 1] = _div(s["P2"][1:] * 5.0, N3) * unit
        P[:, 2] = _div(s["P4"][1:] * 9.0, N3) * unit
        self.Pk = P
        self.Pkphase = _div(s["Pkphase"][1:], N3) * unit
        self.Nmodes3D = N3


class XPk:
    """float64 restatement of ``Pk_library.XPk`` (Pk_library.pyx:534-798): auto and cross
    multipoles of several fields, each with its own window (MAS list).

    Attributes: k1D, Pk1D (kmax_par, F), PkX1D (kmax_par, X), Nmodes1D, kpar, kper,
    Pk2D (n2, F), PkX2D (n2, X), Nmodes2D, k3D, Pk (kmax, 3, F), XPk (kmax, 3, X), Nmodes3D;
    crosses ordered (0,1), (0,2), ..., (1,2), ... (Pk_library.pyx:714-728).
    """

    def __init__(self, delta, BoxSize, axis=2, MAS=None, threads=1):
        fields = len(delta)
        dims = np.asarray(delta[0]).shape[0]
        for d in delta[1:]:
            if np.asarray(d).shape[0] != dims:
                raise SystemExit("Fields have different grid sizes!!!")
        BoxSize = float(BoxSize)
        kF, kN, kmax_par, kmax_per, kmax = frequencies(BoxSize, dims)
        middle = dims // 2
        kx, ky, kz, indep = mode_grid(dims)
        shape = (dims, dims, middle + 1)
        prefact = np.pi / dims
        indep = np.broadcast_to(indep, shape)
        KX, KY, KZ = (np.broadcast_to(v, shape) for v in (kx, ky, kz))
        k2 = (KX * KX + KY * KY + KZ * KZ)[indep]
        kpar_s = (KX, KY, KZ)[axis][indep]
        perp = [v for i, v in enumerate((KX, KY, KZ)) if i != axis]
        k_per = isqrt_array(perp[0] * perp[0] + perp[1] * perp[1])[indep]
        k = np.sqrt(k2.astype(np.float64))
        kidx = isqrt_array(k2)
        with np.errstate(invalid="ignore", divide="ignore"):
            mu = np.where(k == 0, 0.0, kpar_s / k)
        mu2 = mu * mu
        L2 = (3.0 * mu2 - 1.0) / 2.0
        L4 = (35.0 * mu2 * mu2 - 30.0 * mu2 + 3.0) / 8.0
        k_par = np.abs(kpar_s)
        m1 = k2 <= middle * middle
        i2 = (kmax_par + 1) * k_per + k_par
        n1, n2, n3 = kmax_par + 1, (kmax_par + 1) * (kmax_per + 1), kmax + 1
        dks = []
        for f in range(fields):
            p = MAS_function(MAS[f])
            corr = (MAS_correction(prefact * kx, p) * MAS_correction(prefact * ky, p) * MAS_correction(prefact * kz, p))
            dks.append((FFT3Dr(delta[f]) * np.broadcast_to(corr, shape))[indep])
        pairs = [(i, i) for i in range(fields)] + [(i, j) for i in range(fields) for j in range(i + 1, fields)]
        sums = []
        for (a, b) in pairs:
            v = (dks[a] * np.conj(dks[b])).real
            s3 = np.stack([np.bincount(kidx, weights=v * w, minlength=n3) for w in (1.0, L2, L4)], axis=1)
            s1 = np.bincount(k_par[m1], weights=v[m1], minlength=n1)
            s2 = np.bincount(i2, weights=v, minlength=n2)
            sums.append((s3, s1, s2))
        N1 = np.bincount(k_par[m1], minlength=n1).astype(np.float64)
        N2 = np.bincount(i2, minlength=n2).astype(np.float64)
        N3 = np.bincount(kidx, minlength=n3).astype(np.float64)
        k1 = np.bincount(k_par[m1], weights=k_par[m1].astype(np.float64), minlength=n1)
        k3 = np.bincount(kidx, weights=k, minlength=n3)
        fact = (BoxSize / dims ** 2) ** 3
        # 1-D (Pk_library.pyx:734-748)
        self.k1D = _div(k1[1:], N1[1:]) * kF
        kmaxper = np.sqrt(kN ** 2 - self.k1D ** 2)
        area = _div(np.pi * kmaxper ** 2, N1[1:]) / (2.0 * np.pi) ** 2
        P1 = np.stack([s[1][1:] * fact * area for s in sums], axis=1)
        self.Nmodes1D = N1[1:]
        self.Pk1D, self.PkX1D = P1[:, :fields], P1[:, fields:]
        # 2-D (Pk_library.pyx:751-766)
        kp = np.arange(kmax_par + 1)
        kq = np.arange(kmax_per + 1)
        self.kpar = np.tile(0.5 * (2 * kp + 1) * kF, kmax_per + 1)
        self.kper = np.repeat(0.5 * (2 * kq + 1) * kF, kmax_par + 1)
        P2 = np.stack([_div(s[2] * fact, N2) for s in sums], axis=1)
        self.Nmodes2D = N2
        self.Pk2D, self.PkX2D = P2[:, :fields], P2[:, fields:]
        # 3-D (Pk_library.pyx:768-785)
        check_number_modes(N3, dims)
        self.k3D = _div(k3[1:], N3[1:]) * kF
        P3 = np.stack([_div(s[0][1:] * np.array([1.0, 5.0, 9.0]), N3[1:, None]) * fact for s in sums], axis=2)
        self.Nmodes3D = N3[1:]
        self.Pk, self.XPk = P3[:, :, :fields], P3[:, :, fields:]
