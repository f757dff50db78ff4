"""FP64 oracle of Pylians' FFT-shell bispectrum (TEST INFRASTRUCTURE ONLY).

Restates ``class Bk`` of bispectrum_library.pyx:32-196 in float64:

* shells [k_i/kF - 1, k_i/kF + 1) in kF units for k1, k2 and
  k3(theta) = sqrt((k2 sin theta)^2 + (k2 cos theta + k1)^2)   (:59-77);
* window deconvolution of EVERY half-spectrum mode -- the independence cut is
  commented out in the reference (:100-109);
* masked spectra delta_i(k) and I_i(k) = 1 on the shell, each inverse-FFT'd with
  1/N^3 normalisation (:121-129, 142-150, 166-171);
* P_i = sum d_i^2 / sum I_i^2 * (L/N^2)^3   (:133-138, 154-159, 175-180);
* B = sum d1 d2 d3 / sum I1 I2 I3 * (L^2/N^3)^3, Q = B/(P1P2+P1P3+P2P3) (:183-190).

Pinned by (tests/test_oracle_bk.py): brute-force enumeration of closed triangles
k1+k2+k3 = 0 (mod N) over the full-space shells on tiny grids, which equals the
real-space triple sum by the convolution theorem, and Q's closed form.
"""
import numpy as np

from . import pk as _pk


def shell_bounds(BoxSize, dims, k1, k2, theta):
    """k_all, k_min, k_max (kF units) for the bins+2 shells (bispectrum_library.pyx:59-77)."""
    kF = 2.0 * np.pi / BoxSize
    theta = np.asarray(theta, dtype=np.float64)
    k3 = np.sqrt((k2 * np.sin(theta)) ** 2 + (k2 * np.cos(theta) + k1) ** 2)
    k_all = np.concatenate([[k1, k2], k3])
    return k_all, (k_all - kF) / kF, (k_all + kF) / kF


def _div(a, b):
    if b == 0:
        raise ZeroDivisionError("float division")
    return a / b


class Bk:
    """float64 restatement of ``Pk_library.Bk`` (bispectrum_library.pyx:32-196).

    Attributes: B (bins,), Q (bins,), k (bins+2,), Pk (bins+2,).
    """

    def __init__(self, delta, BoxSize, k1, k2, theta, MAS="CIC", threads=1):
        delta = np.asarray(delta)
        dims = delta.shape[0]
        BoxSize = float(BoxSize)
        bins = theta.shape[0]
        k_all, k_min, k_max = shell_bounds(BoxSize, dims, k1, k2, theta)
        p = _pk.MAS_function(MAS)
        dk = _pk.FFT3Dr(delta)
        kx, ky, kz, _ = _pk.mode_grid(dims)
        prefact = np.pi / dims
        corr = (_pk.MAS_correction(prefact * kx, p) * _pk.MAS_correction(prefact * ky, p)
                * _pk.MAS_correction(prefact * kz, p))
        dk = dk * corr
        k = np.sqrt((kx * kx + ky * ky + kz * kz).astype(np.float64))
        k = np.broadcast_to(k, dk.shape)

        def fields(i):
            m = (k >= k_min[i]) & (k < k_max[i])
            d = _pk.IFFT3Dr(np.where(m, dk, 0.0), dims)
            I = _pk.IFFT3Dr(m.astype(np.complex128), dims)
            return d, I

        unitP = (BoxSize / dims ** 2) ** 3
        unitB = (BoxSize ** 2 / dims ** 3) ** 3
        Pk = np.zeros(bins + 2)
        d1, I1 = fields(0)
        Pk[0] = _div(np.sum(d1 * d1), np.sum(I1 * I1)) * unitP
        d2, I2 = fields(1)
        Pk[1] = _div(np.sum(d2 * d2), np.sum(I2 * I2)) * unitP
        d12, I12 = d1 * d2, I1 * I2
        B = np.zeros(bins)
        Q = np.zeros(bins)
        for j in range(bins):
            d3, I3 = fields(j + 2)
            Pk[j + 2] = _div(np.sum(d3 * d3), np.sum(I3 * I3)) * unitP
            B[j] = _div(np.sum(d12 * d3), np.sum(I12 * I3)) * unitB
            Q[j] = B[j] / (Pk[0] * Pk[1] + Pk[0] * Pk[j + 2] + Pk[1] * Pk[j + 2])
        self.B, self.Q, self.k, self.Pk = B, Q, k_all, Pk
