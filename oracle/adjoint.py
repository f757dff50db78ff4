"""FP64 gradient of a weighted power-spectrum loss w.r.t. particle positions
(TEST INFRASTRUCTURE ONLY).  Builder-defined: the reference has no adjoint.

Forward model (float64):  rho = CIC(pos) (MAS_library.pyx:123-166),
delta = rho / rhobar - 1 with rhobar = N_p / N^3 (Pk_snapshot.py:84,88),
delta_k = r2c(delta), P0_b from the Pk mode loop (Pk_library.pyx:314-381, 410-417),
loss = sum_b w_b P0_b over the first len(w) bins (bin b = floor|k| = b+1).

Gradient: with c_k the window factor and g_k = w_b (L/N^2)^3 / Nmodes_b c_k^2,
  dL/d delta(x) = c2r_unnormalised(H)(x),  H_k = g_k delta_k on the half spectrum,
  doubled on the self-conjugate modes (each counted once in the full-space sum but
  twice in d|delta_k|^2);
  dL/d rho = dL/d delta / rhobar;
  dL/d pos_{p,a} = (N/L) * sum_corners G(cell) * (+1 | -1) * prod_{b != a} t_b
  (the derivative of the CIC weights u_a = dist_a - floor, d_a = 1 - u_a).

Pinned by (tests/test_oracle_adjoint.py): central finite differences of ``loss``
(float64) on particles away from cell planes.
"""
import numpy as np

from . import mas as _mas
from . import pk as _pk


def loss(pos, dims, BoxSize, w):
    rho = _mas.deposit(pos, dims, BoxSize, "CIC")
    rhobar = pos.shape[0] / dims ** 3
    P = _pk.Pk(rho / rhobar - 1.0, BoxSize, axis=0, MAS="CIC")
    return float(np.dot(w, P.Pk[: len(w), 0]))


def grad_spectrum(delta, dims, BoxSize, w):
    """Half spectrum H whose unnormalised c2r is dL/d delta."""
    dk = _pk.FFT3Dr(delta)
    kx, ky, kz, indep = _pk.mode_grid(dims)
    shape = dk.shape
    k2 = np.broadcast_to(kx * kx + ky * ky + kz * kz, shape)
    kidx = _pk.isqrt_array(k2)
    nb = len(w)
    nm = np.bincount(kidx[np.broadcast_to(indep, shape)], minlength=nb + 1)[: nb + 1]
    unit = (BoxSize / dims ** 2) ** 3
    g_bin = np.zeros(kidx.max() + 1)
    g_bin[1:nb + 1] = np.asarray(w, dtype=np.float64) * unit / nm[1:nb + 1]
    prefact = np.pi / dims
    c = (_pk.MAS_correction(prefact * kx, 2) * _pk.MAS_correction(prefact * ky, 2)
         * _pk.MAS_correction(prefact * kz, 2))
    g = g_bin[kidx] * np.broadcast_to(c * c, shape)
    m = dims // 2
    if dims % 2 == 0:
        sc = np.broadcast_to(((kx == 0) | (kx == m)) & ((ky == 0) | (ky == m)) & ((kz == 0) | (kz == m)),
                             shape)
        g = np.where(sc, 2.0 * g, g)
    return g * dk


def grad(pos, dims, BoxSize, w):
    """dL/dpos, float64 (N_p, 3)."""
    pos64 = np.asarray(pos, dtype=np.float64)
    n = pos64.shape[0]
    rho = _mas.deposit(pos, dims, BoxSize, "CIC")
    rhobar = n / dims ** 3
    H = grad_spectrum(rho / rhobar - 1.0, dims, BoxSize, w)
    G = _pk.IFFT3Dr(H, dims) * dims ** 3 / rhobar        # dL/d rho
    inv = dims / float(BoxSize)
    idx, t, s = [], [], []
    for a in range(3):
        dist = pos64[:, a] * inv
        base = np.floor(dist)
        u = dist - base
        i0 = base.astype(np.int64) % dims
        idx.append(np.stack([i0, (i0 + 1) % dims], 1))
        t.append(np.stack([1.0 - u, u], 1))
        s.append(np.array([-1.0, 1.0]))
    out = np.zeros((n, 3))
    for l in range(2):
        for m_ in range(2):
            for q in range(2):
                val = G[idx[0][:, l], idx[1][:, m_], idx[2][:, q]]
                out[:, 0] += val * s[0][l] * t[1][:, m_] * t[2][:, q]
                out[:, 1] += val * t[0][:, l] * s[1][m_] * t[2][:, q]
                out[:, 2] += val * t[0][:, l] * t[1][:, m_] * s[2][q]
    return out * inv
