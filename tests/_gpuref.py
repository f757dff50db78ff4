"""Oracle helpers for the GPU parity tests (FP64 CPU restatement as the checker)."""
import numpy as np

from oracle import cpu as C
from oracle import pk as P


def oracle_pk(delta, BoxSize, axis, MAS, keep=False):
    """float64 oracle Pk of a real field (C mode loop for speed; numpy for keep_deltak)."""
    delta = np.asarray(delta, dtype=np.float64)
    N = delta.shape[0]
    if keep:
        return P.Pk(delta, BoxSize, axis=axis, MAS=MAS, keep_deltak=True)
    dk = np.fft.rfftn(delta)
    s = C.pk_sums_c(dk, N, axis, MAS)
    return P.Pk(None, BoxSize, sums=s, dims=N)


def oracle_pk_particles(pos, BoxSize, N, MAS, axis, W=None, with_noise=False):
    """FP64 oracle of the fused path; with with_noise also ``f32_noise_pk`` of the exact overdensity."""
    rho = C.deposit_f64(pos, N, BoxSize, MAS, W)
    mass = float(len(pos)) if W is None else float(np.sum(np.asarray(W, dtype=np.float64)))
    rhobar = mass / N ** 3
    d = rho / rhobar - 1.0
    o = oracle_pk(d, BoxSize, axis, MAS)
    if not with_noise:
        return o
    return o, C.f32_noise_pk(d, BoxSize, axis, MAS)


def f32_noise(delta, BoxSize, axis, MAS):
    return C.f32_noise_pk(delta, BoxSize, axis, MAS)


def _envelope(e, shape=None):
    """Local maximum over neighbouring bins (+-2 in 1-D, 3x3 in the 2-D (k_per, k_par) plane)."""
    e = np.asarray(e, dtype=np.float64)
    if shape is None:
        out = e.copy()
        for d in (1, 2):
            out[d:] = np.maximum(out[d:], e[:-d])
            out[:-d] = np.maximum(out[:-d], e[d:])
        return out
    a = e.reshape(shape)
    pad = np.pad(a, 1, mode="edge")
    out = np.max([pad[i:i + shape[0], j:j + shape[1]] for i in range(3) for j in range(3)], axis=0)
    return out.ravel()


# Allowance in standard deviations of the float32 representation noise of a bin.  The error of a
# bin is a sum of many independent roundings, so it is Gaussian; 6 sigma is exceeded with probability
# 2e-9 per bin, while a wrong term, index or window exceeds it in whole ranges of bins.
NOISE_Z = 6.0


def assert_pk_close(g, o, rtol, dc2d_abs=None, noise=None):
    """Per-bin parity: Nmodes bit-exact, P0 and Pk1D relative, P2/P4 relative to P0 of the bin.

    With `noise` (``f32_noise_pk`` of the exact field), a bin may exceed rtol only by what the float32
    representation of the field does to that bin: up to NOISE_Z standard deviations of it (local
    maximum over neighbouring bins).  Window-suppressed single-mode bins near the Nyquist corner of a
    smooth field are the ones where no float32 grid reaches 1e-5."""
    for name in ("Nmodes3D", "Nmodes1D", "Nmodes2D"):
        assert np.array_equal(getattr(g, name), getattr(o, name)), name
    assert np.allclose(g.k3D, o.k3D, rtol=1e-12, atol=0)
    assert np.allclose(g.k1D, o.k1D, rtol=1e-13, atol=0)
    assert np.allclose(g.kpar, o.kpar, rtol=1e-13) and np.allclose(g.kper, o.kper, rtol=1e-13)
    P0 = o.Pk[:, 0]
    n2shape = (len(o.kpar) // (len(o.k1D) + 1), len(o.k1D) + 1)

    def allowed(std_rel, shape=None):
        if noise is None:
            return rtol
        return np.maximum(rtol, NOISE_Z * _envelope(std_rel, shape))

    e0 = np.abs(g.Pk[:, 0] - P0) / np.abs(P0)
    a0 = allowed(noise.Pk[:, 0] / np.abs(P0) if noise is not None else None)
    assert np.all(e0 <= a0), ("P0", e0.max(), int(np.argmax(e0 - a0)))
    for l in (1, 2):
        e = np.abs(g.Pk[:, l] - o.Pk[:, l]) / np.abs(P0)
        a = allowed(noise.Pk[:, l] / np.abs(P0) if noise is not None else None)
        assert np.all(e <= a), ("P%d" % (2 * l), e.max(), int(np.argmax(e - a)))
    ep = np.abs(g.Pkphase - o.Pkphase) / np.abs(o.Pkphase)
    ap = allowed(noise.Pkphase / np.abs(o.Pkphase) if noise is not None else None)
    assert np.all(ep <= ap), ("Pkphase", ep.max())
    # the last 1-D bin is exactly 0 in the reference (kmaxper = sqrt(kN^2 - k^2) = 0)
    den1 = np.where(o.Pk1D == 0, 1, np.abs(o.Pk1D))
    e1 = np.where(o.Pk1D == 0, np.abs(g.Pk1D), np.abs(g.Pk1D - o.Pk1D) / den1)
    a1 = allowed(noise.Pk1D / den1 if noise is not None else None)
    assert np.all(e1 <= a1), ("Pk1D", e1.max())
    e2 = np.abs(g.Pk2D - o.Pk2D)
    scale = np.where(o.Pk2D == 0, 1, np.abs(o.Pk2D))
    e2r = e2 / scale
    a2 = allowed(noise.Pk2D / scale if noise is not None else None, n2shape)
    # bin (0,0) holds only the DC mode: its power is rounding noise of the zero mean
    ok = e2r <= a2
    ok[0] = e2[0] <= (dc2d_abs if dc2d_abs is not None else rtol * np.abs(o.Pk2D).max())
    assert np.all(ok), ("Pk2D", e2r[1:].max(), int(np.argmax((e2r - a2)[1:])) + 1)
    out = {"P0": float(e0.max()), "P2/P0": float((np.abs(g.Pk[:, 1] - o.Pk[:, 1]) / np.abs(P0)).max()),
           "P4/P0": float((np.abs(g.Pk[:, 2] - o.Pk[:, 2]) / np.abs(P0)).max()),
           "Pk1D": float(e1.max()), "Pk2D": float(e2r[1:].max()), "phase": float(ep.max())}
    if noise is not None:
        out["bins_over_1e-5"] = int(np.sum(e0 > rtol) + np.sum(np.abs(g.Pk[:, 1:] - o.Pk[:, 1:]) / np.abs(P0)[:, None] > rtol))
        out["max_sigmas_P"] = float(np.max(np.abs(g.Pk - o.Pk) / np.maximum(noise.Pk, 1e-300)))
        out["f32_sigma_rel"] = noise.sigma_rel
    return out
