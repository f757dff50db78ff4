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


def f32_floor_pk(delta, BoxSize, axis, MAS):
    """The float32 floor: the exact overdensity rounded once to float32, transformed by a float32 FFT
    (pocketfft in single precision, standing in for the reference's FFTW single-precision plan,
    Pk_library.pyx:120-133), binned in float64.  Window-suppressed corner modes of smooth fields are
    ill-conditioned w.r.t. float32 rounding; there no float32 pipeline can reach 1e-5."""
    import scipy.fft
    d32 = np.asarray(delta, dtype=np.float64).astype(np.float32)
    N = d32.shape[0]
    dk = scipy.fft.rfftn(d32).astype(np.complex128)
    return P.Pk(None, BoxSize, sums=C.pk_sums_c(dk, N, axis, MAS), dims=N)


def oracle_pk_particles(pos, BoxSize, N, MAS, axis, W=None, with_floor=False):
    """FP64 oracle of the fused path; with with_floor also ``f32_floor_pk`` of the exact grid."""
    rho = C.deposit_f64(pos, N, BoxSize, MAS, W)
    mass = float(len(pos)) if W is None else float(np.sum(np.asarray(W, dtype=np.float64)))
    rhobar = mass / N ** 3
    d = rho / rhobar - 1.0
    o = oracle_pk(d, BoxSize, axis, MAS)
    if not with_floor:
        return o
    return o, f32_floor_pk(d, BoxSize, axis, MAS)


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


def assert_pk_close(g, o, rtol, dc2d_abs=None, floor=None):
    """Per-bin parity: Nmodes bit-exact, P0 and Pk1D relative, P2/P4 relative to P0 of the bin.

    With `floor` (``f32_floor_pk``: the exact grid rounded to float32 and transformed by a float32
    FFT), a bin may exceed rtol only where a float32 pipeline does: the allowance is 4x the local
    envelope of the floor's error (two float32 pipelines make independent rounding errors, so a
    single-mode bin can differ by a few times between them)."""
    for name in ("Nmodes3D", "Nmodes1D", "Nmodes2D"):
        assert np.array_equal(getattr(g, name), getattr(o, name)), name
    assert np.allclose(g.k3D, o.k3D, rtol=1e-12, atol=0)
    assert np.allclose(g.k1D, o.k1D, rtol=1e-13, atol=0)
    assert np.allclose(g.kpar, o.kpar, rtol=1e-13) and np.allclose(g.kper, o.kper, rtol=1e-13)
    P0 = o.Pk[:, 0]
    n2shape = (len(o.kpar) // (len(o.k1D) + 1), len(o.k1D) + 1)

    def allowed(err_floor, shape=None):
        if floor is None:
            return rtol
        return np.maximum(rtol, 4.0 * _envelope(err_floor, shape))

    e0 = np.abs(g.Pk[:, 0] - P0) / np.abs(P0)
    a0 = allowed(np.abs(floor.Pk[:, 0] - P0) / np.abs(P0) if floor is not None else None)
    assert np.all(e0 <= a0), ("P0", e0.max(), int(np.argmax(e0 - a0)))
    for l in (1, 2):
        e = np.abs(g.Pk[:, l] - o.Pk[:, l]) / np.abs(P0)
        a = allowed(np.abs(floor.Pk[:, l] - o.Pk[:, l]) / np.abs(P0) if floor is not None else None)
        assert np.all(e <= a), ("P%d" % (2 * l), e.max(), int(np.argmax(e - a)))
    ep = np.abs(g.Pkphase - o.Pkphase) / np.abs(o.Pkphase)
    ap = allowed(np.abs(floor.Pkphase - o.Pkphase) / np.abs(o.Pkphase) if floor is not None else None)
    assert np.all(ep <= ap), ("Pkphase", ep.max())
    # the last 1-D bin is exactly 0 in the reference (kmaxper = sqrt(kN^2 - k^2) = 0)
    den1 = np.where(o.Pk1D == 0, 1, np.abs(o.Pk1D))
    e1 = np.where(o.Pk1D == 0, np.abs(g.Pk1D), np.abs(g.Pk1D - o.Pk1D) / den1)
    a1 = allowed(np.abs(floor.Pk1D - o.Pk1D) / den1 if floor is not None else None)
    assert np.all(e1 <= a1), ("Pk1D", e1.max())
    e2 = np.abs(g.Pk2D - o.Pk2D)
    scale = np.where(o.Pk2D == 0, 1, np.abs(o.Pk2D))
    e2r = e2 / scale
    a2 = allowed(np.abs(floor.Pk2D - o.Pk2D) / scale if floor is not None else None, n2shape)
    # bin (0,0) holds only the DC mode: its power is rounding noise of the zero mean
    ok = e2r <= a2
    ok[0] = e2[0] <= (dc2d_abs if dc2d_abs is not None else rtol * np.abs(o.Pk2D).max())
    assert np.all(ok), ("Pk2D", e2r[1:].max(), int(np.argmax((e2r - a2)[1:])) + 1)
    out = {"P0": float(e0.max()), "P2/P0": float((np.abs(g.Pk[:, 1] - o.Pk[:, 1]) / np.abs(P0)).max()),
           "P4/P0": float((np.abs(g.Pk[:, 2] - o.Pk[:, 2]) / np.abs(P0)).max()),
           "Pk1D": float(e1.max()), "Pk2D": float(e2r[1:].max()), "phase": float(ep.max())}
    if floor is not None:
        out["bins_over_1e-5"] = int(np.sum(e0 > rtol) + np.sum(np.abs(g.Pk[:, 1:] - o.Pk[:, 1:]) / np.abs(P0)[:, None] > rtol))
        out["f32_floor_P4/P0"] = float((np.abs(floor.Pk[:, 2] - o.Pk[:, 2]) / np.abs(P0)).max())
    return out
