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


def oracle_pk_particles(pos, BoxSize, N, MAS, axis, W=None, with_floor=False):
    """FP64 oracle of the fused path.  With with_floor, also the FP64 evaluation on the float32
    rounding of the exact overdensity grid: the error floor of ANY float32 representation of delta
    (the window-suppressed corner modes of clustered fields are ill-conditioned w.r.t. it)."""
    rho = C.deposit_f64(pos, N, BoxSize, MAS, W)
    mass = float(len(pos)) if W is None else float(np.sum(np.asarray(W, dtype=np.float64)))
    rhobar = mass / N ** 3
    d = rho / rhobar - 1.0
    o = oracle_pk(d, BoxSize, axis, MAS)
    if not with_floor:
        return o
    return o, oracle_pk(d.astype(np.float32), BoxSize, axis, MAS)


def assert_pk_close(g, o, rtol, dc2d_abs=None, floor=None):
    """Per-bin parity: Nmodes bit-exact, P0 and Pk1D relative, P2/P4 relative to P0 of the bin.
    With `floor` (oracle on the float32-rounded exact grid), a bin may exceed rtol only where the
    float32 representation itself does, and then by at most 2x that floor's error."""
    for name in ("Nmodes3D", "Nmodes1D", "Nmodes2D"):
        assert np.array_equal(getattr(g, name), getattr(o, name)), name
    assert np.allclose(g.k3D, o.k3D, rtol=1e-12, atol=0)
    assert np.allclose(g.k1D, o.k1D, rtol=1e-13, atol=0)
    assert np.allclose(g.kpar, o.kpar, rtol=1e-13) and np.allclose(g.kper, o.kper, rtol=1e-13)
    P0 = o.Pk[:, 0]

    def allowed(name, err_floor):
        return np.maximum(rtol, 2.0 * err_floor) if floor is not None else rtol

    fl = (lambda a, b, den: np.abs(a - b) / den) if floor is not None else None
    e0 = np.abs(g.Pk[:, 0] - P0) / np.abs(P0)
    a0 = allowed("P0", fl(floor.Pk[:, 0], P0, np.abs(P0)) if floor is not None else 0)
    assert np.all(e0 <= a0), ("P0", e0.max(), int(np.argmax(e0 - a0)))
    for l in (1, 2):
        e = np.abs(g.Pk[:, l] - o.Pk[:, l]) / np.abs(P0)
        a = allowed("P", fl(floor.Pk[:, l], o.Pk[:, l], np.abs(P0)) if floor is not None else 0)
        assert np.all(e <= a), ("P%d" % (2 * l), e.max(), int(np.argmax(e - a)))
    ep = np.abs(g.Pkphase - o.Pkphase) / np.abs(o.Pkphase)
    ap = allowed("ph", fl(floor.Pkphase, o.Pkphase, np.abs(o.Pkphase)) if floor is not None else 0)
    assert np.all(ep <= ap), ("Pkphase", ep.max())
    # the last 1-D bin is exactly 0 in the reference (kmaxper = sqrt(kN^2 - k^2) = 0)
    den1 = np.where(o.Pk1D == 0, 1, np.abs(o.Pk1D))
    e1 = np.where(o.Pk1D == 0, np.abs(g.Pk1D), np.abs(g.Pk1D - o.Pk1D) / den1)
    a1 = allowed("1D", fl(floor.Pk1D, o.Pk1D, den1) if floor is not None else 0)
    assert np.all(e1 <= a1), ("Pk1D", e1.max())
    e2 = np.abs(g.Pk2D - o.Pk2D)
    scale = np.abs(o.Pk2D)
    e2r = e2[1:] / scale[1:]
    a2 = allowed("2D", fl(floor.Pk2D[1:], o.Pk2D[1:], scale[1:]) if floor is not None else 0)
    assert np.all(e2r <= a2), ("Pk2D", e2r.max(), int(np.argmax(e2r - a2)) + 1)
    # bin (0,0) holds only the DC mode: its power is rounding noise of the zero mean
    assert e2[0] <= (dc2d_abs if dc2d_abs is not None else rtol * np.abs(o.Pk2D).max())
    out = {"P0": float(e0.max()), "P2/P0": float((np.abs(g.Pk[:, 1] - o.Pk[:, 1]) / np.abs(P0)).max()),
           "P4/P0": float((np.abs(g.Pk[:, 2] - o.Pk[:, 2]) / np.abs(P0)).max()),
           "Pk1D": float(e1.max()), "Pk2D": float(e2r.max()), "phase": float(ep.max())}
    if floor is not None:
        out["bins_over_1e-5"] = int(np.sum(e0 > rtol) + np.sum(np.abs(g.Pk[:, 1:] - o.Pk[:, 1:]) / np.abs(P0)[:, None] > rtol))
        out["f32_floor_P4/P0"] = float((np.abs(floor.Pk[:, 2] - o.Pk[:, 2]) / np.abs(P0)).max())
    return out
