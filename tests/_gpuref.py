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


def oracle_pk_particles(pos, BoxSize, N, MAS, axis, W=None):
    rho = C.deposit_f64(pos, N, BoxSize, MAS, W)
    mass = float(len(pos)) if W is None else float(np.sum(np.asarray(W, dtype=np.float64)))
    rhobar = mass / N ** 3
    return oracle_pk(rho / rhobar - 1.0, BoxSize, axis, MAS)


def assert_pk_close(g, o, rtol, dc2d_abs=None):
    """Per-bin parity: Nmodes bit-exact, P0 and Pk1D relative, P2/P4 relative to P0 of the bin."""
    for name in ("Nmodes3D", "Nmodes1D", "Nmodes2D"):
        assert np.array_equal(getattr(g, name), getattr(o, name)), name
    assert np.allclose(g.k3D, o.k3D, rtol=1e-12, atol=0)
    assert np.allclose(g.k1D, o.k1D, rtol=1e-13, atol=0)
    assert np.allclose(g.kpar, o.kpar, rtol=1e-13) and np.allclose(g.kper, o.kper, rtol=1e-13)
    P0 = o.Pk[:, 0]
    e0 = np.abs(g.Pk[:, 0] - P0) / np.abs(P0)
    assert e0.max() <= rtol, ("P0", e0.max(), int(e0.argmax()))
    for l in (1, 2):
        e = np.abs(g.Pk[:, l] - o.Pk[:, l]) / np.abs(P0)
        assert e.max() <= rtol, ("P%d" % (2 * l), e.max(), int(e.argmax()))
    ep = np.abs(g.Pkphase - o.Pkphase) / np.abs(o.Pkphase)
    assert ep.max() <= rtol, ("Pkphase", ep.max())
    # the last 1-D bin is exactly 0 in the reference (kmaxper = sqrt(kN^2 - k^2) = 0)
    e1 = np.where(o.Pk1D == 0, np.abs(g.Pk1D), np.abs(g.Pk1D - o.Pk1D) / np.where(o.Pk1D == 0, 1, np.abs(o.Pk1D)))
    assert e1.max() <= rtol, ("Pk1D", e1.max())
    e2 = np.abs(g.Pk2D - o.Pk2D)
    scale = np.abs(o.Pk2D)
    e2r = e2[1:] / scale[1:]
    assert e2r.max() <= rtol, ("Pk2D", e2r.max(), int(e2r.argmax()) + 1)
    # bin (0,0) holds only the DC mode: its power is rounding noise of the zero mean
    assert e2[0] <= (dc2d_abs if dc2d_abs is not None else rtol * np.abs(o.Pk2D).max())
    return {"P0": float(e0.max()), "Pk1D": float(e1.max()), "Pk2D": float(e2r.max()), "phase": float(ep.max())}
