"""FP64 oracle of Pylians mass assignment (TEST INFRASTRUCTURE ONLY).

Restates, in float64 arithmetic:

* the Cython stencils MAS_library.pyx:123-166 (CIC), 179-216 (CICW), 229-266 (CICW_d),
  273-292 (NGP), 305-325 (NGPW), 338-358 (NGPW_d), 369-404 (TSC), 417-452 (TSCW),
  463-497 (PCS), 510-545 (PCSW);
* the dispatcher ``MA`` MAS_library.pyx:57-112, including the 2-D path that
  deposits each weight 1/2/3/4 times and then divides the whole grid by that
  multiplicity when ``renormalize_2D`` (MAS_library.pyx:84-110);
* the OpenMP C stencils MAS_c.c:8-266 (2-D without duplication, ``n_max=1``).

Differences from the reference, all deliberate and mirrored by the GPU kernels:

* arithmetic is float64 (the reference computes ``dist = pos*inv_cell_size`` in
  float32, MAS_library.pyx:127,135, which is ~3e-5 cells off at dist~512);
* the cell index uses ``floor`` instead of C truncation and wraps periodically.
  For positions in [0, BoxSize] (the reference's domain) this is identical;
  for negative positions the reference indexes out of bounds.

Pinned by (tests/test_oracle_mas.py): the reference's own mass-conservation
cases (Test/test_MAS.py:10-238, library/tests/test.py:8-33), closed-form single
particle stencils, the grid-point lattice known answer, a brute-force
all-cells kernel evaluation on tiny grids, and the reference's MAS_c.c compiled
from source (oracle/_ref) compared within float32 rounding.
"""
import numpy as np

SCHEMES = {"NGP": 1, "CIC": 2, "TSC": 3, "PCS": 4}
STENCIL = {"NGP": 1, "CIC": 2, "TSC": 3, "PCS": 4}


def axis_weights(dist, scheme, dims):
    """Per-axis cell indices and 1-D weights for one coordinate axis.

    dist: float64 array (n,) of positions in cell units (pos*dims/BoxSize).
    Returns (idx int64 (n,S), w float64 (n,S)).
    Follows MAS_library.pyx:151-157 (CIC), :290-291 (NGP), :393-399 (TSC),
    :486-492 (PCS).
    """
    dist = np.asarray(dist, dtype=np.float64)
    if scheme == "NGP":
        idx = np.floor(dist + 0.5).astype(np.int64) % dims
        return idx[:, None], np.ones((dist.size, 1))
    if scheme == "CIC":
        base = np.floor(dist)
        u = dist - base
        d = 1.0 - u
        i0 = base.astype(np.int64) % dims
        idx = np.stack([i0, (i0 + 1) % dims], axis=1)
        return idx, np.stack([d, u], axis=1)
    if scheme == "TSC":
        minimum = np.floor(dist - 1.5)
        cells = minimum[:, None] + np.arange(1, 4)[None, :]
        diff = np.abs(cells - dist[:, None])
        w = np.where(diff < 0.5, 0.75 - diff * diff,
                     np.where(diff < 1.5, 0.5 * (1.5 - diff) * (1.5 - diff), 0.0))
        return cells.astype(np.int64) % dims, w
    if scheme == "PCS":
        minimum = np.floor(dist - 2.0)
        cells = minimum[:, None] + np.arange(1, 5)[None, :]
        diff = np.abs(cells - dist[:, None])
        w = np.where(diff < 1.0, (4.0 - 6.0 * diff * diff + 3.0 * diff * diff * diff) / 6.0,
                     np.where(diff < 2.0, (2.0 - diff) * (2.0 - diff) * (2.0 - diff) / 6.0, 0.0))
        return cells.astype(np.int64) % dims, w
    raise ValueError("unknown MAS %r" % (scheme,))


def deposit(pos, dims, BoxSize, scheme, W=None, multiplicity=1, chunk=1 << 18):
    """Sum of stencil contributions of all particles, float64 grid.

    pos: (n, axes) array, axes in {2, 3}. Returns an array of shape (dims,)*axes.
    ``multiplicity`` scales every contribution (the 2-D Cython path deposits
    each weight 2/3/4 times onto the single z=0 plane, MAS_library.pyx:137-139).
    """
    pos = np.asarray(pos)
    n, axes = pos.shape
    inv = dims / float(BoxSize)
    out = np.zeros(dims ** axes, dtype=np.float64)
    S = STENCIL[scheme]
    for s in range(0, n, chunk):
        p = pos[s:s + chunk].astype(np.float64)
        m = p.shape[0]
        flat = np.zeros((m, 1), dtype=np.int64)
        wt = np.ones((m, 1), dtype=np.float64)
        for a in range(axes):
            idx, w = axis_weights(p[:, a] * inv, scheme, dims)
            flat = (flat[:, :, None] * dims + idx[:, None, :]).reshape(m, -1)
            wt = (wt[:, :, None] * w[:, None, :]).reshape(m, -1)
        if W is not None:
            wt = wt * np.asarray(W[s:s + chunk], dtype=np.float64)[:, None]
        out += np.bincount(flat.ravel(), weights=wt.ravel(), minlength=dims ** axes)
    out *= multiplicity
    assert S >= 1
    return out.reshape((dims,) * axes)


def MA(pos, number, BoxSize, MAS="CIC", W=None, renormalize_2D=True):
    """Return the float64 result of MAS_library.MA(pos, number, ...) (MAS_library.pyx:57-112).

    ``number`` is the grid before the call (MA accumulates into it in place and
    never zeroes it, MAS_library.pyx:159-166); the returned array is the grid
    after the call.  Raises SystemExit where the reference calls sys.exit().
    """
    pos = np.asarray(pos)
    number = np.asarray(number, dtype=np.float64)
    coord, coord_aux = pos.shape[1], number.ndim
    if coord != coord_aux:
        raise SystemExit("pos have %d dimensions and the density %d!!!" % (coord, coord_aux))
    if MAS not in SCHEMES:
        raise SystemExit("option not valid!!!")
    dims = number.shape[0]
    if coord == 3:
        return number + deposit(pos, dims, BoxSize, MAS, W)
    # 2-D: weights deposited STENCIL times on the z=0 plane, then the whole
    # grid divided by that multiplicity when renormalize_2D (NGP: no division).
    m = STENCIL[MAS]
    out = number + deposit(pos, dims, BoxSize, MAS, W, multiplicity=m)
    if renormalize_2D and m > 1:
        out = out / m
    return out


def MAS_c(pos, number, BoxSize, MAS="CIC", W=None):
    """float64 result of the OpenMP C path MAS_c.c NGP/CIC/TSC/PCS (2-D: n_max=1, no duplication)."""
    pos = np.asarray(pos)
    number = np.asarray(number, dtype=np.float64)
    return number + deposit(pos, number.shape[0], BoxSize, MAS, W)


def kernel_1d(s, scheme):
    """1-D assignment kernel as a function of the signed distance s (cells) between
    the particle and a grid point.  Used only by the brute-force checks: it is
    the B-spline form of the stencils, independent of the index arithmetic above."""
    a = np.abs(s)
    if scheme == "NGP":
        # grid point i owns [i-0.5, i+0.5): particle at dist picks floor(dist+0.5)
        return ((s >= -0.5) & (s < 0.5)).astype(np.float64)
    if scheme == "CIC":
        return np.clip(1.0 - a, 0.0, None)
    if scheme == "TSC":
        return np.where(a < 0.5, 0.75 - a * a, np.where(a < 1.5, 0.5 * (1.5 - a) ** 2, 0.0))
    if scheme == "PCS":
        return np.where(a < 1.0, (4.0 - 6.0 * a * a + 3.0 * a ** 3) / 6.0,
                        np.where(a < 2.0, (2.0 - a) ** 3 / 6.0, 0.0))
    raise ValueError(scheme)
