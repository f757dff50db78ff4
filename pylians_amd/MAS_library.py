"""Drop-in replacement of Pylians' ``MAS_library`` (mass assignment) on MI355X.

Same call surface as the reference (MAS_library.pyx):

* ``MA(pos, number, BoxSize, MAS='CIC', W=None, verbose=False, renormalize_2D=True)``
  (MAS_library.pyx:57-112): accumulates into ``number`` in place, 2-D or 3-D;
* ``NGP/CIC/TSC/PCS(pos, number, BoxSize)`` and ``NGPW/CICW/TSCW/PCSW(pos, number, BoxSize, W)``
  (MAS_library.pyx:123-545), ``CICW_d`` / ``NGPW_d`` with a float64 grid (:229-266, :338-358);
* the OpenMP-C wrappers ``XXXc3D(pos, number, BoxSize, threads)``,
  ``XXXWc3D(pos, number, W, BoxSize, threads)`` and their 2-D versions (:1136-1221);
* ``CIC_interp(density, BoxSize, pos, den)`` (:558-599);
* ``FLOAT_type()`` (:11-13).

``pos``/``number``/``W`` may be numpy arrays (copied to and from the GPU) or torch tensors on
the GPU (used in place, zero-copy).  Errors follow the reference: a pos/grid rank mismatch or an
unknown scheme prints and raises SystemExit (MAS_library.pyx:64-66, 81-82).
"""
import sys

import numpy as np

from . import _lib

__all__ = ["MA", "NGP", "CIC", "TSC", "PCS", "NGPW", "CICW", "TSCW", "PCSW", "CICW_d", "NGPW_d",
           "NGPc3D", "NGPWc3D", "NGPc2D", "NGPWc2D", "CICc3D", "CICWc3D", "CICc2D", "CICWc2D",
           "TSCc3D", "TSCWc3D", "TSCc2D", "TSCWc2D", "PCSc3D", "PCSWc3D", "PCSc2D", "PCSWc2D",
           "CIC_interp", "FLOAT_type"]


def FLOAT_type():
    """dtype of the C-path grids (MAS_c.h:1 typedef float FLOAT)."""
    return np.float32


def _deposit(pos, number, BoxSize, scheme, W, flags, grid_f64=False, grid_view=None):
    """Run pyl_mas_deposit accumulating into `number` in place (numpy or device tensor)."""
    want = np.float64 if grid_f64 else np.float32
    axes = int(pos.shape[1])
    if _lib.is_dev(number):
        if _lib.dtype_of(number) != want or not number.is_contiguous():
            raise ValueError("number must be a contiguous %s device tensor" % np.dtype(want).name)
        g = number
    else:
        if not isinstance(number, np.ndarray) or number.dtype != want:
            raise ValueError("Buffer dtype mismatch, expected '%s' but got '%s'"
                             % (np.dtype(want).name, getattr(number, "dtype", type(number))))
        g = _lib.to_dev(number, want)
    p = _lib.to_dev(pos, _lib.F32T)
    w = None if W is None else _lib.to_dev(W, _lib.F32T)
    n = int(p.shape[0])
    if w is not None and int(np.prod(w.shape)) != n:
        raise ValueError("W must have one weight per particle")
    dims = int(g.shape[0])
    f = flags | (_lib.GRID_F64 if grid_f64 else 0)
    _lib.check(_lib.lib().pyl_mas_deposit(_lib.ptr(p), _lib.ptr(w), n, axes, dims, float(BoxSize),
                                          scheme, f, _lib.ptr(g), _lib.stream()))
    if not _lib.is_dev(number):
        number[...] = _lib.to_host(g).reshape(number.shape)


def MA(pos, number, BoxSize, MAS="CIC", W=None, verbose=False, renormalize_2D=True):
    """Mass assignment (MAS_library.pyx:57-112); returns None, accumulates into `number`."""
    coord, coord_aux = pos.shape[1], number.ndim
    if coord != coord_aux:
        print("pos have %d dimensions and the density %d!!!" % (coord, coord_aux))
        sys.exit()
    if MAS not in _lib.SCHEMES:
        print("option not valid!!!")
        sys.exit()
    if verbose:
        print("\nUsing %s mass assignment scheme%s" % (MAS, "" if W is None else " with weights"))
    flags = 0
    if coord == 2:
        flags = _lib.DUP2D | (_lib.RENORM2D if renormalize_2D else 0)
    _deposit(pos, number, BoxSize, _lib.SCHEMES[MAS], W, flags)


def _cython_stencil(pos, number, BoxSize, MAS, W=None, grid_f64=False):
    """Per-scheme Cython functions: 3-D grid (a 2-D pos writes the z=0 plane of an (N,N,1) grid
    STENCIL times, MAS_library.pyx:137-139, without renormalisation)."""
    if pos.shape[1] == 3:
        _deposit(pos, number, BoxSize, _lib.SCHEMES[MAS], W, 0, grid_f64)
        return
    if _lib.is_dev(number):
        plane = number[:, :, 0].contiguous()
        _deposit(pos, plane, BoxSize, _lib.SCHEMES[MAS], W, _lib.DUP2D, grid_f64)
        number[:, :, 0] = plane
    else:
        plane = np.ascontiguousarray(number[:, :, 0])
        _deposit(pos, plane, BoxSize, _lib.SCHEMES[MAS], W, _lib.DUP2D, grid_f64)
        number[:, :, 0] = plane


def NGP(pos, number, BoxSize):
    _cython_stencil(pos, number, BoxSize, "NGP")


def CIC(pos, number, BoxSize):
    _cython_stencil(pos, number, BoxSize, "CIC")


def TSC(pos, number, BoxSize):
    _cython_stencil(pos, number, BoxSize, "TSC")


def PCS(pos, number, BoxSize):
    _cython_stencil(pos, number, BoxSize, "PCS")


def NGPW(pos, number, BoxSize, W):
    _cython_stencil(pos, number, BoxSize, "NGP", W)


def CICW(pos, number, BoxSize, W):
    _cython_stencil(pos, number, BoxSize, "CIC", W)


def TSCW(pos, number, BoxSize, W):
    _cython_stencil(pos, number, BoxSize, "TSC", W)


def PCSW(pos, number, BoxSize, W):
    _cython_stencil(pos, number, BoxSize, "PCS", W)


def CICW_d(pos, number, BoxSize, W):
    """CICW with a float64 grid (MAS_library.pyx:229-266)."""
    _cython_stencil(pos, number, BoxSize, "CIC", W, grid_f64=True)


def NGPW_d(pos, number, BoxSize, W):
    """NGPW with a float64 grid (MAS_library.pyx:338-358)."""
    _cython_stencil(pos, number, BoxSize, "NGP", W, grid_f64=True)


# OpenMP C path (MAS_c.c; 2-D deposits each weight once, no renormalisation)
def _c_path(MAS):
    def f3(pos, number, BoxSize, threads):
        _deposit(pos, number, BoxSize, _lib.SCHEMES[MAS], None, 0)

    def fw3(pos, number, W, BoxSize, threads):
        _deposit(pos, number, BoxSize, _lib.SCHEMES[MAS], W, 0)

    return f3, fw3


NGPc3D, NGPWc3D = _c_path("NGP")
NGPc2D, NGPWc2D = _c_path("NGP")
CICc3D, CICWc3D = _c_path("CIC")
CICc2D, CICWc2D = _c_path("CIC")
TSCc3D, TSCWc3D = _c_path("TSC")
TSCc2D, TSCWc2D = _c_path("TSC")
PCSc3D, PCSWc3D = _c_path("PCS")
PCSc2D, PCSWc2D = _c_path("PCS")


def CIC_interp(density, BoxSize, pos, den):
    """CIC interpolation of a 3-D float32 grid at particle positions into `den` (MAS_library.pyx:558-599)."""
    g = _lib.to_dev(density, _lib.F32T)
    p = _lib.to_dev(pos, _lib.F32T)
    n = int(p.shape[0])
    out = den if (_lib.is_dev(den) and _lib.dtype_of(den) == _lib.F32T and den.is_contiguous()) else _lib.empty((n,), _lib.F32T)
    _lib.check(_lib.lib().pyl_interp(_lib.ptr(g), int(g.shape[0]), float(BoxSize), _lib.ptr(p), n, 2,
                                     _lib.ptr(out), _lib.stream()))
    if out is not den:
        if _lib.is_dev(den):
            den.copy_(out)
        else:
            den[...] = _lib.to_host(out)
