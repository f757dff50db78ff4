"""Drop-in replacement of Pylians' ``redshift_space_library`` on MI355X.

``pos_redshift_space(pos, vel, BoxSize, Hubble, redshift, axis)`` (redshift_space_library.pyx:29-43)
moves particles to redshift space in place: ``pos[:, axis] += vel[:, axis] * (1 + z)/H`` in float32,
then ``(pos + BoxSize) % BoxSize`` where a position leaves [0, BoxSize].  ``pos`` and ``vel`` are
C-contiguous float32 (n, cols) arrays, as the reference's typed memoryviews require; numpy arrays
are shifted on the GPU and written back, device tensors are shifted where they lie.
"""
import numpy as np

from . import _lib

__all__ = ["pos_redshift_space"]


def _check(a, name):
    if _lib.is_dev(a):
        if _lib.dtype_of(a) != np.float32 or not a.is_contiguous() or a.ndim != 2:
            raise ValueError("%s must be a contiguous (n, cols) float32 device tensor" % name)
    elif not isinstance(a, np.ndarray) or a.dtype != np.float32:
        raise ValueError("Buffer dtype mismatch, expected 'float' but got '%s'" % getattr(a, "dtype", type(a)))
    elif a.ndim != 2 or not a.flags.c_contiguous:
        raise ValueError("ndarray is not C-contiguous (n, cols)")


def pos_redshift_space(pos, vel, BoxSize, Hubble, redshift, axis):
    """Returns None; shifts `pos` in place (redshift_space_library.pyx:29-43)."""
    _check(pos, "pos")
    _check(vel, "vel")
    if tuple(pos.shape) != tuple(vel.shape):
        raise ValueError("pos and vel must have the same shape")
    n, cols = int(pos.shape[0]), int(pos.shape[1])
    if not 0 <= axis < cols:
        raise IndexError("axis %d out of range for %d columns" % (axis, cols))
    p = pos if _lib.is_dev(pos) else _lib.to_dev(pos, _lib.F32T)
    v = _lib.to_dev(vel, _lib.F32T)
    _lib.check(_lib.lib().pyl_pos_redshift_space(_lib.ptr(p), _lib.ptr(v), n, cols, float(BoxSize), float(Hubble),
                                                 float(redshift), int(axis), _lib.stream()))
    if not _lib.is_dev(pos):
        pos[...] = _lib.to_host(p)
