"""GPU drop-in of smoothing_library (smoothing_library.pyx:20-114): filters on the grid and
field smoothing through the same r2c / c2r passes as Pk_library."""
import numpy as np

from . import _lib

_FILTERS = {"Top-Hat": 0, "Gaussian": 1}


__all__ = ["FT_filter", "field_smoothing"]


def FT_filter(BoxSize, R, dims, Filter, threads=1):
    """Fourier transform (complex64, (dims, dims, dims//2+1)) of the top-hat or Gaussian filter of
    radius R on the grid, normalised to unit sum (smoothing_library.pyx:20-86)."""
    if Filter not in _FILTERS:
        raise Exception("Filter %s not implemented!" % Filter)
    dims = int(dims)
    out = _lib.empty((dims, dims, dims // 2 + 1), _lib.C64T)
    _lib.check(_lib.lib().pyl_ft_filter(dims, float(BoxSize), float(R), _FILTERS[Filter], _lib.ptr(out),
                                        _lib.stream()))
    return _lib.to_host(out)


def field_smoothing(field, filter_k, threads=1):
    """field convolved with the filter whose transform is filter_k (smoothing_library.pyx:94-114),
    float32; numpy in, numpy out (device tensors stay on the device)."""
    if field.shape[0] != filter_k.shape[0]:
        raise Exception("field and filter have different grids!!!")
    f = _lib.to_dev(field, _lib.F32T)
    k = _lib.to_dev(filter_k, _lib.C64T)
    dims = int(f.shape[0])
    if f.ndim != 3 or tuple(k.shape) != (dims, dims, dims // 2 + 1):
        raise ValueError("field must be (dims, dims, dims) and filter_k (dims, dims, dims//2+1)")
    out = _lib.empty((dims, dims, dims), _lib.F32T)
    _lib.check(_lib.lib().pyl_field_smoothing(_lib.ptr(f), _lib.ptr(k), _lib.ptr(out), dims, _lib.stream()))
    return out if _lib.is_dev(field) else _lib.to_host(out)
