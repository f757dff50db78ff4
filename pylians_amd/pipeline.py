"""Fused particles -> P(k) path (the end-to-end recipe of Pk_snapshot.Pk_comp, Pk_snapshot.py:80-91).

``PkParticles(pos, BoxSize, dims, MAS='CIC', axis=0, W=None, precision='fp32')`` deposits the
particles, forms delta = rho/rhobar - 1 with rhobar = sum(W)/N^3 (N_p/N^3 without weights),
transforms and bins, all on the GPU in one stream, without materialising delta on the host.
It exposes the same attributes as ``Pk_library.Pk``.

With ``vel``, the redshift-space shift of pos_redshift_space (redshift_space_library.pyx:29-43)
is applied to each position as the deposit loads it.

``overdensity(pos, BoxSize, dims, MAS='CIC', W=None, precision='fp32')`` stops after the deposit and
returns delta on the device (the input of Pk_library.Bk / XPk / Xi).

``run_device`` is the allocation-free inner call used by the benchmark: device inputs, a
device output vector; with ``flags | PK_ASYNC`` no host synchronisation (the caller then checks
the returned Nmodes, which the binning kernel counted).
"""
import numpy as np

from . import _lib
from .Pk_library import pk_layout, unpack_pk

PRECISION = {"fp32": _lib.F32, "fp64": _lib.F64}


def run_device(pos, W, out, BoxSize, dims, scheme, axis, prec, flags=_lib.PK_ALL, rsd=None):
    """pos: (n,3) float32 device tensor; W: (n,) float32 device tensor or None; out: float64 device
    tensor of pk_layout(dims)[3] values; rsd: None or (vel (n,3) float32 device tensor, Hubble,
    redshift, axis) for the fused redshift-space shift (redshift_space_library.pyx:29-43)."""
    if rsd is None:
        _lib.check(_lib.lib().pyl_pk_particles(_lib.ptr(pos), _lib.ptr(W), int(pos.shape[0]), int(dims),
                                               float(BoxSize), int(scheme), int(axis), int(prec), int(flags),
                                               _lib.ptr(out), _lib.stream()))
        return
    vel, hubble, redshift, rsd_axis = rsd
    _lib.check(_lib.lib().pyl_pk_particles_rsd(_lib.ptr(pos), _lib.ptr(vel), _lib.ptr(W), int(pos.shape[0]),
                                               int(dims), float(BoxSize), float(hubble), float(redshift),
                                               int(rsd_axis), int(scheme), int(axis), int(prec), int(flags),
                                               _lib.ptr(out), _lib.stream()))


class PkParticles:
    def __init__(self, pos, BoxSize, dims, MAS="CIC", axis=0, W=None, precision="fp32", flags=_lib.PK_ALL,
                 deposit=None, vel=None, Hubble=None, redshift=0.0, rsd_axis=None):
        """With `vel`, the particles are moved to redshift space along `rsd_axis` (default: `axis`)
        as pos_redshift_space(pos, vel, BoxSize, Hubble, redshift, rsd_axis) would
        (redshift_space_library.pyx:29-43), fused into the deposit; `pos` is not modified."""
        if MAS not in _lib.SCHEMES:
            raise ValueError("unknown MAS %r" % (MAS,))
        p = _lib.to_dev(pos, _lib.F32T)
        if p.ndim != 2 or p.shape[1] != 3:
            raise ValueError("pos must be (n, 3)")
        w = None if W is None else _lib.to_dev(W, _lib.F32T)
        rsd = None
        if vel is not None:
            if Hubble is None:
                raise ValueError("redshift space needs Hubble")
            v = _lib.to_dev(vel, _lib.F32T)
            if tuple(v.shape) != tuple(p.shape):
                raise ValueError("vel must have the shape of pos")
            rsd = (v, Hubble, redshift, axis if rsd_axis is None else rsd_axis)
        _, _, _, nout = pk_layout(dims)
        out = _lib.empty((nout,), _lib.F64T)
        f = flags | {None: 0, "atomic": _lib.ATOMIC, "binned": _lib.BINNED, "presorted": _lib.SORTED}[deposit]
        try:
            run_device(p, w, out, BoxSize, dims, _lib.SCHEMES[MAS], axis, PRECISION[precision], f, rsd)
        except _lib.PylError as e:
            if e.code == _lib.E_MODES:
                from .Pk_library import _modes_exit
                _modes_exit(e.code)
            raise
        unpack_pk(self, _lib.to_host(out), dims)


def overdensity(pos, BoxSize, dims, MAS="CIC", W=None, precision="fp32", deposit=None):
    """delta = rho / rhobar - 1 on a dims^3 device grid (float32, or float64 with precision='fp64'),
    the MA call plus normalisation of Pk_snapshot.py:84-88 in one entry point: the deposit's float64
    cell sums are normalised before the single rounding."""
    if MAS not in _lib.SCHEMES:
        raise ValueError("unknown MAS %r" % (MAS,))
    p = _lib.to_dev(pos, _lib.F32T)
    if p.ndim != 2 or p.shape[1] != 3:
        raise ValueError("pos must be (n, 3)")
    w = None if W is None else _lib.to_dev(W, _lib.F32T)
    f64 = PRECISION[precision] == _lib.F64
    grid = _lib.empty((dims,) * 3, _lib.F64T if f64 else _lib.F32T)
    f = {None: 0, "atomic": _lib.ATOMIC, "binned": _lib.BINNED, "presorted": _lib.SORTED}[deposit]
    _lib.check(_lib.lib().pyl_mas_overdensity(_lib.ptr(p), _lib.ptr(w), int(p.shape[0]), int(dims), float(BoxSize),
                                              _lib.SCHEMES[MAS], f | (_lib.GRID_F64 if f64 else 0), _lib.ptr(grid),
                                              _lib.stream()))
    return grid
