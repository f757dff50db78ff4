"""Fused particles -> P(k) path (the end-to-end recipe of Pk_snapshot.Pk_comp, Pk_snapshot.py:80-91).

``PkParticles(pos, BoxSize, dims, MAS='CIC', axis=0, W=None, precision='fp32')`` deposits the
particles, forms delta = rho/rhobar - 1 with rhobar = sum(W)/N^3 (N_p/N^3 without weights),
transforms and bins, all on the GPU in one stream, without materialising delta on the host.
It exposes the same attributes as ``Pk_library.Pk``.

``run_device`` is the allocation-free inner call used by the benchmark: device inputs, a
device output vector, no host synchronisation.
"""
import numpy as np
import torch

from . import _lib
from .Pk_library import pk_layout, unpack_pk

PRECISION = {"fp32": _lib.F32, "fp64": _lib.F64}


def run_device(pos, W, out, BoxSize, dims, scheme, axis, prec, flags=_lib.PK_ALL):
    """pos: (n,3) float32 device tensor; W: (n,) float32 device tensor or None; out: float64 device
    tensor of pk_layout(dims)[3] values."""
    _lib.check(_lib.lib().pyl_pk_particles(_lib.ptr(pos), _lib.ptr(W), int(pos.shape[0]), int(dims),
                                           float(BoxSize), int(scheme), int(axis), int(prec), int(flags),
                                           _lib.ptr(out), _lib.stream()))


class PkParticles:
    def __init__(self, pos, BoxSize, dims, MAS="CIC", axis=0, W=None, precision="fp32", flags=_lib.PK_ALL,
                 deposit=None):
        if MAS not in _lib.SCHEMES:
            raise ValueError("unknown MAS %r" % (MAS,))
        p = _lib.to_dev(pos, torch.float32)
        if p.ndim != 2 or p.shape[1] != 3:
            raise ValueError("pos must be (n, 3)")
        w = None if W is None else _lib.to_dev(W, torch.float32)
        _, _, _, nout = pk_layout(dims)
        out = _lib.empty((nout,), torch.float64)
        f = flags | {None: 0, "atomic": _lib.ATOMIC, "binned": _lib.BINNED}[deposit]
        run_device(p, w, out, BoxSize, dims, _lib.SCHEMES[MAS], axis, PRECISION[precision], f)
        unpack_pk(self, _lib.to_host(out), dims)
