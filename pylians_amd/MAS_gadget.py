"""Density field of a Gadget snapshot on the GPU, with the call surface of MAS_gadget.py:7-83.

``density_field_gadget(snapshot_fname, ptypes, dims, MAS='CIC', do_RSD=False, axis=0, verbose=True)``
returns the (dims, dims, dims) float32 mass-assignment grid of the requested particle types
(``[-1]``: all six) over every file of the snapshot, optionally in redshift space along `axis`.

Design: the snapshot is staged once (``staging.Staged``: every file and type of one block end to end
in a single device buffer), moved to redshift space in place on the device, and deposited by ONE
call of the cell-binned deposit, with float64 per-cell accumulation and a single rounding.  The
weighting follows the reference's rule: a single type whose mass the header's mass table gives is
counted (no weights); otherwise every particle carries its mass in Msun/h (mass table or MASS
block).  With `verbose`, the reference's progress lines are printed: the grid's total against the
expected mass (or count), and the time taken.
"""
import time

import numpy as np

from . import _lib, staging
from .MAS_library import MA
from .redshift_space_library import pos_redshift_space

__all__ = ["density_field_gadget"]


def density_field_gadget(snapshot_fname, ptypes, dims, MAS="CIC", do_RSD=False, axis=0, verbose=True):
    t0 = time.time()
    if verbose:
        print("\nComputing density field of particles", ptypes)
    snap = staging.Staged(snapshot_fname, staging.expand_types(ptypes))
    grid = _lib.zeros((dims, dims, dims), _lib.F32T)
    if snap.rows:
        box = snap.box
        pos = snap.positions()
        if do_RSD:
            pos_redshift_space(pos, snap.velocities(), box, snap.hubble_z, snap.head.redshift, axis)
        counted = len(snap.types) == 1 and snap.mass_table[snap.types[0]] != 0
        W = None if counted else snap.weights()
        MA(pos, grid, box, MAS, W=W)
        expected = float(snap.rows) if counted else float(np.sum(_lib.to_host(W), dtype=np.float64))
        del pos, W
    else:
        expected = 0.0
    out = _lib.to_host(grid)
    if verbose:
        print("%.8e should be equal to\n%.8e" % (np.sum(out, dtype=np.float64), expected))
        print("Time taken = %.2f seconds" % (time.time() - t0))
    return out
