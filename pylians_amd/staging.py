"""Device staging of Gadget snapshots for the snapshot drivers (MAS_gadget, Pk_snapshot).

A snapshot is a set of files, each holding a slice of every particle type.  ``Staged`` lays the
slices of the requested types end to end in ONE device buffer per block (positions, velocities,
weights), so that a driver makes a single deposit / P(k) call over the whole snapshot instead of
one per file and type.  Conventions of the reference's readers that the results depend on
(readgadget.py:63-154, readsnap.py:297-298, 375):

* positions: kpc/h on disk, divided by 1e3 in float32 -> Mpc/h;
* velocities: peculiar, i.e. the stored value times sqrt(a) in float32 when z != 0;
* masses: 1e10 Msun/h on disk -> Msun/h, rounded to float32; the mass table where it is set,
  otherwise the MASS block.
"""
import math

import numpy as np

from . import _lib, readgadget

ALL_TYPES = (0, 1, 2, 3, 4, 5)


def expand_types(ptypes):
    """[-1] (or -1 in the list) means every particle type."""
    out = []
    for t in ptypes:
        out += list(ALL_TYPES) if t == -1 else [int(t)]
    return out


class Staged:
    """Segment table of one snapshot for a list of particle types.

    ``segments``: [(file name, type, first row, rows)] in file-major order; ``rows``: total;
    ``mass_table``: the first file's mass table in Msun/h (float64)."""

    def __init__(self, snapshot, types):
        self.files = readgadget.snapshot_files(snapshot)
        h = self.files[0][1]
        self.head = h
        self.types = list(types)
        self.mass_table = np.asarray(h.massarr, dtype=np.float64) * 1e10
        self.segments = []
        row = 0
        for fname, fh in self.files:
            for t in self.types:
                n = int(fh.npart[t])
                if n:
                    self.segments.append((fname, t, row, n))
                    row += n
        self.rows = row
        self.scale_vel = np.float32(math.sqrt(h.time)) if h.redshift != 0 else None
        self.box = h.boxsize / 1e3                                   # Mpc/h
        # H(z) in km/s/(Mpc/h) of the header's cosmology (readgadget.py:49)
        self.hubble_z = 100.0 * np.sqrt(h.omega_m * (1.0 + h.redshift) ** 3 + h.omega_l)

    def _fill(self, width, block, convert):
        """One device buffer of `rows` x `width` float32 filled segment by segment from the files."""
        shape = (self.rows, width) if width > 1 else (self.rows,)
        dev = _lib.empty(shape, _lib.F32T)
        for fname, t, first, n in self.segments:
            _lib.copy_rows(dev, first, convert(t, readgadget.read_segment(fname, block, t)))
        return dev

    def positions(self):
        """(rows, 3) float32 device positions in Mpc/h."""
        return self._fill(3, "POS ", lambda t, a: a / np.float32(1e3))

    def velocities(self):
        """(rows, 3) float32 device peculiar velocities in km/s."""
        s = self.scale_vel
        return self._fill(3, "VEL ", lambda t, a: a if s is None else a * s)

    def weights(self):
        """(rows,) float32 device masses in Msun/h: the mass table value of a type where it is set,
        the MASS block otherwise."""
        dev = _lib.empty((self.rows,), _lib.F32T)
        for fname, t, first, n in self.segments:
            if self.mass_table[t] != 0:
                _lib.copy_rows(dev, first, np.full(n, self.mass_table[t], dtype=np.float32))
            else:
                _lib.copy_rows(dev, first, readgadget.read_segment(fname, "MASS", t) * np.float32(1e10))
        return dev

    def type_mass(self, t):
        """Total mass of type t in Msun/h (float64): mass table x count, or the sum of its MASS block."""
        n = sum(s[3] for s in self.segments if s[1] == t)
        if self.mass_table[t] != 0:
            return self.mass_table[t] * n
        total = 0.0
        for fname, tt, _, _ in self.segments:
            if tt == t:
                total += float(np.sum(readgadget.read_segment(fname, "MASS", t) * np.float32(1e10), dtype=np.float64))
        return total
