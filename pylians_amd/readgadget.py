"""Gadget snapshot reader (format 1 and 2 binary, either byte order, single or multi-file) with the
call surface of the reference's readgadget (readgadget.py:17-154, readsnap.py): ``header(snapshot)``,
``read_field(file, block, ptype)`` and ``read_block(snapshot, block, ptype)`` for the blocks the P(k) driver needs ("POS ", "VEL ",
"MASS", "ID  ").  Arrays are read with numpy memory maps, one contiguous slice per file and type.

HDF5 snapshots need h5py, which this environment does not provide: they raise.
"""
import math
import os

import numpy as np

__all__ = ["header", "read_field", "read_block"]

_HEADER_BYTES = 256
_BLOCK_NUM = {"POS ": 2, "VEL ": 3, "ID  ": 4, "MASS": 5}   # readsnap.py order of the format-1 blocks


def _files(snapshot):
    """(list of files, format tag) of a snapshot name (readgadget.py:5-15)."""
    if os.path.exists(snapshot):
        if snapshot.endswith("hdf5"):
            return [snapshot], "hdf5"
        return [snapshot], "binary"
    if os.path.exists(snapshot + ".0"):
        h = _Header(snapshot + ".0")
        return ["%s.%d" % (snapshot, i) for i in range(max(1, h.filenum))], "binary"
    if os.path.exists(snapshot + ".hdf5") or os.path.exists(snapshot + ".0.hdf5"):
        return [], "hdf5"
    raise Exception("File not found!")


class _Header:
    """The 256-byte Gadget header of one file (readsnap.py:13-75)."""

    def __init__(self, fname):
        with open(fname, "rb") as f:
            first = np.fromfile(f, dtype="<i4", count=1)[0]
            if first in (8, 256):
                endian = "<"
            elif first.byteswap() in (8, 256):
                endian = ">"
            else:
                raise Exception("incorrect file format encountered when reading header of %s" % fname)
            first = int(first if endian == "<" else first.byteswap())
            self.format = 2 if first == 8 else 1
            self.endian = endian
            if self.format == 2:
                f.seek(16, os.SEEK_CUR)   # label record (4 + 4 + 4 + 4) and the header record's length
            raw = f.read(_HEADER_BYTES)
        e = endian
        o = 0

        def take(dt, n):
            nonlocal o
            a = np.frombuffer(raw, dtype=np.dtype(dt).newbyteorder(e), count=n, offset=o)
            o += a.nbytes
            return a.astype(np.dtype(dt).newbyteorder("="))

        self.npart = take("i4", 6)
        self.massarr = take("f8", 6)
        self.time = float(take("f8", 1)[0])
        self.redshift = float(take("f8", 1)[0])
        self.sfr = int(take("i4", 1)[0])
        self.feedback = int(take("i4", 1)[0])
        self.nall = take("u4", 6)
        self.cooling = int(take("i4", 1)[0])
        self.filenum = int(take("i4", 1)[0])
        self.boxsize = float(take("f8", 1)[0])
        self.omega_m = float(take("f8", 1)[0])
        self.omega_l = float(take("f8", 1)[0])
        self.hubble = float(take("f8", 1)[0])


class header:
    """Snapshot header with the reference's attribute names (readgadget.py:18-54)."""

    def __init__(self, snapshot):
        files, fformat = _files(snapshot)
        if fformat == "hdf5":
            raise Exception("HDF5 snapshots need h5py, which is not available")
        h = _Header(files[0])
        for name in ("time", "redshift", "boxsize", "filenum", "omega_m", "omega_l", "hubble", "massarr", "npart",
                     "nall", "cooling", "format"):
            setattr(self, name, getattr(h, name))
        self.Hubble = 100.0 * np.sqrt(self.omega_m * (1.0 + self.redshift) ** 3 + self.omega_l)


def _find_block(fname, h, block):
    """(byte offset, byte size) of a block's payload (readsnap.py:78-131)."""
    size = os.path.getsize(fname)
    dt = np.dtype("u4").newbyteorder(h.endian)
    with open(fname, "rb") as f:
        num = 1
        while f.tell() < size:
            label = None
            if h.format == 2:
                f.seek(4, os.SEEK_CUR)
                label = f.read(4).decode("latin-1")
                f.seek(8, os.SEEK_CUR)
            n = int(np.fromfile(f, dtype=dt, count=1)[0])
            if (label == block) if h.format == 2 else (num == _BLOCK_NUM[block]):
                return f.tell(), n
            f.seek(n, os.SEEK_CUR)
            if int(np.fromfile(f, dtype=dt, count=1)[0]) != n:
                raise Exception("something wrong with the record markers of %s" % fname)
            num += 1
    raise Exception("Error: block %r not found in %s" % (block, fname))


def _read_type(files, block, pt):
    """Block `block` of particle type pt over all files."""
    vec = block in ("POS ", "VEL ")
    parts = []
    ids64 = None
    for fname in files:
        h = _Header(fname)
        n_pt = int(h.npart[pt])
        if block == "MASS" and h.massarr[pt] > 0:
            parts.append(np.full(n_pt, h.massarr[pt], dtype=np.float32))
            continue
        has = np.ones(6, bool) if block != "MASS" else (h.massarr == 0)
        if not has[pt]:
            raise Exception("Error: no data for particle type %d in the block %s" % (pt, block))
        counts = np.where(has, h.npart, 0).astype(np.int64)
        off, nbytes = _find_block(fname, h, block)
        per = nbytes // max(1, int(counts.sum()))
        if block == "ID  ":
            ids64 = per == 8
            dt = np.dtype("u8" if ids64 else "u4")
        else:
            dt = np.dtype("f4")
        dt = dt.newbyteorder(h.endian)
        width = 3 if vec else 1
        first = int(counts[:pt].sum())
        if n_pt == 0:
            parts.append(np.zeros((0, 3) if vec else 0, dtype=dt.newbyteorder("=")))
            continue
        m = np.memmap(fname, dtype=dt, mode="r", offset=off + first * width * dt.itemsize, shape=(n_pt * width,))
        a = np.asarray(m).astype(dt.newbyteorder("="))
        del m
        parts.append(a.reshape(n_pt, 3) if vec else a)
    return np.concatenate(parts) if parts else np.zeros(0, np.float32)


def read_field(snapshot, block, ptype):
    """Block `block` of particle type `ptype` (an int) of ONE file of a snapshot (readgadget.py:63-97);
    velocities times sqrt(a) when z != 0 (readsnap.py:375), masses from the mass table where it is set."""
    if block not in _BLOCK_NUM:
        raise Exception("block not implemented in readgadget!")
    files, fformat = _files(snapshot)
    if fformat == "hdf5":
        raise Exception("HDF5 snapshots need h5py, which is not available")
    out = _read_type(files[:1], block, int(ptype))
    h = _Header(files[0])
    if block == "VEL " and h.redshift != 0:
        out *= np.float32(math.sqrt(h.time))
    return out


def read_block(snapshot, block, ptype, verbose=False):
    """Concatenation over the particle types in `ptype` (a list; -1 = all six types) of one block
    (readgadget.py:106-154).  Velocities are peculiar (times sqrt(a)) when z != 0, as readsnap's
    physical_velocities default (readsnap.py:297-298)."""
    if block not in _BLOCK_NUM:
        raise Exception("block not implemented in readgadget!")
    files, fformat = _files(snapshot)
    if fformat == "hdf5":
        raise Exception("HDF5 snapshots need h5py, which is not available")
    types = []
    for pt in ptype:
        types += list(range(6)) if pt == -1 else [pt]
    arrays = [_read_type(files, block, pt) for pt in types]
    out = np.concatenate(arrays)
    h = _Header(files[0])
    if block == "VEL " and h.redshift != 0:
        out *= np.float32(math.sqrt(h.time))
    return out


def snapshot_files(snapshot):
    """The files of a snapshot, in order, with their headers: [(file name, header), ...]."""
    files, fformat = _files(snapshot)
    if fformat == "hdf5":
        raise Exception("HDF5 snapshots need h5py, which is not available")
    return [(f, _Header(f)) for f in files]


def read_segment(fname, block, ptype):
    """Block `block` of particle type `ptype` of the single file `fname`, as stored (no velocity or
    unit scaling; masses from the mass table where it is set)."""
    if block not in _BLOCK_NUM:
        raise Exception("block not implemented in readgadget!")
    return _read_type([fname], block, int(ptype))
