"""Writes small Gadget format-1/2 snapshots (the public Gadget-2 file layout: a 256-byte header and
Fortran-record blocks POS, VEL, ID, MASS; format 2 labels every block) for the reader tests."""
import numpy as np


def _rec(f, payload, e, label=None):
    if label is not None:
        f.write(np.array([8], e + "u4").tobytes() + label.encode("latin-1") +
                np.array([len(payload) + 8], e + "u4").tobytes() + np.array([8], e + "u4").tobytes())
    n = np.array([len(payload)], e + "u4").tobytes()
    f.write(n + payload + n)


def write_snapshot(base, pos, vel, ids, masses, massarr, npart_files, time, redshift, boxsize,
                   omega_m=0.3, omega_l=0.7, hubble=0.7, fmt=1, endian="<"):
    """pos, vel: lists (per type) of (n, 3) float32; ids: per type uint32; masses: per type float32
    (used where massarr == 0); npart_files: list over files of 6 per-type counts."""
    e = endian
    nfiles = len(npart_files)
    nall = np.sum(npart_files, axis=0)
    start = np.zeros(6, np.int64)
    for fi, npf in enumerate(npart_files):
        fname = base if nfiles == 1 else "%s.%d" % (base, fi)
        npf = np.asarray(npf)
        hdr = (np.array(npf, e + "i4").tobytes() + np.array(massarr, e + "f8").tobytes() +
               np.array([time, redshift], e + "f8").tobytes() + np.array([0, 0], e + "i4").tobytes() +
               np.array(nall, e + "u4").tobytes() + np.array([0, nfiles], e + "i4").tobytes() +
               np.array([boxsize, omega_m, omega_l, hubble], e + "f8").tobytes())
        hdr += b"\0" * (256 - len(hdr))
        sl = [slice(start[t], start[t] + npf[t]) for t in range(6)]
        blocks = {
            "POS ": b"".join(np.asarray(pos[t][sl[t]], e + "f4").tobytes() for t in range(6)),
            "VEL ": b"".join(np.asarray(vel[t][sl[t]], e + "f4").tobytes() for t in range(6)),
            "ID  ": b"".join(np.asarray(ids[t][sl[t]], e + "u4").tobytes() for t in range(6)),
            "MASS": b"".join(np.asarray(masses[t][sl[t]], e + "f4").tobytes() for t in range(6) if massarr[t] == 0),
        }
        with open(fname, "wb") as f:
            _rec(f, hdr, e, "HEAD" if fmt == 2 else None)
            for name in ("POS ", "VEL ", "ID  ", "MASS"):
                if name == "MASS" and not blocks[name]:
                    continue
                _rec(f, blocks[name], e, name if fmt == 2 else None)
        start += npf
