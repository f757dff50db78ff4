"""Density field of a Gadget snapshot on the GPU (MAS_gadget.py:7-83).

``density_field_gadget(snapshot_fname, ptypes, dims, MAS='CIC', do_RSD=False, axis=0, verbose=True)``
walks the files of the snapshot and the requested particle types, moves each chunk of positions to
redshift space when asked, and accumulates it into one float32 grid that stays on the device
until the end.  A single component is deposited unweighted when the mass table gives its mass;
several components, or a component with a MASS block, are weighted by the particle masses in
Msun/h.  Returns the (dims, dims, dims) float32 numpy grid.
"""
import time

import numpy as np
import torch

from . import _lib, readgadget
from .MAS_library import MA
from .redshift_space_library import pos_redshift_space

__all__ = ["density_field_gadget"]


def density_field_gadget(snapshot_fname, ptypes, dims, MAS="CIC", do_RSD=False, axis=0, verbose=True):
    start = time.time()
    if verbose:
        print("\nComputing density field of particles", ptypes)
    density = torch.zeros((dims, dims, dims), dtype=torch.float32, device="cuda")
    head = readgadget.header(snapshot_fname)
    BoxSize = head.boxsize / 1e3          # Mpc/h
    Masses = head.massarr * 1e10          # Msun/h
    filenum = head.filenum
    redshift = head.redshift
    Hubble = head.Hubble
    if ptypes == [-1]:
        ptypes = [0, 1, 2, 3, 4, 5]
    single_component = len(ptypes) == 1
    num = 0.0
    for i in range(filenum):
        snapshot = snapshot_fname if filenum == 1 else "%s.%d" % (snapshot_fname, i)
        npart = readgadget.header(snapshot).npart
        for ptype in ptypes:
            if npart[ptype] == 0:
                continue
            pos = _lib.to_dev(readgadget.read_field(snapshot, "POS ", ptype) / np.float32(1e3), torch.float32)
            if do_RSD:
                vel = _lib.to_dev(readgadget.read_field(snapshot, "VEL ", ptype), torch.float32)
                pos_redshift_space(pos, vel, BoxSize, Hubble, redshift, axis)
                del vel
            if Masses[ptype] != 0:
                if single_component:
                    MA(pos, density, BoxSize, MAS)
                    num += pos.shape[0]
                else:
                    mass = np.ones(int(npart[ptype]), dtype=np.float32) * np.float32(Masses[ptype])
                    MA(pos, density, BoxSize, MAS, W=mass)
                    num += np.sum(mass, dtype=np.float64)
            else:
                mass = readgadget.read_field(snapshot, "MASS", ptype) * np.float32(1e10)
                MA(pos, density, BoxSize, MAS, W=mass)
                num += np.sum(mass, dtype=np.float64)
    out = _lib.to_host(density)
    if verbose:
        print("%.8e should be equal to\n%.8e" % (np.sum(out, dtype=np.float64), num))
        print("Time taken = %.2f seconds" % (time.time() - start))
    return out
