"""End-to-end power spectrum of a Gadget snapshot (Pk_snapshot.py:34-91) on the GPU.

``Pk_comp(snapshot_fname, ptype, dims, do_RSD, axis, cpus, folder_out)`` reads the header and the
positions (and velocities, and masses for ptype = -1), hands them to the fused device path --
redshift-space shift, CIC deposit with delta = rho/rhobar - 1, r2c, deconvolution, binning -- and
writes the reference's table ``[k3D, P0, P2, P4, Nmodes3D]`` under the reference's file name.
"""
import numpy as np

from . import readgadget
from .pipeline import PkParticles

# file-name tags of the particle types (Pk_snapshot.py:20-25)
name_dict = {"0": "GAS", "01": "GCDM", "02": "GNU", "04": "Gstars",
             "1": "CDM", "12": "CDMNU", "14": "CDMStars",
             "2": "NU", "24": "NUStars",
             "4": "Stars",
             "-1": "matter"}


def output_name(folder_out, ptype, do_RSD, axis, redshift):
    z = "%.3f" % redshift
    fout = folder_out + "/Pk_" + name_dict[str(ptype)]
    if do_RSD:
        fout += "_RS_axis=" + str(axis) + "_z=" + z + ".dat"
    else:
        fout += "_z=" + z + ".dat"
    return fout


def Pk_comp(snapshot_fname, ptype, dims, do_RSD, axis, cpus, folder_out):
    """Returns the PkParticles result and writes the table (Pk_snapshot.py:34-91).  ptype -1 = all
    particle types, weighted by their masses (mass table, or the MASS block for hydro runs)."""
    head = readgadget.header(snapshot_fname)
    BoxSize = head.boxsize / 1e3   # Mpc/h
    Masses = head.massarr * 1e10   # Msun/h
    Nall = head.nall
    redshift = head.redshift
    Hubble = 100.0 * np.sqrt(head.omega_m * (1.0 + redshift) ** 3 + head.omega_l)
    fout = output_name(folder_out, ptype, do_RSD, axis, redshift)
    pos = readgadget.read_block(snapshot_fname, "POS ", [ptype]) / 1e3   # float32, Mpc/h
    vel = readgadget.read_block(snapshot_fname, "VEL ", [ptype]) if do_RSD else None
    W = None
    if ptype == -1:
        if Nall[0] == 0:
            W = np.concatenate([np.full(int(Nall[t]), Masses[t], dtype=np.float32) for t in range(6)])
        else:
            W = (readgadget.read_block(snapshot_fname, "MASS", [-1]) * 1e10).astype(np.float32)
    Pk = PkParticles(pos, BoxSize, dims, MAS="CIC", axis=axis, W=W,
                     vel=vel, Hubble=Hubble if do_RSD else None, redshift=redshift, rsd_axis=axis)
    np.savetxt(fout, np.transpose([Pk.k3D, Pk.Pk[:, 0], Pk.Pk[:, 1], Pk.Pk[:, 2], Pk.Nmodes3D]))
    return Pk
