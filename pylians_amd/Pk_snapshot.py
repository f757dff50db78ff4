"""End-to-end power spectra of a Gadget snapshot (Pk_snapshot.py:34-259) on the GPU.

``Pk_comp(snapshot_fname, ptype, dims, do_RSD, axis, cpus, folder_out)`` reads the header and the
positions (and velocities, and masses for ptype = -1), hands them to the fused device path --
redshift-space shift, CIC deposit with delta = rho/rhobar - 1, r2c, deconvolution, binning -- and
writes the reference's table ``[k3D, P0, P2, P4, Nmodes3D]`` under the reference's file name.

``Pk_Gadget(snapshot_fname, dims, particle_type, do_RSD, axis, cpus, folder_out=None)``: one type
goes to Pk_comp; several types get one device overdensity grid each, the auto and cross spectra of
every pair (XPk) and the spectrum of their Omega-weighted sum, each written as a table.
"""
import os

import numpy as np
import torch

from . import _lib, readgadget
from .MAS_library import MA
from .Pk_library import Pk, XPk
from .pipeline import PkParticles
from .redshift_space_library import pos_redshift_space

rho_crit = 2.77536627e11   # h^2 Msun/Mpc^3 (units_library.py:10)

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


def _div_sub(field, div, sub):
    _lib.check(_lib.lib().pyl_field_div_sub(_lib.ptr(field), int(field.numel()), float(div), float(sub), _lib.F32,
                                            _lib.stream()))


def _axpy(y, x, a):
    _lib.check(_lib.lib().pyl_field_axpy(_lib.ptr(y), _lib.ptr(x), int(y.numel()), float(a), _lib.F32,
                                         _lib.stream()))


def _table(fname, k, P, Nmodes):
    np.savetxt(fname, np.transpose([k, P[:, 0], P[:, 1], P[:, 2], Nmodes]))


def Pk_Gadget(snapshot_fname, dims, particle_type, do_RSD, axis, cpus, folder_out=None):
    """Power spectra of one or several particle types (Pk_snapshot.py:107-259); returns None.
    Overdensities follow the reference's float32 steps: CIC counts, /= N_p/dims^3, -= 1; the total
    is sum(Omega_i delta_i) / sum(Omega_i) with the numpy float32 rounding of each step."""
    if folder_out is None:
        folder_out = os.getcwd()
    if len(particle_type) == 1:
        Pk_comp(snapshot_fname, particle_type[0], dims, do_RSD, axis, cpus, folder_out)
        return None
    head = readgadget.header(snapshot_fname)
    BoxSize = head.boxsize / 1e3
    Nall = head.nall
    Masses = head.massarr * 1e10
    redshift = head.redshift
    Hubble = 100.0 * np.sqrt(head.omega_m * (1.0 + redshift) ** 3 + head.omega_l)
    z = "%.3f" % redshift
    dims3 = dims ** 3
    Omega_c = Masses[1] * Nall[1] / BoxSize ** 3 / rho_crit
    Omega_n = Masses[2] * Nall[2] / BoxSize ** 3 / rho_crit
    Omega_g, Omega_s = 0.0, 0.0
    if Nall[0] > 0:
        if Masses[0] > 0:
            Omega_g = Masses[0] * Nall[0] / BoxSize ** 3 / rho_crit
            Omega_s = Masses[4] * Nall[4] / BoxSize ** 3 / rho_crit
        else:
            mass = readgadget.read_block(snapshot_fname, "MASS", [0]) * np.float32(1e10)
            Omega_g = np.sum(mass, dtype=np.float64) / BoxSize ** 3 / rho_crit
            mass = readgadget.read_block(snapshot_fname, "MASS", [4]) * np.float32(1e10)
            Omega_s = np.sum(mass, dtype=np.float64) / BoxSize ** 3 / rho_crit
            del mass
    print("Omega_gas    = ", Omega_g)
    print("Omega_cdm    = ", Omega_c)
    print("Omega_nu     = ", Omega_n)
    print("Omega_star   = ", Omega_s)
    print("Omega_m      = ", Omega_g + Omega_c + Omega_n + Omega_s)
    print("Omega_m snap = ", head.omega_m)
    Omega_dict = {0: Omega_g, 1: Omega_c, 2: Omega_n, 4: Omega_s}
    index_dict = {0: 0, 1: 1, 2: 2, 4: 3}   # the stars' delta is delta[3]
    suffix = ("_RS_axis=" + str(axis) + "_z=" + z + ".dat") if do_RSD else ("_z=" + z + ".dat")

    delta = {}
    for ptype in particle_type:
        pos = _lib.to_dev(readgadget.read_block(snapshot_fname, "POS ", [ptype]) / np.float32(1e3), torch.float32)
        if do_RSD:
            vel = _lib.to_dev(readgadget.read_block(snapshot_fname, "VEL ", [ptype]), torch.float32)
            pos_redshift_space(pos, vel, BoxSize, Hubble, redshift, axis)
            del vel
        mean_number = len(pos) * 1.0 / dims3
        d = torch.zeros((dims, dims, dims), dtype=torch.float32, device="cuda")
        MA(pos, d, BoxSize, "CIC")
        del pos
        _div_sub(d, mean_number, 1.0)
        delta[index_dict[ptype]] = d

    for i, ptype1 in enumerate(particle_type):
        for ptype2 in particle_type[i + 1:]:
            index1, index2 = index_dict[ptype1], index_dict[ptype2]
            fout1 = folder_out + "/Pk_" + name_dict[str(ptype1)] + suffix
            fout2 = folder_out + "/Pk_" + name_dict[str(ptype2)] + suffix
            fout12 = folder_out + "/Pk_" + name_dict[str(ptype1) + str(ptype2)] + suffix
            print("\nComputing the auto- and cross-power spectra of types: ", ptype1, "-", ptype2)
            print("saving results in:")
            print(fout1, "\n", fout2, "\n", fout12)
            data = XPk([delta[index1], delta[index2]], BoxSize, axis=axis, MAS=["CIC", "CIC"], threads=cpus)
            _table(fout12, data.k3D, data.XPk[:, :, 0], data.Nmodes3D)
            _table(fout1, data.k3D, data.Pk[:, :, 0], data.Nmodes3D)
            _table(fout2, data.k3D, data.Pk[:, :, 1], data.Nmodes3D)

    print("\ncomputing P(k) of all components")
    delta_tot = torch.zeros((dims, dims, dims), dtype=torch.float32, device="cuda")
    Omega_tot = 0.0
    fout = folder_out + "/Pk_"
    for ptype in particle_type:
        _axpy(delta_tot, delta[index_dict[ptype]], Omega_dict[ptype])
        Omega_tot += Omega_dict[ptype]
        fout += name_dict[str(ptype)] + "+"
    _div_sub(delta_tot, Omega_tot, 0.0)
    del delta
    fout = fout[:-1]
    data = Pk(delta_tot, BoxSize, axis=axis, MAS="CIC", threads=cpus)
    del delta_tot
    _table(fout + suffix, data.k3D, data.Pk, data.Nmodes3D)

