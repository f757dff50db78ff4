"""Power spectra of Gadget snapshots on the GPU, with the call surface of Pk_snapshot.py:34-267.

``Pk_comp(snapshot_fname, ptype, dims, do_RSD, axis, cpus, folder_out)``: P0/P2/P4 of one particle
type (or of all matter, ptype = -1, mass weighted) in real or redshift space, written as the
reference's table ``[k3D, P0, P2, P4, Nmodes3D]`` under the reference's file name; returns the
result object.

``Pk_Gadget(snapshot_fname, dims, particle_type, do_RSD, axis, cpus, folder_out=None)``: one type is
``Pk_comp``; for several types, the auto and cross spectra of every pair and the spectrum of the
Omega-weighted total are written as tables; returns None.

Design.  Each type is staged once into device memory (``staging.Staged``) and never returns to the
host before the tables.  ``Pk_comp`` is one fused call (RSD shift, weighted deposit with
delta = rho/rhobar - 1, r2c, deconvolution, binning).  ``Pk_Gadget`` builds one device overdensity
per type (the reference's float32 steps rho / (N_t / dims^3) - 1), makes ONE multi-field XPk call for
all autos and crosses, and obtains the total from that matrix instead of a further grid and FFT:
delta_tot = sum_i w_i delta_i / sum_i w_i with w_i = Omega_i, and the binned power of a linear
combination is the same combination of the binned autos and crosses,
P_tot = sum_ij w_i w_j P_ij / (sum_i w_i)^2, multipole by multipole (every field carries the same
CIC deconvolution).
"""
import os

import numpy as np

from . import _lib, staging
from .MAS_library import MA
from .Pk_library import XPk
from .pipeline import PkParticles
from .redshift_space_library import pos_redshift_space

__all__ = ["Pk_comp", "Pk_Gadget", "name_dict", "rho_crit"]

rho_crit = 2.77536627e11   # h^2 Msun/Mpc^3 (units_library.py:10)

# file-name tags of particle types and pairs: output format of Pk_snapshot.py:20-25
name_dict = {"0": "GAS", "01": "GCDM", "02": "GNU", "04": "Gstars",
             "1": "CDM", "12": "CDMNU", "14": "CDMStars",
             "2": "NU", "24": "NUStars",
             "4": "Stars",
             "-1": "matter"}

_SPECIES = ((0, "Omega_gas    = "), (1, "Omega_cdm    = "), (2, "Omega_nu     = "), (4, "Omega_star   = "))


def _suffix(do_RSD, axis, redshift):
    z = "%.3f" % redshift
    return ("_RS_axis=%d_z=%s.dat" % (axis, z)) if do_RSD else ("_z=%s.dat" % z)


def output_name(folder_out, ptype, do_RSD, axis, redshift):
    """Table file of one particle type (or -1) as the reference names it."""
    return folder_out + "/Pk_" + name_dict[str(ptype)] + _suffix(do_RSD, axis, redshift)


def _write_table(fname, k, P, Nmodes):
    """[k, P0, P2, P4, Nmodes] columns; P is (kmax, 3)."""
    np.savetxt(fname, np.column_stack([k, P[:, 0], P[:, 1], P[:, 2], Nmodes]))


def _print_ranges(pos):
    if _lib.is_tensor(pos):
        lo = pos.min(dim=0).values.cpu().numpy()
        hi = pos.max(dim=0).values.cpu().numpy()
    else:
        host = _lib.to_host(pos)
        lo, hi = host.min(axis=0), host.max(axis=0)
    for c, name in enumerate("XYZ"):
        print("%.3f < %s [Mpc/h] < %.3f%s" % (lo[c], name, hi[c], "\n" if c == 2 else ""))


def Pk_comp(snapshot_fname, ptype, dims, do_RSD, axis, cpus, folder_out):
    print("Computing power spectrum...")
    snap = staging.Staged(snapshot_fname, staging.expand_types([ptype]))
    z = snap.head.redshift
    pos = snap.positions()
    _print_ranges(pos)
    vel = None
    if do_RSD:
        print("moving particles to redshift-space...")
        vel = snap.velocities()
    W = snap.weights() if ptype == -1 else None
    res = PkParticles(pos, snap.box, dims, MAS="CIC", axis=axis, W=W, vel=vel,
                      Hubble=snap.hubble_z if do_RSD else None, redshift=z, rsd_axis=axis)
    if do_RSD:
        print("done")
    del pos, vel, W
    _write_table(output_name(folder_out, ptype, do_RSD, axis, z), res.k3D, res.Pk, res.Nmodes3D)
    return res


def _overdensity(snapshot_fname, ptype, dims, do_RSD, axis):
    """float32 device overdensity of one type: CIC counts / (N_t / dims^3) - 1, each step rounded to
    float32 as the reference's numpy in-place operations round it."""
    snap = staging.Staged(snapshot_fname, [ptype])
    pos = snap.positions()
    if do_RSD:
        pos_redshift_space(pos, snap.velocities(), snap.box, snap.hubble_z, snap.head.redshift, axis)
    grid = _lib.zeros((dims, dims, dims), _lib.F32T)
    MA(pos, grid, snap.box, "CIC")
    del pos
    _lib.check(_lib.lib().pyl_field_div_sub(_lib.ptr(grid), int(grid.numel()), snap.rows / dims ** 3, 1.0,
                                            _lib.F32, _lib.stream()))
    return grid


def Pk_Gadget(snapshot_fname, dims, particle_type, do_RSD, axis, cpus, folder_out=None):
    if folder_out is None:
        folder_out = os.getcwd()
    if len(particle_type) == 1:
        Pk_comp(snapshot_fname, particle_type[0], dims, do_RSD, axis, cpus, folder_out)
        return None

    print("\nREADING SNAPSHOTS PROPERTIES")
    types = [int(t) for t in particle_type]
    census = staging.Staged(snapshot_fname, sorted(set(types) | {0, 1, 2, 4}))
    head = census.head
    omega = {t: census.type_mass(t) / census.box ** 3 / rho_crit for t in census.types}
    for t, label in _SPECIES:
        print(label, omega[t])
    print("Omega_m      = ", sum(omega[t] for t, _ in _SPECIES))
    print("Omega_m snap = ", head.omega_m)
    suffix = _suffix(do_RSD, axis, head.redshift)

    deltas = [_overdensity(snapshot_fname, t, dims, do_RSD, axis) for t in types]
    spec = XPk(deltas, census.box, axis=axis, MAS=["CIC"] * len(types), threads=cpus)
    del deltas
    F = len(types)
    pair = {}
    c = 0
    for i in range(F):
        for j in range(i + 1, F):
            pair[(i, j)] = c
            c += 1

    def tag(*ts):
        return folder_out + "/Pk_" + name_dict["".join(str(t) for t in ts)] + suffix

    for (i, j), c in pair.items():
        f1, f2, f12 = tag(types[i]), tag(types[j]), tag(types[i], types[j])
        print("\nComputing the auto- and cross-power spectra of types: ", types[i], "-", types[j])
        print("saving results in:")
        print(f1, "\n", f2, "\n", f12)
        _write_table(f12, spec.k3D, spec.XPk[:, :, c], spec.Nmodes3D)
        _write_table(f1, spec.k3D, spec.Pk[:, :, i], spec.Nmodes3D)
        _write_table(f2, spec.k3D, spec.Pk[:, :, j], spec.Nmodes3D)

    print("\ncomputing P(k) of all components")
    w = np.array([omega[t] for t in types])
    total = np.zeros_like(spec.Pk[:, :, 0])
    for i in range(F):
        total += w[i] * w[i] * spec.Pk[:, :, i]
    for (i, j), c in pair.items():
        total += 2.0 * w[i] * w[j] * spec.XPk[:, :, c]
    total /= w.sum() ** 2
    name = folder_out + "/Pk_" + "+".join(name_dict[str(t)] for t in types) + suffix
    _write_table(name, spec.k3D, total, spec.Nmodes3D)
    return None
