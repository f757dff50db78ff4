"""Pk_comp (Pk_snapshot.py:34-91) end to end on the GPU from a Gadget snapshot: the table matches
the FP64 oracle run on the same particles (read, kpc -> Mpc, peculiar velocities, RSD shift)."""
import math
import os

import numpy as np
import pytest

from oracle import rsd
from pylians_amd import synth
from tests._gadget import write_snapshot
from tests._gpuref import assert_pk_close, oracle_pk, oracle_pk_particles

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("do_RSD", [False, True])
def test_pk_comp(gpu, tmp_path, do_RSD):
    from pylians_amd import Pk_snapshot
    n, N, L = 48, 48, 200.0
    pos = synth.zeldovich(n, L, seed=4, rms_disp=1.0, sigma_v=0.0) * np.float32(1e3)   # kpc/h
    rng = np.random.default_rng(4)
    vel = (rng.standard_normal(pos.shape) * 200.0).astype(np.float32)
    ids = np.arange(len(pos), dtype=np.uint32)
    z, a = 0.5, 1.0 / 1.5
    empty3, empty = np.zeros((0, 3), np.float32), np.zeros(0, np.float32)
    P = [empty3, pos, empty3, empty3, empty3, empty3]
    V = [empty3, vel, empty3, empty3, empty3, empty3]
    I = [empty.astype(np.uint32), ids] + [empty.astype(np.uint32)] * 4
    M = [empty] * 6
    half = len(pos) // 2
    base = str(tmp_path / "snapdir")
    write_snapshot(base, P, V, I, M, [0, 1.0, 0, 0, 0, 0], [[0, half, 0, 0, 0, 0], [0, len(pos) - half, 0, 0, 0, 0]],
                   a, z, L * 1e3, fmt=2)
    Pk = Pk_snapshot.Pk_comp(base, 1, N, do_RSD, 0, 1, str(tmp_path))
    fout = Pk_snapshot.output_name(str(tmp_path), 1, do_RSD, 0, z)
    assert os.path.exists(fout)
    tab = np.loadtxt(fout)
    assert tab.shape == (len(Pk.k3D), 5)
    # the oracle on the same particles
    p = pos / np.float32(1e3)
    if do_RSD:
        H = 100.0 * np.sqrt(0.3 * 1.5 ** 3 + 0.7)
        p = rsd.pos_redshift_space(p, vel * np.float32(math.sqrt(a)), L, H, z, 0)
    o, nz = oracle_pk_particles(p, L, N, "CIC", 0, with_noise=True)
    assert_pk_close(Pk, o, rtol=1e-5, noise=nz)
    assert np.array_equal(tab[:, 4], o.Nmodes3D)


def _two_species_snapshot(tmp_path, fmt=1, gas_mass_block=True):
    """Two-file snapshot: gas (type 0, MASS block), CDM (type 1, mass table 1.0) and neutrinos
    (type 2, mass table 0.25), positions in kpc/h, z = 1."""
    L, z, a = 150.0, 1.0, 0.5
    rng = np.random.default_rng(21)
    cdm = synth.zeldovich(24, L, seed=21, rms_disp=1.5, sigma_v=0.0) * np.float32(1e3)
    nu = (rng.random((6000, 3)) * L * 1e3).astype(np.float32)
    gas = (rng.random((5000, 3)) * L * 1e3).astype(np.float32)
    P = [gas, cdm, nu] + [np.zeros((0, 3), np.float32)] * 3
    V = [(rng.standard_normal(p.shape) * 250.0).astype(np.float32) for p in P]
    I = [np.arange(len(p), dtype=np.uint32) for p in P]
    M = [(rng.random(len(gas)) * 0.1 + 0.05).astype(np.float32)] + [np.zeros(len(p), np.float32) for p in P[1:]]
    massarr = [0.0 if gas_mass_block else 0.1, 1.0, 0.25, 0, 0, 0]
    counts = [len(p) for p in P]
    first = [c // 3 for c in counts]
    files = [first, [c - f for c, f in zip(counts, first)]]
    base = str(tmp_path / "snap_002")
    write_snapshot(base, P, V, I, M, massarr, files, a, z, L * 1e3, fmt=fmt)
    H = 100.0 * np.sqrt(0.3 * (1.0 + z) ** 3 + 0.7)
    return base, L, z, a, H, P, V, M, massarr


def _oracle_positions(p, v, L, H, z, a, do_RSD, axis):
    q = p / np.float32(1e3)
    if do_RSD:
        q = rsd.pos_redshift_space(q, v * np.float32(math.sqrt(a)), L, H, z, axis)
    return q


@pytest.mark.parametrize("ptypes,do_RSD", [([1], False), ([1], True), ([0], False), ([1, 2], True), ([-1], False)])
def test_density_field_gadget(gpu, tmp_path, ptypes, do_RSD):
    """MAS_gadget.density_field_gadget (MAS_gadget.py:7-83): the float32 grid accumulated over files
    and types equals the float64 oracle deposit of the same (shifted) positions and masses."""
    from oracle import cpu as C
    from pylians_amd import MAS_gadget
    base, L, z, a, H, P, V, M, massarr = _two_species_snapshot(tmp_path)
    N, axis = 32, 1
    g = MAS_gadget.density_field_gadget(base, ptypes, N, MAS="CIC", do_RSD=do_RSD, axis=axis, verbose=False)
    assert g.dtype == np.float32 and g.shape == (N, N, N)
    types = [0, 1, 2, 3, 4, 5] if ptypes == [-1] else ptypes
    o = np.zeros((N, N, N))
    for t in types:
        if len(P[t]) == 0:
            continue
        q = _oracle_positions(P[t], V[t], L, H, z, a, do_RSD, axis)
        if massarr[t] != 0 and len(types) == 1:
            w = None
        elif massarr[t] != 0:
            w = np.full(len(q), np.float32(massarr[t] * 1e10), np.float32)
        else:
            w = M[t] * np.float32(1e10)
        o += C.deposit_f64(q, N, L, "CIC", w)
    assert np.all(np.abs(g - o) <= 1e-5 * np.abs(o) + 1e-6 * np.abs(o).max())


def _check_table(fname, o, noise, scale=None):
    """A [k, P0, P2, P4, Nmodes] table against the float64 oracle: k to 1e-12, Nmodes exact, P0 and
    P2/P4 relative to `scale` (P0 of the table; for a cross spectrum sqrt(P0_a P0_b) of the autos,
    as the cross P0 itself can be near zero) within max(1e-5, NOISE_Z sigma of the float32
    representation noise)."""
    from tests._gpuref import NOISE_Z, _envelope
    tab = np.loadtxt(fname)
    k, Pk, Nm = o
    assert tab.shape == (len(k), 5)
    assert np.allclose(tab[:, 0], k, rtol=1e-12, atol=0)
    assert np.array_equal(tab[:, 4], Nm)
    P0 = np.abs(Pk[:, 0]) if scale is None else scale
    for l in range(3):
        err = np.abs(tab[:, 1 + l] - Pk[:, l]) / P0
        allowed = np.maximum(1e-5, NOISE_Z * _envelope(noise[:, l] / P0))
        assert np.all(err <= allowed), (fname, l, float((err / allowed).max()))


@pytest.mark.parametrize("do_RSD", [False, True])
def test_pk_gadget_two_species(gpu, tmp_path, do_RSD):
    """Pk_Gadget (Pk_snapshot.py:107-259) for CDM + neutrinos: auto, cross and Omega-weighted total
    tables against the float64 oracle run on the overdensities built with the reference's numpy
    float32 steps (rho / mean - 1, sum of Omega * delta, / Omega_tot)."""
    from oracle import cpu as C
    from oracle import pk as PO
    from pylians_amd import Pk_snapshot
    base, L, z, a, H, P, V, M, massarr = _two_species_snapshot(tmp_path)
    N, axis = 32, 0
    assert Pk_snapshot.Pk_Gadget(base, N, [1, 2], do_RSD, axis, 1, str(tmp_path)) is None
    suffix = ("_RS_axis=%d_z=%.3f.dat" % (axis, z)) if do_RSD else ("_z=%.3f.dat" % z)
    deltas, om = [], []
    for t in (1, 2):
        q = _oracle_positions(P[t], V[t], L, H, z, a, do_RSD, axis)
        rho = C.deposit_f64(q, N, L, "CIC").astype(np.float32)
        d = rho / np.float32(len(q) * 1.0 / N ** 3)
        d -= np.float32(1.0)
        deltas.append(d)
        om.append(massarr[t] * 1e10 * len(q) / L ** 3 / Pk_snapshot.rho_crit)
    x = PO.XPk([d.astype(np.float64) for d in deltas], L, axis=axis, MAS=["CIC", "CIC"])
    nz = C.f32_noise_xpk(deltas, L, axis, ["CIC", "CIC"])
    _check_table(str(tmp_path / ("Pk_CDM" + suffix)), (x.k3D, x.Pk[:, :, 0], x.Nmodes3D), nz.Pk[:, :, 0])
    _check_table(str(tmp_path / ("Pk_NU" + suffix)), (x.k3D, x.Pk[:, :, 1], x.Nmodes3D), nz.Pk[:, :, 1])
    _check_table(str(tmp_path / ("Pk_CDMNU" + suffix)), (x.k3D, x.XPk[:, :, 0], x.Nmodes3D), nz.XPk[:, :, 0],
                 scale=np.sqrt(np.abs(x.Pk[:, 0, 0] * x.Pk[:, 0, 1])))
    tot = np.zeros((N, N, N), np.float32)
    for d, w in zip(deltas, om):
        tot += np.float32(w) * d
    tot /= np.float32(om[0] + om[1])
    o = oracle_pk(tot, L, axis, "CIC")
    _check_table(str(tmp_path / ("Pk_CDM+NU" + suffix)), (o.k3D, o.Pk, o.Nmodes3D),
                 C.f32_noise_pk(tot, L, axis, "CIC").Pk)


def test_pk_gadget_single_species_is_pk_comp(gpu, tmp_path):
    """Pk_Gadget with one particle type writes Pk_comp's table (Pk_snapshot.py:113-117); two runs of
    the path agree to the summation order of the float64 shell sums."""
    from pylians_amd import Pk_snapshot
    base, L, z, a, H, P, V, M, massarr = _two_species_snapshot(tmp_path)
    out1, out2 = tmp_path / "a", tmp_path / "b"
    out1.mkdir()
    out2.mkdir()
    assert Pk_snapshot.Pk_Gadget(base, 32, [1], False, 0, 1, str(out1)) is None
    Pk_snapshot.Pk_comp(base, 1, 32, False, 0, 1, str(out2))
    name = "Pk_CDM_z=%.3f.dat" % z
    t1, t2 = np.loadtxt(out1 / name), np.loadtxt(out2 / name)
    assert np.array_equal(t1[:, [0, 4]], t2[:, [0, 4]])
    assert np.allclose(t1, t2, rtol=1e-12, atol=1e-12 * np.abs(t2[:, 1]).max())
