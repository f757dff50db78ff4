"""Redshift space fused into the deposit: the same result as the path run on the positions that
pos_redshift_space (redshift_space_library.pyx:29-43, oracle/rsd.py) produces -- the same float32
positions and density grid; the float64 shell sums agree to their summation order (1e-12)."""
import numpy as np
import pytest

from oracle import rsd
from pylians_amd import synth

pytestmark = pytest.mark.gpu


def _case(n, L, seed):
    pos = synth.uniform(n, L, seed=seed)
    rng = np.random.default_rng(seed)
    vel = (rng.standard_normal((n, 3)) * 300.0).astype(np.float32)
    vel[:7, 0] = np.float32(2.0e5)      # a few particles that cross the box several times
    vel[7:14, 0] = np.float32(-2.0e5)
    return pos, vel


@pytest.mark.parametrize("dims,MAS,prec", [(64, "CIC", "fp32"), (48, "TSC", "fp32"), (64, "PCS", "fp64"),
                                           (32, "NGP", "fp32")])
def test_fused_rsd_equals_shifted_positions(gpu, dims, MAS, prec):
    from pylians_amd import pipeline
    L, H, z = 250.0, 100.0, 0.5
    pos, vel = _case(40000, L, dims)
    ref_pos = rsd.pos_redshift_space(pos, vel, L, H, z, 0)
    assert not np.array_equal(ref_pos, pos)
    g = pipeline.PkParticles(pos, L, dims, MAS=MAS, axis=0, precision=prec, vel=vel, Hubble=H, redshift=z)
    r = pipeline.PkParticles(ref_pos, L, dims, MAS=MAS, axis=0, precision=prec)
    assert np.array_equal(g.Nmodes3D, r.Nmodes3D) and np.array_equal(g.k3D, r.k3D)
    for name in ("Pk", "Pk1D", "Pkphase"):
        assert np.allclose(getattr(g, name), getattr(r, name), rtol=1e-12, atol=0), name
    # 2-D bin 0 holds only the DC mode: rounding noise of the zero mean
    assert np.allclose(g.Pk2D[1:], r.Pk2D[1:], rtol=1e-12, atol=0)
    assert abs(g.Pk2D[0] - r.Pk2D[0]) <= 1e-12 * np.abs(r.Pk2D).max()


def test_fused_rsd_weighted_other_axis(gpu):
    from pylians_amd import pipeline
    L, H, z = 300.0, 70.0, 1.0
    pos, vel = _case(30000, L, 5)
    W = (np.random.default_rng(5).random(len(pos)) + 0.5).astype(np.float32)
    ref_pos = rsd.pos_redshift_space(pos, vel, L, H, z, 2)
    g = pipeline.PkParticles(pos, L, 64, MAS="CIC", axis=2, W=W, vel=vel, Hubble=H, redshift=z, rsd_axis=2)
    r = pipeline.PkParticles(ref_pos, L, 64, MAS="CIC", axis=2, W=W)
    assert np.allclose(g.Pk, r.Pk, rtol=1e-12, atol=0) and np.allclose(g.Pk2D[1:], r.Pk2D[1:], rtol=1e-12, atol=0)


@pytest.mark.parametrize("cols,axis", [(3, 0), (3, 2), (2, 1)])
def test_pos_redshift_space_dropin(gpu, cols, axis):
    """redshift_space_library.pos_redshift_space in place: bit-exact with the float32 oracle, for a
    numpy array (written back) and a device tensor (shifted where it lies)."""
    import torch
    from pylians_amd import redshift_space_library as RSL
    L, H, z = 500.0, 68.0, 2.0
    pos, vel = _case(50000, L, 11 + cols)
    pos, vel = np.ascontiguousarray(pos[:, :cols]), np.ascontiguousarray(vel[:, :cols])
    pos[20, axis] = np.float32(L)        # x = L stays (only x > L wraps)
    vel[20, axis] = np.float32(0.0)
    want = rsd.pos_redshift_space(pos, vel, L, H, z, axis)
    p = pos.copy()
    assert RSL.pos_redshift_space(p, vel, L, H, z, axis) is None
    assert np.array_equal(p, want)
    t = torch.from_numpy(pos).cuda()
    RSL.pos_redshift_space(t, torch.from_numpy(vel).cuda(), L, H, z, axis)
    assert np.array_equal(t.cpu().numpy(), want)
    with pytest.raises(ValueError):
        RSL.pos_redshift_space(pos.astype(np.float64), vel, L, H, z, axis)


def test_field_div_sub_axpy_match_numpy(gpu):
    """The field steps pyl_field_div_sub / pyl_field_axpy (the overdensity step of the snapshot drivers
    and the Omega-weighted sum) round like numpy float32 in-place operations."""
    import torch
    from pylians_amd import _lib

    def div_sub(t, d, s):
        _lib.check(_lib.lib().pyl_field_div_sub(_lib.ptr(t), int(t.numel()), float(d), float(s), _lib.F32,
                                                _lib.stream()))

    def axpy(y, x, a):
        _lib.check(_lib.lib().pyl_field_axpy(_lib.ptr(y), _lib.ptr(x), int(y.numel()), float(a), _lib.F32,
                                             _lib.stream()))

    rng = np.random.default_rng(3)
    a = (rng.random(100003) * 37.0).astype(np.float32)
    b = rng.standard_normal(100003).astype(np.float32)
    mean, om = 8.0 / 3.0 + 1e-3, 0.2718281828
    t = torch.from_numpy(a).cuda()
    div_sub(t, mean, 1.0)
    want = a.copy()
    want /= np.float32(mean)
    want -= np.float32(1.0)
    assert np.array_equal(t.cpu().numpy(), want)
    y = torch.from_numpy(want).cuda()
    axpy(y, torch.from_numpy(b).cuda(), om)
    want2 = want + np.float32(om) * b
    assert np.array_equal(y.cpu().numpy(), want2)


def test_presorted_fused_rsd_matches_binned(gpu):
    """The presort applies the redshift-space shift while grouping: same spectra as the plain binned
    deposit of the same call (the float64 cell sums differ only in summation order)."""
    from pylians_amd import pipeline
    L, H, z = 250.0, 100.0, 0.5
    pos, vel = _case(60000, L, 3)
    W = (np.random.default_rng(3).random(len(pos)) + 0.5).astype(np.float32)
    g = pipeline.PkParticles(pos, L, 64, MAS="CIC", axis=1, W=W, vel=vel, Hubble=H, redshift=z, rsd_axis=1,
                             deposit="presorted")
    r = pipeline.PkParticles(pos, L, 64, MAS="CIC", axis=1, W=W, vel=vel, Hubble=H, redshift=z, rsd_axis=1,
                             deposit="binned")
    assert np.array_equal(g.Nmodes3D, r.Nmodes3D)
    for name in ("Pk", "Pk1D", "Pkphase"):
        assert np.allclose(getattr(g, name), getattr(r, name), rtol=1e-9, atol=0), name
    assert np.allclose(g.Pk2D[1:], r.Pk2D[1:], rtol=1e-9, atol=0)


def test_pos_redshift_space_edges(gpu):
    """No particles; an axis outside the columns; mismatched shapes."""
    import torch
    from pylians_amd import _lib
    from pylians_amd import redshift_space_library as RSL
    RSL.pos_redshift_space(np.zeros((0, 3), np.float32), np.zeros((0, 3), np.float32), 100.0, 70.0, 0.0, 1)
    with pytest.raises(IndexError):
        RSL.pos_redshift_space(np.zeros((4, 2), np.float32), np.zeros((4, 2), np.float32), 100.0, 70.0, 0.0, 2)
    with pytest.raises(ValueError):
        RSL.pos_redshift_space(np.zeros((4, 3), np.float32), np.zeros((5, 3), np.float32), 100.0, 70.0, 0.0, 0)
    p = torch.zeros((4, 3), device="cuda")
    rc = _lib.lib().pyl_pos_redshift_space(_lib.ptr(p), _lib.ptr(p), 4, 3, 100.0, 70.0, 0.0, 3, _lib.stream())
    assert rc == _lib.E_ARG
    rc = _lib.lib().pyl_pos_redshift_space(_lib.ptr(p), _lib.ptr(p), 4, 3, 100.0, 0.0, 0.0, 0, _lib.stream())
    assert rc == _lib.E_ARG
