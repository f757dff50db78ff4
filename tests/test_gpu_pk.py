"""GPU parity of Pk_library.Pk (1-D, 2-D, 3-D, multipoles, phase, Nmodes) against the FP64 oracle."""
import numpy as np
import pytest
import torch

from oracle import cpu as C
from pylians_amd import synth
from tests._gpuref import NOISE_Z, _envelope, assert_pk_close, f32_noise, oracle_pk

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def PKL(gpu):
    from pylians_amd import Pk_library
    return Pk_library


def density_field(N, L, seed, MAS="CIC"):
    pos = synth.zeldovich(N, L, seed=seed, rms_disp=0.02 * L, sigma_v=0.0)
    rho = C.deposit_f64(pos, N, L, MAS)
    return (rho / rho.mean() - 1.0).astype(np.float32)


@pytest.mark.parametrize("N", [16, 32, 64, 96, 128])
@pytest.mark.parametrize("axis", [0, 1, 2])
def test_pk_float32_parity(PKL, N, axis):
    L = 500.0
    delta = density_field(N, L, seed=N + axis)
    g = PKL.Pk(delta, L, axis=axis, MAS="CIC", threads=4)
    o = oracle_pk(delta, L, axis, "CIC")
    errs = assert_pk_close(g, o, rtol=1e-5, noise=f32_noise(delta, L, axis, "CIC"))
    print("N=%d axis=%d max rel err vs FP64 oracle: %s" % (N, axis, errs))


@pytest.mark.parametrize("MAS", ["NGP", "TSC", "PCS", "None"])
def test_pk_windows(PKL, MAS):
    N, L = 64, 250.0
    delta = density_field(N, L, seed=3, MAS=MAS if MAS != "None" else "CIC")
    g = PKL.Pk(delta, L, axis=2, MAS=MAS)
    o = oracle_pk(delta, L, 2, MAS)
    assert_pk_close(g, o, rtol=1e-5, noise=f32_noise(delta, L, 2, MAS))


@pytest.mark.parametrize("N", [32, 64])
def test_pk_float64_path(PKL, N):
    L = 1000.0
    delta = density_field(N, L, seed=N).astype(np.float64)
    g = PKL.Pk(delta, L, axis=0, MAS="TSC")
    o = oracle_pk(delta, L, 0, "TSC")
    assert_pk_close(g, o, rtol=1e-10)


def test_keep_deltak(PKL):
    N, L = 32, 100.0
    delta = density_field(N, L, seed=1)
    g = PKL.Pk(delta, L, axis=1, MAS="PCS", keep_deltak=True)
    o = oracle_pk(delta, L, 1, "PCS", keep=True)
    assert g.delta_k.dtype == np.complex64
    assert np.abs(g.delta_k - o.delta_k).max() <= 1e-5 * np.abs(o.delta_k).max()


def test_device_input_and_odd_free_sizes(PKL):
    N, L = 48, 300.0
    delta = density_field(N, L, seed=4)
    g = PKL.Pk(torch.from_numpy(delta).cuda(), L, axis=0, MAS="CIC")
    o = oracle_pk(delta, L, 0, "CIC")
    assert_pk_close(g, o, rtol=1e-5, noise=f32_noise(delta, L, 0, "CIC"))


def test_plane_wave_closed_form(PKL):
    N, L, n = 32, 200.0, 4
    x = np.arange(N)
    delta = np.ascontiguousarray(np.broadcast_to(np.cos(2 * np.pi * n * x / N)[None, :, None],
                                                 (N, N, N)).astype(np.float32))
    p = PKL.Pk(delta, L, axis=1, MAS="None")
    b = n - 1
    assert p.Pk[b, 0] == pytest.approx(L ** 3 / (4 * p.Nmodes3D[b]), rel=1e-5)
    assert p.Pk[b, 1] == pytest.approx(5 * p.Pk[b, 0], rel=1e-5)
    assert p.Pk[b, 2] == pytest.approx(9 * p.Pk[b, 0], rel=1e-5)
    assert int(p.Nmodes3D.sum()) + 1 == (N ** 3 - 8) // 2 + 8


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_xpk_parity(PKL, axis):
    from oracle import pk as P
    N, L = 32, 300.0
    d1 = density_field(N, L, seed=11, MAS="CIC")
    d2 = density_field(N, L, seed=12, MAS="TSC")
    d3 = (0.5 * d1 + d2).astype(np.float32)
    g = PKL.XPk([d1, d2, d3], L, axis=axis, MAS=["CIC", "TSC", "PCS"])
    o = P.XPk([d1.astype(np.float64), d2.astype(np.float64), d3.astype(np.float64)], L, axis=axis,
              MAS=["CIC", "TSC", "PCS"])
    for name in ("Nmodes3D", "Nmodes1D", "Nmodes2D"):
        assert np.array_equal(getattr(g, name), getattr(o, name)), name
    assert np.allclose(g.k3D, o.k3D, rtol=1e-12) and np.allclose(g.k1D, o.k1D, rtol=1e-13)
    # 1e-5, or NOISE_Z standard deviations of the float32 representation noise of the bin
    nz = C.f32_noise_xpk([d1, d2, d3], L, axis, ["CIC", "TSC", "PCS"])
    n2shape = (len(o.kpar) // (len(o.k1D) + 1), len(o.k1D) + 1)

    def ok(err, std, shape=None):
        return np.all(err <= np.maximum(1e-5, NOISE_Z * _envelope(std, shape)))

    for c in range(3):
        P0 = np.abs(o.Pk[:, 0, c])
        for l in range(3):
            assert ok(np.abs(g.Pk[:, l, c] - o.Pk[:, l, c]) / P0, nz.Pk[:, l, c] / P0), ("auto", c, l)
        den = np.maximum(np.abs(o.Pk1D[:, c]), 1e-30)
        assert ok(np.abs(g.Pk1D[:, c] - o.Pk1D[:, c]) / den, nz.Pk1D[:, c] / den), ("auto 1D", c)
    for x, (a, b) in enumerate([(0, 1), (0, 2), (1, 2)]):
        scale = np.sqrt(np.abs(o.Pk[:, 0, a] * o.Pk[:, 0, b]))
        for l in range(3):
            assert ok(np.abs(g.XPk[:, l, x] - o.XPk[:, l, x]) / scale, nz.XPk[:, l, x] / scale), ("cross", x, l)
        s2 = np.sqrt(np.abs(o.Pk2D[:, a] * o.Pk2D[:, b]))
        s2[0] = 1.0
        e2 = np.abs(g.PkX2D[:, x] - o.PkX2D[:, x]) / s2
        a2 = np.maximum(1e-5, NOISE_Z * _envelope(nz.PkX2D[:, x] / s2, n2shape))
        assert np.all((e2 <= a2)[1:]), ("cross 2D", x)


@pytest.mark.parametrize("axis", [0, 2])
def test_pk_float32_parity_192(PKL, axis):
    """A 3-smooth grid large enough for the radix-3 launch shapes of pass Z and of the binning pass."""
    N, L = 192, 600.0
    delta = density_field(N, L, seed=7 + axis)
    g = PKL.Pk(delta, L, axis=axis, MAS="CIC")
    o = oracle_pk(delta, L, axis, "CIC")
    assert_pk_close(g, o, rtol=1e-5, noise=f32_noise(delta, L, axis, "CIC"))


def test_xpk_five_fields(PKL):
    """Five fields: launch groups whose pairs touch more than four fields (the widest register
    layout of the binning kernel); every auto and cross multipole against the float64 XPk."""
    from oracle import pk as P
    N, L = 32, 200.0
    ds = [density_field(N, L, seed=40 + i, MAS="CIC") for i in range(5)]
    MAS = ["CIC", "TSC", "CIC", "PCS", "NGP"]
    g = PKL.XPk(ds, L, axis=2, MAS=MAS)
    o = P.XPk([d.astype(np.float64) for d in ds], L, axis=2, MAS=MAS)
    nz = C.f32_noise_xpk(ds, L, 2, MAS)
    assert np.array_equal(g.Nmodes3D, o.Nmodes3D)

    def ok(err, std):
        return np.all(err <= np.maximum(1e-5, NOISE_Z * _envelope(std)))

    for c in range(5):
        P0 = np.abs(o.Pk[:, 0, c])
        for l in range(3):
            assert ok(np.abs(g.Pk[:, l, c] - o.Pk[:, l, c]) / P0, nz.Pk[:, l, c] / P0), ("auto", c, l)
    pairs = [(a, b) for a in range(5) for b in range(a + 1, 5)]
    for x, (a, b) in enumerate(pairs):
        scale = np.sqrt(np.abs(o.Pk[:, 0, a] * o.Pk[:, 0, b]))
        for l in range(3):
            assert ok(np.abs(g.XPk[:, l, x] - o.XPk[:, l, x]) / scale, nz.XPk[:, l, x] / scale), ("cross", x, l)


@pytest.mark.parametrize("axis", [0, 2])
@pytest.mark.parametrize("N", [512, 640, 768, 1024])
def test_pk_large_grids_float32_vs_float64(PKL, N, axis):
    """Grids of 512 to 1024 points, line of sight along x and along z (Pk_library.pyx:266-425): where two
    binning blocks per CU would hold 2 lines each (float64 from 512 points, float32 from 640), the
    binning runs one block per CU with 8 or 4 lines; float32 and float64 then mostly take different
    shapes, and the kernel's mode counts are checked on the device against the mode geometry.  On
    white noise without a window the mode counts agree exactly and every
    3-D, phase and 1-D bin to 1e-5; a 2-D bin holds a few modes, some of whose powers lie far below the
    mean, and float32 rounding moves a mode's power by an amount proportional to its amplitude times
    the field's, so 2-D bins are held to 1e-5 of the larger of their own and the mean 2-D power."""
    gen = torch.Generator(device="cuda")
    gen.manual_seed(N)
    d32 = torch.randn((N, N, N), device="cuda", generator=gen)
    p32 = PKL.Pk(d32, 1000.0, axis, "None")
    d64 = d32.double()
    del d32
    p64 = PKL.Pk(d64, 1000.0, axis, "None")
    del d64
    torch.cuda.empty_cache()
    for name in ("Nmodes3D", "Nmodes1D", "Nmodes2D"):
        assert np.array_equal(getattr(p32, name), getattr(p64, name)), name
    P0 = np.abs(p64.Pk[:, 0])
    for l in range(3):
        assert np.max(np.abs(p32.Pk[:, l] - p64.Pk[:, l]) / P0) <= 1e-5, l
    assert np.max(np.abs(p32.Pkphase - p64.Pkphase) / np.abs(p64.Pkphase)) <= 1e-5
    nz = p64.Pk1D != 0
    assert np.max(np.abs(p32.Pk1D[nz] - p64.Pk1D[nz]) / np.abs(p64.Pk1D[nz])) <= 1e-5
    assert np.all(np.abs(p32.Pk1D[~nz]) <= 1e-5 * np.abs(p64.Pk1D).max())
    b = np.abs(p64.Pk2D)
    assert np.all(np.abs(p32.Pk2D - p64.Pk2D) <= 1e-5 * np.maximum(b, b.mean()))
