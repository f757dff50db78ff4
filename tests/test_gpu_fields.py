"""GPU parity of correct_MAS, Xi, FT_filter and field_smoothing against the float64 oracle
(oracle/fields.py).  float32 fields: per cell / per bin within 1e-5 of the largest value (the
float32 FFT's rounding is ~1e-7 of it); float64 fields: 1e-10."""
import numpy as np
import pytest

from oracle import fields as F
from pylians_amd import synth

pytestmark = pytest.mark.gpu


def _field(N, L, seed, MAS="CIC"):
    from oracle import cpu as C
    pos = synth.zeldovich(N, L, seed=seed, rms_disp=0.02 * L, sigma_v=0.0)
    rho = C.deposit_f64(pos, N, L, MAS)
    return (rho / rho.mean() - 1.0).astype(np.float32)


@pytest.mark.parametrize("N,MAS", [(32, "CIC"), (64, "TSC"), (48, "PCS")])
def test_correct_mas(gpu, N, MAS):
    from pylians_amd import Pk_library as PKL
    L = 300.0
    d = _field(N, L, N, MAS)
    g = PKL.correct_MAS(d, L, MAS)
    o = F.correct_MAS(d, L, MAS)
    assert g.dtype == np.float32
    assert np.max(np.abs(g - o)) <= 1e-5 * np.max(np.abs(o))
    g64 = PKL.correct_MAS(d.astype(np.float64), L, MAS)
    assert np.max(np.abs(g64 - o)) <= 1e-10 * np.max(np.abs(o))


def test_correct_mas_dependent_half_plane(gpu):
    from pylians_amd import Pk_library as PKL
    N, n = 16, 3
    x = np.arange(N)
    W = (np.pi * n / N / np.sin(np.pi * n / N)) ** 2
    wx = (np.cos(2 * np.pi * n * x / N)[:, None, None] * np.ones((1, N, N))).astype(np.float64)
    assert np.allclose(PKL.correct_MAS(wx, 100.0, "CIC"), (W + 1) / 2 * wx, atol=1e-12)


@pytest.mark.parametrize("N,axis,MAS", [(32, 2, "CIC"), (64, 0, "TSC"), (40, 1, "None")])
def test_xi(gpu, N, axis, MAS):
    from pylians_amd import Pk_library as PKL
    L = 250.0
    d = _field(N, L, N + axis, MAS if MAS != "None" else "CIC")
    g = PKL.Xi(d, L, MAS=MAS, axis=axis)
    o = F.Xi(d, L, MAS=MAS, axis=axis)
    assert np.array_equal(g.Nmodes3D, o.Nmodes3D)
    assert np.allclose(g.r3D, o.r3D, rtol=1e-12, atol=0)
    scale = np.max(np.abs(o.xi[:, 0]))
    assert np.max(np.abs(g.xi - o.xi)) <= 1e-5 * scale
    g64 = PKL.Xi(d.astype(np.float64), L, MAS=MAS, axis=axis)
    assert np.max(np.abs(g64.xi - o.xi)) <= 1e-10 * scale


@pytest.mark.parametrize("Filter", ["Top-Hat", "Gaussian"])
def test_filter_and_smoothing(gpu, Filter):
    from pylians_amd import smoothing_library as SL
    N, L, R = 64, 200.0, 12.0
    gk = SL.FT_filter(L, R, N, Filter)
    ok = F.FT_filter(L, R, N, Filter)
    assert gk.dtype == np.complex64 and gk.shape == (N, N, N // 2 + 1)
    assert np.max(np.abs(gk - ok)) <= 1e-6
    d = _field(N, L, 7)
    gs = SL.field_smoothing(d, gk)
    os_ = F.field_smoothing(d, ok)
    assert np.max(np.abs(gs - os_)) <= 1e-5 * np.max(np.abs(os_))
    with pytest.raises(Exception):
        SL.FT_filter(L, R, N, "Box")
