"""GPU r2c / c2r FFTs (FFT3Dr_f/_d, IFFT3Dr_f/_d drop-ins) against numpy's pocketfft in float64."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

SIZES = [8, 12, 16, 20, 24, 32, 40, 48, 64, 80, 96, 128, 160, 192, 256]


@pytest.fixture(scope="module")
def PKL(gpu):
    from pylians_amd import Pk_library
    return Pk_library


@pytest.mark.parametrize("N", SIZES)
def test_r2c_c2r_float32(PKL, N):
    rng = np.random.default_rng(N)
    a = rng.standard_normal((N, N, N)).astype(np.float32)
    ref = np.fft.rfftn(a.astype(np.float64))
    out = PKL.FFT3Dr_f(a, 1)
    assert out.dtype == np.complex64 and out.shape == (N, N, N // 2 + 1)
    scale = np.abs(ref).max()
    assert np.abs(out - ref).max() <= 2e-6 * scale * np.log2(N)
    back = PKL.IFFT3Dr_f(out, 1)
    assert back.dtype == np.float32 and back.shape == (N, N, N)
    ref_back = np.fft.irfftn(out.astype(np.complex128), s=(N, N, N), axes=(0, 1, 2))
    assert np.abs(back - ref_back).max() <= 2e-6 * np.abs(ref_back).max() * np.log2(N)


@pytest.mark.parametrize("N", SIZES)
def test_r2c_c2r_float64(PKL, N):
    rng = np.random.default_rng(N + 1)
    a = rng.standard_normal((N, N, N))
    ref = np.fft.rfftn(a)
    out = PKL.FFT3Dr_d(a, 1)
    assert np.abs(out - ref).max() <= 1e-13 * np.abs(ref).max() * np.log2(N)
    back = PKL.IFFT3Dr_d(out, 1)
    assert np.abs(back - a).max() <= 1e-13 * np.log2(N) * np.abs(a).max()


def test_plane_wave_known_answer(PKL):
    N, n = 64, 5
    x = np.arange(N)
    a = np.broadcast_to(np.cos(2 * np.pi * n * x / N)[:, None, None], (N, N, N)).astype(np.float32)
    out = PKL.FFT3Dr_f(np.ascontiguousarray(a), 1)
    expect = np.zeros_like(out)
    expect[n, 0, 0] = expect[N - n, 0, 0] = N ** 3 / 2
    assert np.abs(out - expect).max() <= 1e-6 * N ** 3


def test_device_tensors_zero_copy(PKL):
    N = 32
    a = torch.randn(N, N, N, device="cuda")
    out = PKL.FFT3Dr_f(a)
    assert out.is_cuda and out.dtype == torch.complex64
    ref = torch.fft.rfftn(a.double())
    assert (out.to(torch.complex128) - ref).abs().max().item() <= 1e-5 * ref.abs().max().item()
    back = PKL.IFFT3Dr_f(out)
    assert (back - a).abs().max().item() <= 1e-5


def test_unsupported_size_raises(PKL):
    from pylians_amd._lib import PylError
    with pytest.raises(PylError):
        PKL.FFT3Dr_f(np.zeros((14, 14, 14), np.float32))
