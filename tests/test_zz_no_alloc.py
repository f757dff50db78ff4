"""pyl_pk_noalloc (no device allocation; scratch in a fixed per-GPU array) against pyl_pk on the same
fields, through the C ABI: two calls in a row, the second on a smaller grid, float32 and float64,
with and without the corrected spectrum; plus the size limit and its error."""
import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _field(N, seed, dtype):
    g = torch.Generator(device="cpu").manual_seed(seed)
    return torch.randn((N, N, N), generator=g, dtype=torch.float64).to(dtype).cuda()


def _run(fn, delta, N, L, axis, mas_p, prec, keep):
    from pylians_amd import _lib
    from pylians_amd.Pk_library import pk_layout
    _, _, _, nout = pk_layout(N)
    out = torch.full((nout,), float("nan"), dtype=torch.float64, device="cuda")
    dk = None
    if keep:
        ct = torch.complex64 if prec == _lib.F32 else torch.complex128
        dk = torch.empty((N, N, N // 2 + 1), dtype=ct, device="cuda")
    _lib.check(getattr(_lib.lib(), fn)(_lib.ptr(delta), N, L, axis, mas_p, prec, 0, _lib.ptr(out), _lib.ptr(dk),
                                        _lib.stream()))
    torch.cuda.synchronize()
    return out.cpu().numpy(), (None if dk is None else dk.cpu().numpy())


@pytest.mark.parametrize("prec,keep", [(4, False), (8, True)])
def test_noalloc_matches_pyl_pk_twice(gpu, prec, keep):
    from pylians_amd import _lib
    dt = torch.float32 if prec == _lib.F32 else torch.float64
    for N, L, axis, mas_p, seed in ((128, 1000.0, 2, 2, 1), (48, 300.0, 0, 3, 2)):
        d = _field(N, seed, dt)
        a, ka = _run("pyl_pk_noalloc", d, N, L, axis, mas_p, prec, keep)
        b, kb = _run("pyl_pk", d, N, L, axis, mas_p, prec, keep)
        assert np.all(np.isfinite(a))
        # identical kernels and launch shapes: the outputs agree to the float64 summation order of the
        # binning (LDS atomics), and Nmodes (the kernel's own counts) exactly
        from pylians_amd.Pk_library import pk_layout
        _, _, kmax, _ = pk_layout(N)
        assert np.array_equal(a[5 * kmax:6 * kmax], b[5 * kmax:6 * kmax])
        np.testing.assert_allclose(a, b, rtol=1e-12, atol=0)
        if keep:
            assert np.array_equal(ka, kb)


def test_noalloc_rejects_large_grid(gpu):
    from pylians_amd import _lib
    d = _field(160, 3, torch.float32)
    out = torch.zeros(10, dtype=torch.float64, device="cuda")
    rc = _lib.lib().pyl_pk_noalloc(_lib.ptr(d), 160, 100.0, 0, 2, _lib.F32, 0, _lib.ptr(out), None, _lib.stream())
    assert rc == _lib.E_UNSUPPORTED
    assert b"exceeds" in _lib.lib().pyl_last_error()
