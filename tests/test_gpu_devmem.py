"""Device-memory entry points of the C ABI (no counterpart in the reference; they serve the numpy
drop-in path): pyl_dev_alloc / pyl_dev_free, pyl_memcpy_h2d / pyl_memcpy_d2h through the library's
pinned staging buffers (3 x 32 MiB), pyl_memcpy_d2d, pyl_dev_zero.  Round trips are bit-exact for
sizes below, at and across the staging chunk, more chunks than buffers, and unaligned host addresses;
copies queued on a stream stay ordered with the kernels on it."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MiB = 1 << 20


def _lib():
    from pylians_amd import _lib
    return _lib


def _alloc(nbytes):
    L = _lib()
    p = ctypes.c_void_p()
    L.check(L.lib().pyl_dev_alloc(ctypes.byref(p), nbytes))
    return p


@pytest.mark.parametrize("nbytes,offset", [(1, 0), (1000, 3), (32 * MiB - 5, 1), (32 * MiB, 0),
                                           (100 * MiB + 7, 5), (200 * MiB + 12, 0)])
def test_staged_round_trip(gpu, nbytes, offset):
    L = _lib()
    rng = np.random.default_rng(nbytes)
    host = np.frombuffer(rng.bytes(nbytes + offset), dtype=np.uint8)
    src = host[offset:]                      # an unaligned host address when offset > 0
    d = _alloc(nbytes)
    try:
        print("\nround trip of %d bytes (host offset %d)" % (nbytes, offset), flush=True)
        L.check(L.lib().pyl_memcpy_h2d(d, src.ctypes.data_as(ctypes.c_void_p), nbytes, None))
        back = np.empty(nbytes + offset, np.uint8)[offset:]
        L.check(L.lib().pyl_memcpy_d2h(back.ctypes.data_as(ctypes.c_void_p), d, nbytes, None))
        assert np.array_equal(back, src)
        # d2d and zero
        d2 = _alloc(nbytes)
        try:
            L.check(L.lib().pyl_memcpy_d2d(d2, d, nbytes, None))
            L.check(L.lib().pyl_dev_zero(d, nbytes, None))
            z = np.ones(nbytes, np.uint8)
            L.check(L.lib().pyl_memcpy_d2h(z.ctypes.data_as(ctypes.c_void_p), d, nbytes, None))
            assert not z.any()
            L.check(L.lib().pyl_memcpy_d2h(z.ctypes.data_as(ctypes.c_void_p), d2, nbytes, None))
            assert np.array_equal(z, src)
        finally:
            L.check(L.lib().pyl_dev_free(d2))
    finally:
        L.check(L.lib().pyl_dev_free(d))


def test_staged_copies_follow_stream_order(gpu):
    """h2d into a tensor, a torch kernel on the same stream, d2h: each step sees the previous one, and
    the host buffer of the h2d may be overwritten as soon as the call returns."""
    import torch
    L = _lib()
    n = 3 * 32 * MiB // 4 + 1001                 # 96 MiB + 4004 bytes of float32: four chunks
    a = np.arange(n, dtype=np.float32)
    t = torch.empty(n, dtype=torch.float32, device="cuda")
    st = L.stream()
    L.check(L.lib().pyl_memcpy_h2d(L.ptr(t), a.ctypes.data_as(ctypes.c_void_p), a.nbytes, st))
    a[:] = -1.0                                   # the call returned: the source may be reused
    t.mul_(2.0)
    out = np.empty(n, np.float32)
    L.check(L.lib().pyl_memcpy_d2h(out.ctypes.data_as(ctypes.c_void_p), L.ptr(t), out.nbytes, st))
    assert np.array_equal(out, 2.0 * np.arange(n, dtype=np.float32))
    # the Python helpers take the same path
    b = np.random.default_rng(0).random((1000, 3)).astype(np.float32)
    assert np.array_equal(L.to_host(L.to_dev(b, np.float32)), b)
