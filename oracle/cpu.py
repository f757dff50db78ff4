"""ctypes front-end of the C checkers (TEST INFRASTRUCTURE ONLY).

* ``deposit_f64``  -> liboracle.so ``oracle_deposit_f64`` (float64 stencils, MAS_library.pyx:123-545)
* ``pk_sums_c``    -> liboracle.so ``oracle_pkloop_f64`` / ``oracle_pkloop_f32ref``
  (mode loop, Pk_library.pyx:314-381)
* ``ref_mas_c``    -> oracle/_ref/libmas_c_ref.so, the reference's MAS_c.c compiled from
  source (MAS_c.c:8-266, called as MAS_library.pyx:1136-1221 does)
"""
import ctypes
import os

import numpy as np

from . import pk as _pk

_HERE = os.path.dirname(os.path.abspath(__file__))
_lib = None
_ref = None
_SCHEME = {"NGP": 1, "CIC": 2, "TSC": 3, "PCS": 4}


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so missing: run `make -C oracle`")
        _lib = ctypes.CDLL(path)
        _lib.oracle_deposit_f64.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_long,
                                            ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                            ctypes.c_int, ctypes.c_double, ctypes.c_void_p]
        for f in (_lib.oracle_pkloop_f64, _lib.oracle_pkloop_f32ref):
            f.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int] + [ctypes.c_void_p] * 9
    return _lib


def ref_available():
    return os.path.exists(os.path.join(_HERE, "_ref", "libmas_c_ref.so"))


def ref():
    global _ref
    if _ref is None:
        path = os.path.join(_HERE, "_ref", "libmas_c_ref.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/_ref/libmas_c_ref.so missing: run `make -C oracle ref`")
        _ref = ctypes.CDLL(path)
        for name in ("NGP", "CIC", "TSC", "PCS"):
            getattr(_ref, name).argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                            ctypes.c_long, ctypes.c_int, ctypes.c_int,
                                            ctypes.c_float, ctypes.c_int]
    return _ref


def deposit_f64(pos, dims, BoxSize, scheme, W=None, multiplicity=1.0):
    pos = np.ascontiguousarray(pos, dtype=np.float32)
    n, axes = pos.shape
    grid = np.zeros((dims,) * axes, dtype=np.float64)
    Wc = None if W is None else np.ascontiguousarray(W, dtype=np.float32)
    lib().oracle_deposit_f64(_p(pos), None if Wc is None else _p(Wc), n, axes, dims,
                             float(BoxSize), _SCHEME[scheme], float(multiplicity), _p(grid))
    return grid


def pk_sums_c(delta_k, dims, axis, MAS, faithful_f32=False):
    """Raw histograms of the mode loop (same keys as oracle.pk.pk_sums)."""
    middle = dims // 2
    _, _, kmax_par, kmax_per, kmax = _pk.frequencies(1.0, dims)
    if faithful_f32:
        dk = np.ascontiguousarray(delta_k, dtype=np.complex64)
        fn = lib().oracle_pkloop_f32ref
    else:
        dk = np.ascontiguousarray(delta_k, dtype=np.complex128)
        fn = lib().oracle_pkloop_f64
    assert dk.shape == (dims, dims, middle + 1)
    n1, n2, n3 = kmax_par + 1, (kmax_par + 1) * (kmax_per + 1), kmax + 1
    o = {k: np.zeros(n) for k, n in (("k1D", n1), ("Pk1D", n1), ("Nmodes1D", n1), ("Pk2D", n2),
                                     ("Nmodes2D", n2), ("k3D", n3), ("Pkphase", n3), ("Nmodes3D", n3))}
    P3 = np.zeros((n3, 3))
    fn(_p(dk), dims, axis, _pk.MAS_function(MAS), _p(o["k1D"]), _p(o["Pk1D"]), _p(o["Nmodes1D"]),
       _p(o["Pk2D"]), _p(o["Nmodes2D"]), _p(o["k3D"]), _p(P3), _p(o["Pkphase"]), _p(o["Nmodes3D"]))
    o["P0"], o["P2"], o["P4"] = P3[:, 0].copy(), P3[:, 1].copy(), P3[:, 2].copy()
    return o


def ref_mas_c(pos, number, BoxSize, MAS="CIC", W=None, threads=1):
    """Run the reference MAS_c.c stencil in place on a float32 grid (MAS_library.pyx:1136-1221)."""
    pos = np.ascontiguousarray(pos, dtype=np.float32)
    assert number.dtype == np.float32 and number.flags.c_contiguous
    Wc = None if W is None else np.ascontiguousarray(W, dtype=np.float32)
    getattr(ref(), MAS)(_p(pos), _p(number), None if Wc is None else _p(Wc), pos.shape[0],
                        number.shape[0], pos.shape[1], float(BoxSize), int(threads))
    return number


def cpu_reference_pipeline(pos, dims, BoxSize, axis, threads, MAS="CIC"):
    """The reference's CPU path on this host, timed stage by stage (seconds):
    MAS_c.c deposit with OpenMP `threads` (the reference's own C, oracle/_ref), overdensity
    normalisation in float32 (Pk_snapshot.py:84-88), float32 r2c with scipy.fft (pocketfft, in
    place of pyFFTW/FFTW, which this image lacks) on `threads` workers, and the serial mode loop
    restated in C with the reference's float32 window multiply (Pk_library.pyx:314-381).
    Returns (oracle.pk.Pk-like result, {stage: seconds})."""
    import time

    import scipy.fft

    t = {}
    t0 = time.perf_counter()
    grid = np.zeros((dims, dims, dims), dtype=np.float32)
    ref_mas_c(pos, grid, BoxSize, MAS, None, threads)
    t1 = time.perf_counter()
    mean = len(pos) * 1.0 / dims ** 3
    grid /= np.float32(mean)
    grid -= np.float32(1.0)
    dk = scipy.fft.rfftn(grid, workers=threads)
    t2 = time.perf_counter()
    sums = pk_sums_c(dk, dims, axis, MAS, faithful_f32=True)
    res = _pk.Pk(None, BoxSize, sums=sums, dims=dims)
    t3 = time.perf_counter()
    t["deposit"], t["normalise_fft"], t["modes"], t["total"] = t1 - t0, t2 - t1, t3 - t2, t3 - t0
    return res, t


def fp64_oracle_pipeline(pos, dims, BoxSize, axis, MAS="CIC", workers=8, with_floor=False):
    """FP64 oracle of the end-to-end path at full size (C deposit, float64 r2c, C float64 mode loop).
    With with_floor, also returns the float32 floor: the exact overdensity rounded once to float32,
    transformed with a float32 FFT (pocketfft single precision) and binned in float64."""
    import scipy.fft

    rho = deposit_f64(pos, dims, BoxSize, MAS)
    rho /= len(pos) * 1.0 / dims ** 3
    rho -= 1.0
    d32 = rho.astype(np.float32) if with_floor else None
    dk = scipy.fft.rfftn(rho, workers=workers)
    del rho
    o = _pk.Pk(None, BoxSize, sums=pk_sums_c(dk, dims, axis, MAS), dims=dims)
    if not with_floor:
        return o
    del dk
    dk = scipy.fft.rfftn(d32, workers=workers).astype(np.complex128)
    del d32
    return o, _pk.Pk(None, BoxSize, sums=pk_sums_c(dk, dims, axis, MAS), dims=dims)
