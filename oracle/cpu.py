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
        _lib.oracle_deposit_f64_par.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_long, ctypes.c_int,
                                                ctypes.c_double, ctypes.c_int, ctypes.c_double, ctypes.c_void_p,
                                                ctypes.c_int]
        for f in (_lib.oracle_pkloop_f64, _lib.oracle_pkloop_f32ref):
            f.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int] + [ctypes.c_void_p] * 9
        _lib.oracle_pkloop_var.argtypes = ([ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int]
                                           + [ctypes.c_void_p] * 4 + [ctypes.c_int])
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


def host_threads():
    """Threads for the oracle's parallel loops: this process's CPU share, at most 16."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def deposit_f64(pos, dims, BoxSize, scheme, W=None, multiplicity=1.0, threads=None):
    """FP64 oracle deposit.  3-D inputs of more than 2^20 particles use the plane-split parallel
    loop, whose result is bit-identical to the serial one."""
    pos = np.ascontiguousarray(pos, dtype=np.float32)
    n, axes = pos.shape
    grid = np.zeros((dims,) * axes, dtype=np.float64)
    Wc = None if W is None else np.ascontiguousarray(W, dtype=np.float32)
    threads = host_threads() if threads is None else int(threads)
    if axes == 3 and n > (1 << 20) and threads > 1:
        lib().oracle_deposit_f64_par(_p(pos), None if Wc is None else _p(Wc), n, dims, float(BoxSize),
                                     _SCHEME[scheme], float(multiplicity), _p(grid), threads)
        return grid
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


def fp64_oracle_pipeline(pos, dims, BoxSize, axis, MAS="CIC", workers=8, with_floor=False, with_noise=False,
                         with_sigma=False):
    """FP64 oracle of the end-to-end path at full size (C deposit, float64 r2c, C float64 mode loop).
    With with_floor, also returns the float32 floor: the exact overdensity rounded once to float32,
    transformed with a float32 FFT (pocketfft single precision) and binned in float64.  With
    with_noise (small grids), returns ``f32_noise_pk`` of the exact overdensity instead; with
    with_sigma (any size), ``f32_sigma_pk``, its first-order closed form."""
    import scipy.fft

    rho = deposit_f64(pos, dims, BoxSize, MAS)
    rho /= len(pos) * 1.0 / dims ** 3
    rho -= 1.0
    if with_noise:
        return (_pk.Pk(None, BoxSize, sums=pk_sums_c(np.fft.rfftn(rho), dims, axis, MAS), dims=dims),
                f32_noise_pk(rho, BoxSize, axis, MAS))
    if with_sigma:
        d32 = rho.astype(np.float32)
        dk = scipy.fft.rfftn(rho, workers=workers)
        del rho
        sums = pk_sums_c(dk, dims, axis, MAS)
        o = _pk.Pk(None, BoxSize, sums=sums, dims=dims)
        return o, f32_sigma_pk(dk, d32, BoxSize, axis, MAS, sums, workers=workers)
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


class _Std:
    pass


def f32_noise_pk(delta, BoxSize, axis, MAS="CIC", samples=16, seed=0):
    """Per-bin standard deviation of the Pk outputs caused by float32 representation noise.

    A float32 pipeline (the exact overdensity rounded once to float32, a float32 r2c) adds to every
    mode delta_k an error e_k that is white: the mean |e_k|^2 = sigma^2 is measured here with
    pocketfft in single precision on this field (standing in for the reference's FFTW float plan,
    Pk_library.pyx:120-133).  Real white noise of variance sigma^2 / N^3 per cell has exactly that
    spectrum, with the Hermitian symmetry of a real field, so adding `samples` draws of it to the
    exact spectrum and binning in float64 gives the spread the float32 representation causes in every
    output bin.  Returned attributes Pk (kmax, 3), Pk1D, Pk2D, Pkphase are RMS deviations from the
    noise-free bins; sigma_rel = sigma / rms |delta_k|.
    """
    import scipy.fft

    d = np.asarray(delta, dtype=np.float64)
    N = d.shape[0]
    ex = np.fft.rfftn(d)
    fl = scipy.fft.rfftn(d.astype(np.float32)).astype(np.complex128)
    sigma2 = float(np.mean(np.abs(fl - ex) ** 2))
    del fl
    base = _pk.Pk(None, BoxSize, sums=pk_sums_c(ex, N, axis, MAS), dims=N)
    rng = np.random.default_rng(seed)
    names = ("Pk", "Pk1D", "Pk2D", "Pkphase")
    acc = {k: np.zeros_like(getattr(base, k)) for k in names}
    s = np.sqrt(sigma2 / N ** 3)
    for _ in range(samples):
        dk = ex + np.fft.rfftn(rng.normal(0.0, s, d.shape))
        X = _pk.Pk(None, BoxSize, sums=pk_sums_c(dk, N, axis, MAS), dims=N)
        for k in names:
            acc[k] += (getattr(X, k) - getattr(base, k)) ** 2
    out = _Std()
    for k in names:
        setattr(out, k, np.sqrt(acc[k] / samples))
    out.sigma_rel = float(np.sqrt(sigma2 / np.mean(np.abs(ex) ** 2)))
    return out


def f32_sigma_pk(dk, d32, BoxSize, axis, MAS, sums, workers=8):
    """First-order closed form of ``f32_noise_pk``: per-bin standard deviation of the Pk outputs under
    white noise e_k on the raw spectrum with the mean |e_k|^2 = sigma^2 that a float32 pipeline adds
    (measured: ``d32``, the exact overdensity rounded once to float32, through pocketfft in single
    precision, against ``dk``, the float64 spectrum of the exact field).  To first order a mode's
    |C (delta_k + e_k)|^2 moves by 2 C^2 Re(conj(delta_k) e_k), of variance 2 C^2 |C delta_k|^2 sigma^2
    (C: window correction), so each bin's variance is sigma^2 times a weighted mode sum
    (liboracle ``oracle_pkloop_var``), and the normalisation of Pk_library.pyx:387-421 is linear in the
    sums.  ``sums``: the float64 mode-loop sums of ``dk`` (for Nmodes and k)."""
    import scipy.fft

    N = dk.shape[0]
    fl = scipy.fft.rfftn(d32, workers=workers)
    sigma2 = 0.0
    for i in range(0, N, 16):   # plane blocks: no full-size temporaries
        sigma2 += float(np.sum(np.abs(fl[i:i + 16].astype(np.complex128) - dk[i:i + 16]) ** 2))
    sigma2 /= dk.size
    del fl
    dkc = np.ascontiguousarray(dk, dtype=np.complex128)
    _, _, kmax_par, kmax_per, kmax = _pk.frequencies(1.0, N)
    V1 = np.zeros(kmax_par + 1)
    V2 = np.zeros((kmax_par + 1) * (kmax_per + 1))
    V3 = np.zeros((kmax + 1, 3))
    Vp = np.zeros(kmax + 1)
    lib().oracle_pkloop_var(_p(dkc), N, axis, _pk.MAS_function(MAS), _p(V1), _p(V2), _p(V3), _p(Vp),
                            int(workers))
    sd = dict(sums)
    sd["Pk1D"] = np.sqrt(sigma2 * V1)
    sd["Pk2D"] = np.sqrt(sigma2 * V2)
    sd["P0"], sd["P2"], sd["P4"] = (np.sqrt(sigma2 * V3[:, i]) for i in range(3))
    sd["Pkphase"] = np.sqrt(sigma2 * Vp)
    X = _pk.Pk(None, BoxSize, sums=sd, dims=N)
    out = _Std()
    for k in ("Pk", "Pk1D", "Pk2D", "Pkphase"):
        setattr(out, k, getattr(X, k))
    out.sigma_rel = float(np.sqrt(sigma2 / np.mean(np.abs(dk) ** 2)))
    return out


def f32_noise_xpk(deltas, BoxSize, axis, MAS, samples=16, seed=0):
    """``f32_noise_pk`` for ``XPk``: each field gets real white noise of the variance its float32
    r2c adds (pocketfft single precision), and the float64 XPk restatement is re-run on the noisy
    fields.  Returns RMS deviations Pk (kmax, 3, F), XPk (kmax, 3, X), Pk1D (kmax_par, F), Pk2D (n2, F)
    and PkX2D (n2, X)."""
    import scipy.fft

    ds = [np.asarray(d, dtype=np.float64) for d in deltas]
    N = ds[0].shape[0]
    sd = []
    for d in ds:
        ex = np.fft.rfftn(d)
        fl = scipy.fft.rfftn(d.astype(np.float32)).astype(np.complex128)
        sd.append(np.sqrt(float(np.mean(np.abs(fl - ex) ** 2)) / N ** 3))
    base = _pk.XPk(ds, BoxSize, axis=axis, MAS=MAS)
    rng = np.random.default_rng(seed)
    names = ("Pk", "XPk", "Pk1D", "Pk2D", "PkX2D")
    acc = {k: np.zeros_like(getattr(base, k)) for k in names}
    for _ in range(samples):
        X = _pk.XPk([d + rng.normal(0.0, s, d.shape) for d, s in zip(ds, sd)], BoxSize, axis=axis, MAS=MAS)
        for k in names:
            acc[k] += (getattr(X, k) - getattr(base, k)) ** 2
    out = _Std()
    for k in names:
        setattr(out, k, np.sqrt(acc[k] / samples))
    return out
