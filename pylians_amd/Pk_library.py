"""Drop-in replacement of Pylians' ``Pk_library`` power-spectrum path on MI355X.

Same call surface as the reference (Pk_library.pyx, bispectrum_library.pyx; the package
``__init__`` re-exports the bispectrum, Pk_library/__init__.py:1-4):

* ``frequencies``, ``frequencies_2D``, ``MAS_function``, ``MAS_correction``,
  ``check_number_modes`` (Pk_library.pyx:59-117);
* ``FFT3Dr_f / FFT3Dr_d`` (unnormalised r2c) and ``IFFT3Dr_f / IFFT3Dr_d`` (1/N^3-normalised
  c2r), Pk_library.pyx:120-181, on hand-written gfx950 Stockham FFTs;
* ``Pk(delta, BoxSize, axis=2, MAS='CIC', threads=1, keep_deltak=False)`` with the attributes
  k1D, Pk1D, Nmodes1D, kpar, kper, Pk2D, Nmodes2D, k3D, Pk, Pkphase, Nmodes3D [, delta_k]
  (Pk_library.pyx:266-425);
* ``Bk(delta, BoxSize, k1, k2, theta, MAS='CIC', threads=1)`` with B, Q, k, Pk
  (bispectrum_library.pyx:32-196).

Inputs may be numpy arrays or torch tensors on the GPU.  A float32 field runs the float32
path (float32 FFT, float64 shell sums, as the reference); a float64 field runs the float64
path.  ``threads`` is accepted for signature parity and ignored.
"""
import sys
import time

import numpy as np

from . import _lib

VERBOSE = False
# extra pyl_pk flags OR-ed into every Pk call; the tests of the mode check set PK_DROP_KZ0 here
_EXTRA_FLAGS = 0

__all__ = ["frequencies", "frequencies_2D", "MAS_function", "MAS_correction", "check_number_modes",
           "check_number_modes_2D", "FFT3Dr_f", "FFT3Dr_d", "IFFT3Dr_f", "IFFT3Dr_d", "Pk", "XPk", "Bk",
           "pk_layout"]


def frequencies(BoxSize, dims):
    """kF, kN, kmax_par, kmax_per, kmax (Pk_library.pyx:59-64)."""
    kF = 2.0 * np.pi / BoxSize
    middle = dims // 2
    kN = middle * kF
    kmax_par = middle
    kmax_per = int(np.sqrt(middle ** 2 + middle ** 2))
    kmax = int(np.sqrt(middle ** 2 + middle ** 2 + middle ** 2))
    return kF, kN, kmax_par, kmax_per, kmax


def frequencies_2D(BoxSize, dims):
    """2-D variant (Pk_library.pyx:67-72)."""
    kF = 2.0 * np.pi / BoxSize
    middle = dims // 2
    kN = middle * kF
    return kF, kN, middle, middle, int(np.sqrt(middle ** 2 + middle ** 2))


def MAS_function(MAS):
    """Window exponent: 1/2/3/4 for NGP/CIC/TSC/PCS, 0 otherwise (Pk_library.pyx:75-81)."""
    return {"NGP": 1, "CIC": 2, "TSC": 3, "PCS": 4}.get(MAS, 0)


def MAS_correction(x, MAS_index):
    """(x/sin x)^p, 1 at x == 0 (Pk_library.pyx:86-87)."""
    return 1.0 if x == 0.0 else (x / np.sin(x)) ** MAS_index


def check_number_modes(Nmodes, dims):
    """Exit when not all independent modes were counted (Pk_library.pyx:90-102)."""
    own = 1 if dims % 2 == 1 else 8
    indep = (dims ** 3 - own) // 2 + own
    if int(np.sum(Nmodes)) != indep:
        print("WARNING: Not all modes counted")
        print("Counted  %d independent modes" % int(np.sum(Nmodes)))
        print("Expected %d independent modes" % indep)
        sys.exit()


def check_number_modes_2D(Nmodes, dims):
    """2-D variant (Pk_library.pyx:105-117)."""
    own = 1 if dims % 2 == 1 else 4
    indep = (dims ** 2 - own) // 2 + own
    if int(np.sum(Nmodes)) != indep:
        print("WARNING: Not all modes counted")
        sys.exit()


def _modes_exit(rc):
    """check_number_modes' report and exit (Pk_library.pyx:98-102) for a PYL_E_MODES return."""
    import re
    msg = _lib.lib().pyl_last_error().decode(errors="replace")
    m = re.search(r"counted (\d+) independent modes, expected (\d+)", msg)
    print("WARNING: Not all modes counted")
    if m:
        print("Counted  %d independent modes" % int(m.group(1)))
        print("Expected %d independent modes" % int(m.group(2)))
    sys.exit()


def _grid3(a, dtype, name):
    if a.ndim != 3 or not (a.shape[0] == a.shape[1] == a.shape[2]):
        raise ValueError("%s must be a (dims, dims, dims) array" % name)
    return _lib.to_dev(a, dtype)


def _ret(like, t):
    """numpy result for numpy input, device tensor for device input."""
    return t if _lib.is_dev(like) else _lib.to_host(t)


def _fft_r2c(a, real_t, cplx_t, prec):
    d = _grid3(a, real_t, "a")
    N = int(d.shape[0])
    out = _lib.empty((N, N, N // 2 + 1), cplx_t)
    _lib.check(_lib.lib().pyl_fft_r2c(_lib.ptr(d), _lib.ptr(out), N, prec, _lib.stream()))
    return _ret(a, out)


def _fft_c2r(a, real_t, cplx_t, prec):
    d = _lib.to_dev(a, cplx_t)
    N = int(d.shape[0])
    if d.ndim != 3 or d.shape[1] != N or d.shape[2] != N // 2 + 1:
        raise ValueError("a must be a (dims, dims, dims//2+1) complex array")
    out = _lib.empty((N, N, N), real_t)
    _lib.check(_lib.lib().pyl_fft_c2r(_lib.ptr(d), _lib.ptr(out), N, prec, 1, _lib.stream()))
    return _ret(a, out)


def FFT3Dr_f(a, threads=1):
    """Unnormalised r2c FFT of a float32 (N,N,N) field -> complex64 (N,N,N//2+1) (Pk_library.pyx:120-133)."""
    return _fft_r2c(a, _lib.F32T, _lib.C64T, _lib.F32)


def FFT3Dr_d(a, threads=1):
    """float64 version (Pk_library.pyx:136-149)."""
    return _fft_r2c(a, _lib.F64T, _lib.C128T, _lib.F64)


def IFFT3Dr_f(a, threads=1):
    """1/N^3-normalised c2r FFT complex64 (N,N,N//2+1) -> float32 (N,N,N) (Pk_library.pyx:152-165)."""
    return _fft_c2r(a, _lib.F32T, _lib.C64T, _lib.F32)


def IFFT3Dr_d(a, threads=1):
    """float64 version (Pk_library.pyx:168-181)."""
    return _fft_c2r(a, _lib.F64T, _lib.C128T, _lib.F64)


def pk_layout(dims):
    """(kmax_par, kmax_per, kmax, out_len) of the pyl_pk output vector."""
    import ctypes
    a, b, c, n = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_int64()
    _lib.check(_lib.lib().pyl_pk_layout(int(dims), ctypes.byref(a), ctypes.byref(b), ctypes.byref(c),
                                        ctypes.byref(n)))
    return a.value, b.value, c.value, n.value


def unpack_pk(obj, out, dims):
    """Set the Pk attributes from the pyl_pk output vector (layout in include/pylians_hip.h)."""
    kmax_par, kmax_per, kmax, _ = pk_layout(dims)
    n2 = (kmax_par + 1) * (kmax_per + 1)
    o = 0

    def take(n):
        nonlocal o
        v = out[o:o + n]
        o += n
        return v

    obj.k3D = take(kmax)
    P0, P2, P4 = take(kmax), take(kmax), take(kmax)
    obj.Pk = np.stack([P0, P2, P4], axis=1)
    obj.Pkphase = take(kmax)
    obj.Nmodes3D = take(kmax)
    obj.k1D, obj.Pk1D, obj.Nmodes1D = take(kmax_par), take(kmax_par), take(kmax_par)
    obj.kpar, obj.kper, obj.Pk2D, obj.Nmodes2D = take(n2), take(n2), take(n2), take(n2)
    for name in ("Nmodes1D", "Nmodes2D", "Nmodes3D"):
        if np.any(getattr(obj, name) == 0):
            raise ZeroDivisionError("float division (empty %s bin)" % name)


class Pk:
    """Power spectrum of a density field (Pk_library.pyx:266-425), computed on the GPU."""

    def __init__(self, delta, BoxSize, axis=2, MAS="CIC", threads=1, keep_deltak=False):
        start = time.time()
        if VERBOSE:
            print("\nComputing power spectrum of the field...")
        f64 = _lib.dtype_of(delta) == np.float64
        real_t, cplx_t = (_lib.F64T, _lib.C128T) if f64 else (_lib.F32T, _lib.C64T)
        d = _grid3(delta, real_t, "delta")
        dims = int(d.shape[0])
        _, _, _, nout = pk_layout(dims)
        out = _lib.empty((nout,), _lib.F64T)
        dk = _lib.empty((dims, dims, dims // 2 + 1), cplx_t) if keep_deltak else None
        rc = _lib.lib().pyl_pk(_lib.ptr(d), dims, float(BoxSize), int(axis), MAS_function(MAS),
                               _lib.F64 if f64 else _lib.F32, _lib.PK_ALL | _EXTRA_FLAGS, _lib.ptr(out),
                               _lib.ptr(dk), _lib.stream())
        if rc == _lib.E_MODES:
            _modes_exit(rc)
        _lib.check(rc)
        unpack_pk(self, _lib.to_host(out), dims)
        if keep_deltak:
            self.delta_k = _ret(delta, dk)
        if VERBOSE:
            print("Time taken = %.2f seconds" % (time.time() - start))


class XPk:
    """Auto and cross power spectra of several fields (Pk_library.pyx:534-798), on the GPU.

    Attributes as the reference: k1D, Pk1D (kmax_par, F), PkX1D (kmax_par, X), Nmodes1D, kpar, kper,
    Pk2D (n2, F), PkX2D (n2, X), Nmodes2D, k3D, Pk (kmax, 3, F), XPk (kmax, 3, X), Nmodes3D with
    X = F(F-1)/2 crosses ordered (0,1), (0,2), ..., (1,2), ...
    """

    def __init__(self, delta, BoxSize, axis=2, MAS=None, threads=1):
        import ctypes
        fields = len(delta)
        dims = int(delta[0].shape[0])
        for d in delta[1:]:
            if int(d.shape[0]) != dims:
                print("Fields have different grid sizes!!!")
                sys.exit()
        f64 = all(_lib.dtype_of(x if hasattr(x, "dtype") else np.asarray(x)) == np.float64 for x in delta)
        rt = _lib.F64T if f64 else _lib.F32T
        devs = [_grid3(x, rt, "delta") for x in delta]
        ptrs = (ctypes.c_void_p * fields)(*[_lib.ptr(t).value for t in devs])
        mas = (ctypes.c_int * fields)(*[MAS_function(m) for m in MAS])
        n = ctypes.c_int64()
        _lib.check(_lib.lib().pyl_xpk_layout(dims, fields, ctypes.byref(n)))
        out = np.empty(n.value)   # pyl_xpk writes every value of the layout
        rc = _lib.lib().pyl_xpk(ptrs, fields, dims, float(BoxSize), int(axis), mas, _lib.F64 if f64 else _lib.F32,
                                out.ctypes.data_as(ctypes.c_void_p), _lib.stream())
        if rc == _lib.E_MODES:
            _modes_exit(rc)
        _lib.check(rc)
        kmax_par, kmax_per, kmax, _ = pk_layout(dims)
        n2 = (kmax_par + 1) * (kmax_per + 1)
        o = 0

        def take(m):
            nonlocal o
            v = out[o:o + m]
            o += m
            return v

        self.k3D, self.Nmodes3D = take(kmax), take(kmax)
        self.k1D, self.Nmodes1D = take(kmax_par), take(kmax_par)
        self.kpar, self.kper, self.Nmodes2D = take(n2), take(n2), take(n2)
        C = fields * (fields + 1) // 2
        # the pairs are the last index of each block (autos, then crosses): views, no copies
        P3 = take(3 * kmax * C).reshape(kmax, 3, C)
        P1 = take(kmax_par * C).reshape(kmax_par, C)
        P2 = take(n2 * C).reshape(n2, C)
        self.Pk, self.XPk = P3[:, :, :fields], P3[:, :, fields:]
        self.Pk1D, self.PkX1D = P1[:, :fields], P1[:, fields:]
        self.Pk2D, self.PkX2D = P2[:, :fields], P2[:, fields:]
        check_number_modes(np.concatenate([[1.0], self.Nmodes3D]), dims)


class Bk:
    """FFT-shell bispectrum for (k1, k2, theta) (bispectrum_library.pyx:32-196), on the GPU."""

    def __init__(self, delta, BoxSize, k1, k2, theta, MAS="CIC", threads=1):
        import ctypes
        f64 = _lib.dtype_of(delta) == np.float64
        d = _grid3(delta, _lib.F64T if f64 else _lib.F32T, "delta")
        dims = int(d.shape[0])
        theta = np.ascontiguousarray(theta, dtype=np.float64)
        bins = theta.shape[0]
        B = np.zeros(bins)
        Q = np.zeros(bins)
        k = np.zeros(bins + 2)
        P = np.zeros(bins + 2)
        c = lambda a: a.ctypes.data_as(ctypes.c_void_p)
        rc = _lib.lib().pyl_bk(_lib.ptr(d), dims, float(BoxSize), float(k1), float(k2), c(theta), bins,
                               MAS_function(MAS), _lib.F64 if f64 else _lib.F32, c(B), c(Q), c(k), c(P),
                               _lib.stream())
        if rc == _lib.E_ARG and b"division by zero" in _lib.lib().pyl_last_error():
            raise ZeroDivisionError(_lib.lib().pyl_last_error().decode())
        _lib.check(rc)
        self.B, self.Q, self.k, self.Pk = B, Q, k, P


def correct_MAS(delta, BoxSize, MAS="CIC", threads=1):
    """The field with its mass-assignment window removed (Pk_library.pyx:1749-1806): the independent
    modes of FFT(delta) times prod_i (x_i/sin x_i)^p, back to real space (1/N^3-normalised), on the
    GPU.  float32 like the reference; float64 input stays float64."""
    f64 = _lib.dtype_of(delta) == np.float64
    real_t = _lib.F64T if f64 else _lib.F32T
    d = _grid3(delta, real_t, "delta")
    dims = int(d.shape[0])
    out = _lib.empty((dims, dims, dims), real_t)
    _lib.check(_lib.lib().pyl_correct_mas(_lib.ptr(d), _lib.ptr(out), dims, MAS_function(MAS),
                                          _lib.F64 if f64 else _lib.F32, _lib.stream()))
    return _ret(delta, out)


class Xi:
    """Correlation-function multipoles of a density field (Pk_library.pyx:2035-2151), on the GPU.

    Attributes r3D (mean |r| of the cells in each bin, Mpc/h), xi (kmax, 3) monopole, quadrupole,
    hexadecapole, Nmodes3D (cells per bin); bins floor(|r|/cell), the r = 0 bin dropped."""

    def __init__(self, delta, BoxSize, MAS="CIC", axis=2, threads=1):
        import ctypes
        BoxSize = float(np.float32(BoxSize))   # a C float in the reference (:2036)
        f64 = _lib.dtype_of(delta) == np.float64
        d = _grid3(delta, _lib.F64T if f64 else _lib.F32T, "delta")
        dims = int(d.shape[0])
        kmax = ctypes.c_int()
        n = ctypes.c_int64()
        _lib.check(_lib.lib().pyl_xi_layout(dims, ctypes.byref(kmax), ctypes.byref(n)))
        kmax = kmax.value
        out = np.zeros(n.value, dtype=np.float64)
        _lib.check(_lib.lib().pyl_xi(_lib.ptr(d), dims, BoxSize, MAS_function(MAS), int(axis),
                                     _lib.F64 if f64 else _lib.F32, out.ctypes.data_as(ctypes.c_void_p),
                                     _lib.stream()))
        self.r3D = out[:kmax].copy()
        self.xi = out[kmax:4 * kmax].reshape(kmax, 3).copy()
        self.Nmodes3D = out[4 * kmax:].copy()
