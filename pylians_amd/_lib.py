"""ctypes loader of libpylians_hip.so and device-buffer plumbing.

Device buffers come from one of two places, decided per call:

* torch, when the process already imported it (the caller passes torch tensors, or uses torch
  anyway): inputs and outputs are torch tensors and the library works on torch's current stream;
* the library itself otherwise (``DevArray``: pyl_dev_alloc / pyl_memcpy_*, the default stream), so
  that a numpy caller of the drop-in modules never imports torch.

There is one HIP runtime per process either way.  When torch is installed but not imported, the HIP
runtime bundled with it is loaded by path (RTLD_GLOBAL) before the library, whose DT_NEEDED
libamdhip64.so.7 then resolves to it, and a later ``import torch`` finds the same file already
loaded instead of bringing a second runtime.  There is no CPU fallback: every compute call goes
through the HIP library and fails loudly without a GPU.
"""
import ctypes
import importlib.util
import os
import sys

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpylians_hip.so")

# constants of include/pylians_hip.h
F32, F64 = 4, 8
GRID_F64, DUP2D, RENORM2D, ZERO, ATOMIC, BINNED, SORTED = 1, 2, 4, 8, 16, 32, 64
ACC_F32W, ACC_F64 = 1024, 2048
PK_1D, PK_2D, PK_PHASE, PK_ALL, PK_ASYNC, PK_DROP_KZ0 = 1, 2, 4, 7, 256, 512
E_ARG, E_UNSUPPORTED, E_MODES, E_ALLOC, E_LAUNCH = 1, 2, 3, 4, 5
SCHEMES = {"NGP": 1, "CIC": 2, "TSC": 3, "PCS": 4}

# element types of device buffers (numpy dtypes; torch dtypes are accepted where a dtype is asked for)
F32T, F64T, C64T, C128T = np.dtype(np.float32), np.dtype(np.float64), np.dtype(np.complex64), np.dtype(np.complex128)

_c = ctypes
_vp, _i, _i64, _d, _f, _l = _c.c_void_p, _c.c_int, _c.c_int64, _c.c_double, _c.c_float, _c.c_long

# name -> argtypes; every entry point of include/pylians_hip.h
SIGNATURES = {
    "pyl_abi_version": [],
    "pyl_last_error": [],
    "pyl_device_count": [_vp],
    "pyl_dev_alloc": [_vp, _i64],
    "pyl_dev_free": [_vp],
    "pyl_memcpy_h2d": [_vp, _vp, _i64, _vp],
    "pyl_memcpy_d2h": [_vp, _vp, _i64, _vp],
    "pyl_memcpy_d2d": [_vp, _vp, _i64, _vp],
    "pyl_dev_zero": [_vp, _i64, _vp],
    "pyl_stream_sync": [_vp],
    "pyl_mas_deposit": [_vp, _vp, _i64, _i, _i, _d, _i, _i, _vp, _vp],
    "pyl_mas_overdensity": [_vp, _vp, _i64, _i, _d, _i, _i, _vp, _vp],
    "pyl_mas_NGP": [_vp, _vp, _vp, _l, _i, _i, _f, _i],
    "pyl_mas_CIC": [_vp, _vp, _vp, _l, _i, _i, _f, _i],
    "pyl_mas_TSC": [_vp, _vp, _vp, _l, _i, _i, _f, _i],
    "pyl_mas_PCS": [_vp, _vp, _vp, _l, _i, _i, _f, _i],
    "pyl_fft_r2c": [_vp, _vp, _i, _i, _vp],
    "pyl_fft_c2r": [_vp, _vp, _i, _i, _i, _vp],
    "pyl_pk_layout": [_i, _vp, _vp, _vp, _vp],
    "pyl_pk": [_vp, _i, _d, _i, _i, _i, _i, _vp, _vp, _vp],
    "pyl_pk_noalloc": [_vp, _i, _d, _i, _i, _i, _i, _vp, _vp, _vp],
    "pyl_pk_particles": [_vp, _vp, _i64, _i, _d, _i, _i, _i, _i, _vp, _vp],
    "pyl_pk_particles_rsd": [_vp, _vp, _vp, _i64, _i, _d, _f, _f, _i, _i, _i, _i, _i, _vp, _vp],
    "pyl_pos_redshift_space": [_vp, _vp, _i64, _i, _f, _f, _f, _i, _vp],
    "pyl_field_div_sub": [_vp, _i64, _d, _d, _i, _vp],
    "pyl_field_axpy": [_vp, _vp, _i64, _d, _i, _vp],
    "pyl_correct_mas": [_vp, _vp, _i, _i, _i, _vp],
    "pyl_xi_layout": [_i, _vp, _vp],
    "pyl_xi": [_vp, _i, _d, _i, _i, _i, _vp, _vp],
    "pyl_ft_filter": [_i, _f, _f, _i, _vp, _vp],
    "pyl_field_smoothing": [_vp, _vp, _vp, _i, _vp],
    "pyl_bk": [_vp, _i, _d, _d, _d, _vp, _i, _i, _i, _vp, _vp, _vp, _vp, _vp],
    "pyl_pk_adjoint": [_vp, _i64, _i, _d, _i, _vp, _i, _vp, _vp, _vp],
    "pyl_interp": [_vp, _i, _d, _vp, _i64, _i, _vp, _vp],
    "pyl_pk_batch": [_vp, _i, _vp],
    "pyl_graph_create": [_vp, _i, _vp, _vp],
    "pyl_graph_launch": [_vp, _vp],
    "pyl_graph_destroy": [_vp],
    "pyl_xpk_layout": [_i, _i, _vp],
    "pyl_xpk": [_vp, _i, _i, _d, _i, _vp, _i, _vp, _vp],
    "pyl_tune": [_i],
    "pyl_tune_export": [_c.c_char_p, _i],
    "pyl_tune_import": [_c.c_char_p],
    "pyl_profile": [_i],
    "pyl_profile_read": [_c.c_char_p, _i],
}

_lib = None


class PkJob(ctypes.Structure):
    """pyl_pk_job of include/pylians_hip.h"""
    _fields_ = [("pos", _vp), ("W", _vp), ("n", _i64), ("dims", _i), ("BoxSize", _d), ("scheme", _i),
                ("axis", _i), ("prec", _i), ("flags", _i), ("out", _vp)]


class PylError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("libpylians_hip error %d: %s" % (code, msg))
        self.code = code
        self.msg = msg


def _preload_hip_runtime():
    """Load the HIP runtime torch ships with (when torch is installed and not yet imported), so that
    the library and a later torch share one runtime."""
    if "torch" in sys.modules:
        return
    try:
        spec = importlib.util.find_spec("torch")
    except (ImportError, ValueError):
        spec = None
    if spec is None or not spec.submodule_search_locations:
        return
    for d in spec.submodule_search_locations:
        f = os.path.join(d, "lib", "libamdhip64.so")
        if os.path.exists(f):
            ctypes.CDLL(f, mode=ctypes.RTLD_GLOBAL)
            return


def lib():
    """The loaded library (raises if it is not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libpylians_hip.so is not built: run `python -m pylians_amd.build`")
        _preload_hip_runtime()
        L = ctypes.CDLL(LIB_PATH)
        for name, argt in SIGNATURES.items():
            fn = getattr(L, name)
            fn.argtypes = argt
            fn.restype = _c.c_char_p if name == "pyl_last_error" else _c.c_int
        _lib = L
    return _lib


def check(rc):
    if rc != 0:
        raise PylError(rc, lib().pyl_last_error().decode(errors="replace"))


# ---------------------------------------------------------------- backends
def torch_mode():
    """True when device buffers are torch tensors: the process has imported torch."""
    return "torch" in sys.modules


def _torch():
    import torch
    return torch


def is_tensor(a):
    """a is a torch tensor (without importing torch)."""
    t = sys.modules.get("torch")
    return t is not None and isinstance(a, t.Tensor)


def is_dev(a):
    """a is a device array: a torch tensor on the GPU or a DevArray."""
    return isinstance(a, DevArray) or (is_tensor(a) and a.is_cuda)


_NP2T = {"float32": "float32", "float64": "float64", "complex64": "complex64", "complex128": "complex128",
         "int32": "int32", "int64": "int64", "uint8": "uint8"}


def np_dtype(dtype):
    """numpy dtype of a numpy or torch dtype."""
    if is_tensor_dtype(dtype):
        return np.dtype(str(dtype).replace("torch.", ""))
    return np.dtype(dtype)


def is_tensor_dtype(dtype):
    t = sys.modules.get("torch")
    return t is not None and isinstance(dtype, t.dtype)


def dtype_of(a):
    """numpy dtype of a numpy array, torch tensor or DevArray."""
    return np_dtype(a.dtype)


def _torch_dtype(dtype):
    return getattr(_torch(), _NP2T[np_dtype(dtype).name])


def require_gpu():
    if torch_mode():
        if not _torch().cuda.is_available():
            raise RuntimeError("pylians_amd runs on a ROCm GPU (gfx950); no GPU is visible and there is no CPU fallback")
        return
    n = _c.c_int(0)
    check(lib().pyl_device_count(_c.byref(n)))
    if n.value < 1:
        raise RuntimeError("pylians_amd runs on a ROCm GPU (gfx950); no GPU is visible and there is no CPU fallback")


# Freed DevArray buffers are kept per size for reuse (at most _POOL_MAX bytes): hipFree synchronises
# the device, and the drop-in calls allocate the same few sizes again and again.
_POOL = {}
_POOL_BYTES = 0
_POOL_MAX = 8 << 30


def _pool_take(nbytes):
    global _POOL_BYTES
    lst = _POOL.get(nbytes)
    if lst:
        _POOL_BYTES -= nbytes
        return lst.pop()
    p = _c.c_void_p()
    rc = lib().pyl_dev_alloc(_c.byref(p), nbytes)
    if rc == E_ALLOC and _POOL_BYTES:
        release_pool()
        rc = lib().pyl_dev_alloc(_c.byref(p), nbytes)
    check(rc)
    return p.value


def _pool_give(ptr_, nbytes):
    global _POOL_BYTES
    if _lib is None:
        return
    if _POOL_BYTES + nbytes <= _POOL_MAX:
        _POOL.setdefault(nbytes, []).append(ptr_)
        _POOL_BYTES += nbytes
    else:
        _lib.pyl_dev_free(_c.c_void_p(ptr_))


def release_pool():
    """Free the device buffers kept for reuse."""
    global _POOL_BYTES
    for nbytes, lst in _POOL.items():
        for p in lst:
            lib().pyl_dev_free(_c.c_void_p(p))
    _POOL.clear()
    _POOL_BYTES = 0


class DevArray:
    """A C-contiguous device buffer owned by the library: shape, numpy dtype, device address."""

    def __init__(self, shape, dtype):
        self.shape = tuple(int(x) for x in shape)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape, dtype=np.int64)) * self.dtype.itemsize
        self._ptr = _pool_take(self.nbytes) if self.nbytes else None

    @property
    def ndim(self):
        return len(self.shape)

    def numel(self):
        return int(np.prod(self.shape, dtype=np.int64))

    def is_contiguous(self):
        return True

    def data_ptr(self):
        return self._ptr or 0

    def numpy(self):
        out = np.empty(self.shape, self.dtype)
        if self.nbytes:
            check(lib().pyl_memcpy_d2h(out.ctypes.data_as(_c.c_void_p), _c.c_void_p(self._ptr), self.nbytes,
                                       stream()))
        return out

    def __del__(self):
        if getattr(self, "_ptr", None):
            _pool_give(self._ptr, self.nbytes)
            self._ptr = None


def stream():
    """Stream of the library calls: torch's current stream, or the default stream."""
    if torch_mode():
        return _c.c_void_p(_torch().cuda.current_stream().cuda_stream)
    return None


def ptr(t):
    if t is None:
        return None
    return _c.c_void_p(t.data_ptr())


def to_dev(a, dtype):
    """Contiguous device array of `dtype` from a numpy array or a tensor: a torch tensor in torch
    mode (or for a tensor input), else a DevArray."""
    require_gpu()
    if is_tensor(a):
        t = a if a.is_cuda else a.cuda()
        return t.to(_torch_dtype(dtype)).contiguous()
    if isinstance(a, DevArray):
        return a if a.dtype == np_dtype(dtype) else to_dev(a.numpy(), dtype)
    arr = np.ascontiguousarray(a, dtype=np_dtype(dtype))
    d = empty(arr.shape, arr.dtype)
    if arr.nbytes:
        # staged through the library's pinned buffers; returns once `arr` may be freed
        check(lib().pyl_memcpy_h2d(ptr(d), arr.ctypes.data_as(_c.c_void_p), arr.nbytes, stream()))
    return d


def empty(shape, dtype):
    require_gpu()
    if torch_mode():
        return _torch().empty(shape, dtype=_torch_dtype(dtype), device="cuda")
    return DevArray(shape, dtype)


def zeros(shape, dtype):
    """A cleared device array."""
    require_gpu()
    if torch_mode():
        return _torch().zeros(shape, dtype=_torch_dtype(dtype), device="cuda")
    d = DevArray(shape, dtype)
    check(lib().pyl_dev_zero(ptr(d), d.nbytes, None))
    return d


def copy_rows(dev, first, host):
    """dev[first:first + len(host)] = host (host converted to the dtype of dev)."""
    arr = np.ascontiguousarray(host, dtype=dtype_of(dev))
    if arr.shape[0] == 0:
        return
    if is_tensor(dev) and not dev.is_contiguous():
        dev[first:first + arr.shape[0]].copy_(_torch().from_numpy(arr))
        return
    row = (dev.element_size() * dev.numel() if is_tensor(dev) else dev.nbytes) // dev.shape[0]
    check(lib().pyl_memcpy_h2d(_c.c_void_p(dev.data_ptr() + first * row), arr.ctypes.data_as(_c.c_void_p),
                               arr.nbytes, stream()))


def to_host(t):
    if isinstance(t, DevArray):
        return t.numpy()
    t = t.detach()
    if not t.is_cuda:
        return t.numpy()
    t = t.contiguous()
    out = np.empty(tuple(t.shape), np_dtype(t.dtype))
    if out.nbytes:
        # staged through the library's pinned buffers, after the work queued on torch's stream
        check(lib().pyl_memcpy_d2h(out.ctypes.data_as(_c.c_void_p), ptr(t), out.nbytes, stream()))
    return out


def profile(enable):
    """Enable (and reset) or disable the library's per-stage HIP-event timers."""
    check(lib().pyl_profile(1 if enable else 0))


def profile_read():
    """{stage: (count, total_ms)} from the library's stage timers."""
    n = lib().pyl_profile_read(None, 0)
    buf = ctypes.create_string_buffer(n + 1)
    lib().pyl_profile_read(buf, n + 1)
    out = {}
    for line in buf.value.decode().splitlines():
        name, cnt, ms = line.split()
        out[name] = (int(cnt), float(ms))
    return out
