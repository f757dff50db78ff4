"""ctypes loader of libpylians_hip.so and device-buffer plumbing.

torch is imported first so that the process has exactly one HIP runtime (torch's); the
library's DT_NEEDED libamdhip64.so.7 then resolves to it and device pointers, streams
and events are shared.  torch is used only for device memory and streams.  There is no CPU
fallback: every compute call goes through the HIP library and fails loudly without a GPU.
"""
import ctypes
import os

import numpy as np
import torch

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpylians_hip.so")

# constants of include/pylians_hip.h
F32, F64 = 4, 8
GRID_F64, DUP2D, RENORM2D, ZERO, ATOMIC, BINNED, SORTED = 1, 2, 4, 8, 16, 32, 64
ACC_F32W, ACC_F64 = 1024, 2048
PK_1D, PK_2D, PK_PHASE, PK_ALL, PK_ASYNC, PK_DROP_KZ0 = 1, 2, 4, 7, 256, 512
E_ARG, E_UNSUPPORTED, E_MODES, E_ALLOC, E_LAUNCH = 1, 2, 3, 4, 5
SCHEMES = {"NGP": 1, "CIC": 2, "TSC": 3, "PCS": 4}

_c = ctypes
_vp, _i, _i64, _d, _f, _l = _c.c_void_p, _c.c_int, _c.c_int64, _c.c_double, _c.c_float, _c.c_long

# name -> argtypes; every entry point of include/pylians_hip.h
SIGNATURES = {
    "pyl_abi_version": [],
    "pyl_last_error": [],
    "pyl_mas_deposit": [_vp, _vp, _i64, _i, _i, _d, _i, _i, _vp, _vp],
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


def lib():
    """The loaded library (raises if it is not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libpylians_hip.so is not built: run `python -m pylians_amd.build`")
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


def require_gpu():
    if not torch.cuda.is_available():
        raise RuntimeError("pylians_amd runs on a ROCm GPU (gfx950); no GPU is visible and there is no CPU fallback")


def stream():
    return _c.c_void_p(torch.cuda.current_stream().cuda_stream)


def ptr(t):
    return None if t is None else _c.c_void_p(t.data_ptr())


def is_dev(a):
    return isinstance(a, torch.Tensor) and a.is_cuda


def to_dev(a, dtype):
    """Contiguous device tensor of `dtype` (torch dtype) from a numpy array or a tensor."""
    require_gpu()
    if isinstance(a, torch.Tensor):
        t = a if a.is_cuda else a.cuda()
        return t.to(dtype).contiguous()
    arr = np.ascontiguousarray(a, dtype=torch.empty((), dtype=dtype).numpy().dtype)
    return torch.from_numpy(arr).cuda()


def empty(shape, dtype):
    require_gpu()
    return torch.empty(shape, dtype=dtype, device="cuda")


def to_host(t):
    return t.detach().cpu().numpy()


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
