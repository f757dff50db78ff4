"""Mode counting by the binning kernels (Pk_library.pyx:376-381) and the mode check that uses those
counts (check_number_modes, Pk_library.pyx:90-102, 411).

The Nmodes that Pk / PkParticles / XPk return are the counts of the GPU kernel that binned the modes.
They are compared with the oracle's own per-mode count (its C mode loop adds 1 per mode, like the
reference) and with the closed-form total (N^3-8)/2+8.  A run with PYL_PK_DROP_KZ0 (the binning kernel
leaves out the kz = 0 plane: modes go missing, nothing faults) must end in the reference's
SystemExit, and the C ABI must return PYL_E_MODES."""
import ctypes

import numpy as np
import pytest
import torch

from oracle import cpu as C
from pylians_amd import synth
from tests._gpuref import oracle_pk

pytestmark = pytest.mark.gpu


def _field(N, L, seed):
    pos = synth.zeldovich(N, L, seed=seed, rms_disp=0.02 * L)
    rho = C.deposit_f64(pos, N, L, "CIC")
    return (rho / rho.mean() - 1.0).astype(np.float32)


def _expected_total(N):
    own = 1 if N % 2 else 8
    return (N ** 3 - own) // 2 + own


@pytest.mark.parametrize("N", [16, 24, 64, 160])
@pytest.mark.parametrize("axis", [0, 2])
def test_kernel_counts_equal_oracle_counts(gpu, N, axis):
    from pylians_amd import Pk_library
    L = 300.0
    delta = _field(N, L, seed=N + axis)
    g = Pk_library.Pk(delta, L, axis=axis, MAS="CIC")
    o = oracle_pk(delta, L, axis, "CIC")
    for name in ("Nmodes3D", "Nmodes1D", "Nmodes2D"):
        assert np.array_equal(getattr(g, name), getattr(o, name)), name
    assert int(g.Nmodes3D.sum()) + 1 == _expected_total(N)   # + the DC mode, dropped from the output


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_dropped_plane_exits_like_the_reference(gpu, axis, capsys):
    from pylians_amd import Pk_library, _lib
    N, L = 32, 200.0
    delta = _field(N, L, seed=3)
    Pk_library._EXTRA_FLAGS = _lib.PK_DROP_KZ0
    try:
        with pytest.raises(SystemExit):
            Pk_library.Pk(delta, L, axis=axis, MAS="CIC")
    finally:
        Pk_library._EXTRA_FLAGS = 0
    out = capsys.readouterr().out
    assert "Not all modes counted" in out and "Expected %d independent modes" % _expected_total(N) in out
    # the kz = 0 plane holds N^2/2 + 2 independent modes
    assert "Counted  %d independent modes" % (_expected_total(N) - (N * N // 2 + 2)) in out
    # the same field without the injection passes
    Pk_library.Pk(delta, L, axis=axis, MAS="CIC")


def test_dropped_plane_through_the_c_abi(gpu):
    from pylians_amd import _lib
    from pylians_amd.Pk_library import pk_layout, unpack_pk
    N, L = 32, 200.0
    d = torch.from_numpy(_field(N, L, seed=5)).cuda()
    out = torch.zeros(pk_layout(N)[3], dtype=torch.float64, device="cuda")
    L_ = _lib.lib()
    args = (_lib.ptr(d), N, L, 0, 2, _lib.F32)
    assert L_.pyl_pk(*args, _lib.PK_ALL, _lib.ptr(out), None, _lib.stream()) == 0
    assert L_.pyl_pk(*args, _lib.PK_ALL | _lib.PK_DROP_KZ0, _lib.ptr(out), None, _lib.stream()) == _lib.E_MODES
    assert b"Not all modes counted" in L_.pyl_last_error()
    # asynchronous: no error code, the returned counts show the loss
    assert L_.pyl_pk(*args, _lib.PK_ALL | _lib.PK_DROP_KZ0 | _lib.PK_ASYNC, _lib.ptr(out), None,
                     _lib.stream()) == 0
    torch.cuda.synchronize()

    class R:
        pass

    r = R()
    with pytest.raises(ZeroDivisionError):   # bins fed only by the kz = 0 plane are empty
        unpack_pk(r, out.cpu().numpy(), N)
    o = out.cpu().numpy()
    kmax = pk_layout(N)[2]
    n3 = o[5 * kmax:6 * kmax]
    assert int(n3.sum()) + 1 < _expected_total(N)


def test_fused_particles_path_checks_modes(gpu, capsys):
    from pylians_amd import _lib, pipeline
    L, N = 250.0, 32
    pos = synth.uniform(20 ** 3, L, seed=9)
    g = pipeline.PkParticles(pos, L, N, MAS="TSC", axis=1)
    assert int(g.Nmodes3D.sum()) + 1 == _expected_total(N)
    with pytest.raises(SystemExit):
        pipeline.PkParticles(pos, L, N, MAS="TSC", axis=1, flags=_lib.PK_ALL | _lib.PK_DROP_KZ0)
    assert "Not all modes counted" in capsys.readouterr().out


def test_xpk_counts_are_the_kernels(gpu):
    from pylians_amd import Pk_library
    N, L = 48, 400.0
    a, b = _field(N, L, seed=1), _field(N, L, seed=2)
    x = Pk_library.XPk([a, b], L, axis=1, MAS=["CIC", "CIC"])
    o = oracle_pk(a, L, 1, "CIC")
    for name in ("Nmodes3D", "Nmodes1D", "Nmodes2D"):
        assert np.array_equal(getattr(x, name), getattr(o, name)), name
