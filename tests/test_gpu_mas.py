"""GPU parity of the mass assignment (drop-in MAS_library) against the FP64 oracle."""
import ctypes

import numpy as np
import pytest
import torch

from oracle import cpu as C
from oracle import mas as M
from pylians_amd import synth
from tests._cases import SCHEMES, places_ok, ref_mas_case

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def MASL(gpu):
    from pylians_amd import MAS_library
    return MAS_library


@pytest.mark.parametrize("scheme", SCHEMES)
@pytest.mark.parametrize("weighted", [False, True])
def test_reference_unit_cases(MASL, scheme, weighted):
    """Test/test_MAS.py:10-238 through MA, the Cython-style and the MAS_c-style entry points."""
    W = np.full(1000, 3.0, np.float32) if weighted else None
    norm = 1000 * (3.0 if weighted else 1.0)
    places = 20 if scheme == "NGP" else 8
    for axes in (3, 2):
        pos = ref_mas_case(axes)
        d = np.zeros((64,) * axes, np.float32)
        MASL.MA(pos, d, 1.0, scheme, W=W)
        assert places_ok(np.sum(d, dtype=np.float64) / norm, places)
        d = np.zeros((64,) * axes, dtype=MASL.FLOAT_type())
        suffix = ("W" if weighted else "") + ("c3D" if axes == 3 else "c2D")
        fn = getattr(MASL, scheme + suffix)
        fn(pos, d, W, 1.0, 2) if weighted else fn(pos, d, 1.0, 2)
        assert places_ok(np.sum(d, dtype=np.float64) / norm, places)


@pytest.mark.parametrize("scheme", SCHEMES)
@pytest.mark.parametrize("weighted", [False, True])
@pytest.mark.parametrize("axes", [3, 2])
def test_per_cell_parity(MASL, scheme, weighted, axes):
    rng = np.random.default_rng(7)
    n, dims, L = 200000, 64, 1000.0
    pos = rng.random((n, axes), dtype=np.float32) * np.float32(L)
    W = (rng.random(n, dtype=np.float32) + np.float32(0.25)) if weighted else None
    g = np.zeros((dims,) * axes, np.float32)
    MASL.MA(pos, g, L, scheme, W=W)
    o = M.MA(pos, np.zeros((dims,) * axes), L, scheme, W)
    err = np.abs(g - o)
    # per cell: relative on rho (float32 weights and accumulation), no cell exempt
    assert np.all(err <= 1e-5 * np.abs(o) + 1e-12), (err / np.maximum(np.abs(o), 1e-30)).max()


@pytest.mark.parametrize("scheme", SCHEMES)
def test_zeldovich_parity_and_float64_grid(MASL, scheme):
    pos = synth.zeldovich(48, 300.0, seed=5, rms_disp=3.0, sigma_v=200.0)
    dims = 48
    g = np.zeros((dims,) * 3, np.float32)
    MASL.MA(pos, g, 300.0, scheme)
    o = C.deposit_f64(pos, dims, 300.0, scheme)
    assert np.all(np.abs(g - o) <= 1e-5 * np.abs(o) + 1e-12)
    if scheme in ("CIC", "NGP"):
        W = np.random.default_rng(1).random(len(pos), dtype=np.float32)
        gd = np.zeros((dims,) * 3, np.float64)
        getattr(MASL, scheme + "W_d")(pos, gd, 300.0, W)
        od = C.deposit_f64(pos, dims, 300.0, scheme, W)
        assert np.allclose(gd, od, rtol=1e-12, atol=1e-12)


def test_accumulates_and_device_inplace(MASL):
    rng = np.random.default_rng(9)
    pos = rng.random((5000, 3), dtype=np.float32) * np.float32(10.0)
    start = rng.random((16, 16, 16)).astype(np.float32)
    g = start.copy()
    MASL.MA(pos, g, 10.0, "TSC")
    o = M.MA(pos, start.astype(np.float64), 10.0, "TSC")
    assert np.allclose(g, o, rtol=1e-5, atol=0)
    t = torch.from_numpy(start.copy()).cuda()
    MASL.MA(torch.from_numpy(pos).cuda(), t, 10.0, "TSC")
    assert np.allclose(t.cpu().numpy(), o, rtol=1e-5, atol=0)


def test_2d_renormalisation_semantics(MASL):
    rng = np.random.default_rng(12)
    pos = rng.random((3000, 2), dtype=np.float32)
    prev = rng.random((32, 32)).astype(np.float32)
    for ren in (True, False):
        g = prev.copy()
        MASL.MA(pos, g, 1.0, "PCS", renormalize_2D=ren)
        o = M.MA(pos, prev.astype(np.float64), 1.0, "PCS", renormalize_2D=ren)
        assert np.allclose(g, o, rtol=2e-6, atol=1e-7)


def test_errors(MASL):
    with pytest.raises(SystemExit):
        MASL.MA(np.zeros((3, 3), np.float32), np.zeros((4, 4), np.float32), 1.0)
    with pytest.raises(SystemExit):
        MASL.MA(np.zeros((3, 3), np.float32), np.zeros((4, 4, 4), np.float32), 1.0, "QQQ")
    with pytest.raises(ValueError):
        MASL.MA(np.zeros((3, 3), np.float32), np.zeros((4, 4, 4), np.float64), 1.0)


def test_mas_c_host_entry_points(gpu):
    """pyl_mas_CIC etc. take host pointers exactly like MAS_c.h:3-10."""
    from pylians_amd import _lib
    L = _lib.lib()
    rng = np.random.default_rng(21)
    pos = rng.random((20000, 3), dtype=np.float32) * np.float32(50.0)
    W = rng.random(20000, dtype=np.float32)
    for s in SCHEMES:
        g = np.zeros((32,) * 3, np.float32)
        rc = getattr(L, "pyl_mas_" + s)(pos.ctypes.data_as(ctypes.c_void_p), g.ctypes.data_as(ctypes.c_void_p),
                                       W.ctypes.data_as(ctypes.c_void_p), len(pos), 32, 3, 50.0, 4)
        assert rc == 0
        o = C.deposit_f64(pos, 32, 50.0, s, W)
        assert np.allclose(g, o, rtol=1e-5, atol=1e-9)
        if C.ref_available():
            r = C.ref_mas_c(pos, np.zeros((32,) * 3, np.float32), 50.0, s, W, threads=4)
            assert np.allclose(g, r, rtol=1e-4, atol=1e-5)


def _dev_deposit(pos, W, dims, L, scheme, flags, start=None, f64=False):
    from pylians_amd import _lib
    p = torch.from_numpy(np.ascontiguousarray(pos)).cuda()
    w = None if W is None else torch.from_numpy(np.ascontiguousarray(W)).cuda()
    g = torch.from_numpy(np.zeros((dims,) * 3, np.float64 if f64 else np.float32) if start is None else start.copy()).cuda()
    _lib.check(_lib.lib().pyl_mas_deposit(_lib.ptr(p), _lib.ptr(w), len(pos), 3, dims, float(L), scheme,
                                          flags | (_lib.GRID_F64 if f64 else 0), _lib.ptr(g), _lib.stream()))
    return g.cpu().numpy()


@pytest.mark.parametrize("scheme", SCHEMES)
@pytest.mark.parametrize("weighted", [False, True])
def test_binned_is_correctly_rounded(gpu, scheme, weighted):
    """The cell-binned deposit accumulates in float64 (LDS) and rounds each cell once."""
    from pylians_amd import _lib
    rng = np.random.default_rng(31)
    n, dims, L = 300000, 64, 200.0
    pos = synth.zeldovich(64, L, seed=4, rms_disp=3.0, sigma_v=100.0)[:n] if not weighted else \
        rng.random((n, 3), dtype=np.float32) * np.float32(L)
    W = (rng.random(len(pos), dtype=np.float32) + np.float32(0.1)) if weighted else None
    o = C.deposit_f64(pos, dims, L, scheme, W)
    s = _lib.SCHEMES[scheme]
    g = _dev_deposit(pos, W, dims, L, s, _lib.BINNED | _lib.ZERO)
    assert np.array_equal(g, o.astype(np.float32)) or np.max(np.abs(g - o) / np.maximum(np.abs(o), 1e-30)) <= 6.5e-8
    ga = _dev_deposit(pos, W, dims, L, s, _lib.ATOMIC | _lib.ZERO)
    assert np.all(np.abs(ga - o) <= 1e-5 * np.abs(o) + 1e-12)
    # accumulate into an existing grid: one rounding of old + sum
    start = rng.random((dims,) * 3).astype(np.float32)
    gacc = _dev_deposit(pos, W, dims, L, s, _lib.BINNED, start=start)
    assert np.max(np.abs(gacc - (start + o)) / (start + o)) <= 6.5e-8
    g64 = _dev_deposit(pos, W, dims, L, s, _lib.BINNED | _lib.ZERO, f64=True)
    assert np.allclose(g64, o, rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize("scheme", SCHEMES)
@pytest.mark.parametrize("weighted", [False, True])
def test_binned_float32_weights(gpu, scheme, weighted):
    """PYL_ACC_F32W: 1-D weights and products in float32 from the float64 cell fraction, float64 cell
    sums.  Every contribution is within a few float32 ulps, so every cell (a sum of positive terms) is
    within 1e-6 relative of the FP64 oracle, in the plain, accumulating and overdensity modes."""
    from pylians_amd import _lib
    rng = np.random.default_rng(32)
    n, dims, L = 300000, 64, 200.0
    pos = synth.zeldovich(64, L, seed=5, rms_disp=3.0, sigma_v=100.0)[:n] if not weighted else \
        rng.random((n, 3), dtype=np.float32) * np.float32(L)
    W = (rng.random(len(pos), dtype=np.float32) + np.float32(0.1)) if weighted else None
    o = C.deposit_f64(pos, dims, L, scheme, W)
    s = _lib.SCHEMES[scheme]
    g = _dev_deposit(pos, W, dims, L, s, _lib.BINNED | _lib.ZERO | _lib.ACC_F32W)
    rel = np.abs(g - o) / np.maximum(np.abs(o), 1e-30)
    assert rel.max() <= 1e-6, rel.max()
    # the same cells through the presorted path
    gs = _dev_deposit(pos, W, dims, L, s, _lib.SORTED | _lib.ZERO | _lib.ACC_F32W)
    assert np.max(np.abs(gs - o) / np.maximum(np.abs(o), 1e-30)) <= 1e-6
    start = rng.random((dims,) * 3).astype(np.float32)
    gacc = _dev_deposit(pos, W, dims, L, s, _lib.BINNED | _lib.ACC_F32W, start=start)
    assert np.max(np.abs(gacc - (start + o)) / (start + o)) <= 1e-6


def test_binned_overflow_path_clustered(gpu):
    """All particles inside a few cells: buckets overflow and the exact-count path takes over."""
    from pylians_amd import _lib
    rng = np.random.default_rng(5)
    n, dims, L = 400000, 64, 64.0
    pos = np.concatenate([rng.random((n // 2, 3), dtype=np.float32) * np.float32(2.0) + np.float32(10.0),
                          rng.random((n // 2, 3), dtype=np.float32) * np.float32(L)]).astype(np.float32)
    for scheme in (2, 4):
        name = {2: "CIC", 4: "PCS"}[scheme]
        o = C.deposit_f64(pos, dims, L, name)
        g = _dev_deposit(pos, None, dims, L, scheme, _lib.BINNED | _lib.ZERO)
        assert np.max(np.abs(g - o) / np.maximum(np.abs(o), 1e-30)) <= 6.5e-8
        assert abs(np.sum(g, dtype=np.float64) / n - 1.0) < 1e-6


def test_binned_periodic_edges(gpu):
    from pylians_amd import _lib
    dims, L = 32, 32.0
    pos = np.array([[0.0, 0.0, 0.0], [32.0, 32.0, 32.0], [31.99, 0.01, 31.5], [15.999, 16.0, 16.001]] * 50, np.float32)
    for s, name in ((1, "NGP"), (2, "CIC"), (3, "TSC"), (4, "PCS")):
        o = C.deposit_f64(pos, dims, L, name)
        g = _dev_deposit(pos, None, dims, L, s, _lib.BINNED | _lib.ZERO)
        assert np.max(np.abs(g - o) / np.maximum(np.abs(o), 1e-30)) <= 6.5e-8, name


@pytest.mark.parametrize("scheme", SCHEMES)
@pytest.mark.parametrize("weighted", [False, True])
def test_presorted_binned_is_correctly_rounded(gpu, scheme, weighted):
    """The binned deposit after the supertile presort (PYL_SORTED): same single rounding of the float64
    cell sums, for unordered and ordered inputs, overwrite and accumulate, float32 and float64 grids."""
    from pylians_amd import _lib
    rng = np.random.default_rng(37)
    n, dims, L = 300000, 80, 150.0
    pos = rng.random((n, 3), dtype=np.float32) * np.float32(L) if weighted else \
        synth.zeldovich(66, L, seed=9, rms_disp=3.0, sigma_v=100.0)[:n]
    W = (rng.random(len(pos), dtype=np.float32) + np.float32(0.1)) if weighted else None
    o = C.deposit_f64(pos, dims, L, scheme, W)
    s = _lib.SCHEMES[scheme]
    g = _dev_deposit(pos, W, dims, L, s, _lib.SORTED | _lib.ZERO)
    assert np.max(np.abs(g - o) / np.maximum(np.abs(o), 1e-30)) <= 6.5e-8
    start = rng.random((dims,) * 3).astype(np.float32)
    gacc = _dev_deposit(pos, W, dims, L, s, _lib.SORTED, start=start)
    assert np.max(np.abs(gacc - (start + o)) / (start + o)) <= 6.5e-8
    g64 = _dev_deposit(pos, W, dims, L, s, _lib.SORTED | _lib.ZERO, f64=True)
    assert np.allclose(g64, o, rtol=1e-13, atol=1e-13)


def test_presorted_clustered_and_edges(gpu):
    """Presort with overflowing buckets (half the particles in a few cells) and on the periodic edges
    and the x = L face, against the float64 oracle."""
    from pylians_amd import _lib
    rng = np.random.default_rng(6)
    n, dims, L = 400000, 64, 64.0
    pos = np.concatenate([rng.random((n // 2, 3), dtype=np.float32) * np.float32(2.0) + np.float32(10.0),
                          rng.random((n // 2, 3), dtype=np.float32) * np.float32(L)]).astype(np.float32)
    edge = np.array([[0.0, 0.0, 0.0], [64.0, 64.0, 64.0], [63.99, 0.01, 63.5], [31.999, 32.0, 32.001]] * 50, np.float32)
    for scheme, name in ((2, "CIC"), (4, "PCS")):
        for p in (pos, edge):
            o = C.deposit_f64(p, dims, L, name)
            g = _dev_deposit(p, None, dims, L, scheme, _lib.SORTED | _lib.ZERO)
            assert np.max(np.abs(g - o) / np.maximum(np.abs(o), 1e-30)) <= 6.5e-8, name


def test_presorted_tiny_inputs(gpu):
    """Presort with fewer particles than one workgroup step, and with a single particle."""
    from pylians_amd import _lib
    dims, L = 32, 10.0
    for n in (1, 5, 700):
        pos = np.random.default_rng(n).random((n, 3), dtype=np.float32) * np.float32(L)
        o = C.deposit_f64(pos, dims, L, "TSC")
        g = _dev_deposit(pos, None, dims, L, 3, _lib.SORTED | _lib.ZERO)
        assert np.max(np.abs(g - o) / np.maximum(np.abs(o), 1e-30)) <= 6.5e-8, n


@pytest.mark.parametrize("ordered", [True, False])
@pytest.mark.parametrize("weighted", [False, True])
def test_dense_cic_register_runs(gpu, ordered, weighted):
    """CIC with at least 3 particles per cell takes register runs of the stencil block in the tiles
    (MAS_library.pyx:123-166): lattice-ordered particles (most records extend a run) and unordered
    ones (every record starts one), plain and presorted, overwrite and accumulate, float32 and float64
    grids -- each cell the single rounding of its float64 sum, as the default accumulation."""
    from pylians_amd import _lib
    rng = np.random.default_rng(43)
    dims, L = 48, 120.0
    if ordered:
        pos = synth.zeldovich(96, L, seed=12, rms_disp=2.0, sigma_v=50.0)      # 8 particles per cell
    else:
        pos = rng.random((5 * dims ** 3, 3), dtype=np.float32) * np.float32(L)  # 5 per cell
    W = (rng.random(len(pos), dtype=np.float32) + np.float32(0.1)) if weighted else None
    assert len(pos) >= 3 * dims ** 3
    o = C.deposit_f64(pos, dims, L, "CIC", W)
    for flags in (_lib.BINNED, _lib.SORTED):
        g = _dev_deposit(pos, W, dims, L, 2, flags | _lib.ZERO)
        assert np.max(np.abs(g - o) / np.maximum(np.abs(o), 1e-30)) <= 6.5e-8, flags
    start = rng.random((dims,) * 3).astype(np.float32)
    gacc = _dev_deposit(pos, W, dims, L, 2, _lib.BINNED, start=start)
    assert np.max(np.abs(gacc - (start + o)) / (start + o)) <= 6.5e-8
    g64 = _dev_deposit(pos, W, dims, L, 2, _lib.BINNED | _lib.ZERO, f64=True)
    assert np.allclose(g64, o, rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize("scheme", SCHEMES)
@pytest.mark.parametrize("weighted", [False, True])
def test_overdensity_entry_point(gpu, scheme, weighted):
    """pyl_mas_overdensity: delta = rho * dims^3 / M - 1 (Pk_snapshot.py:84-88) from every deposit path
    (atomic, binned, presorted, the autotuner's choice), into float32 and float64 grids.  float32: one
    rounding of the float64 value; float64: the oracle's value to 1e-12."""
    from pylians_amd import _lib
    rng = np.random.default_rng(47)
    dims, L = 48, 150.0
    pos = synth.zeldovich(64, L, seed=6, rms_disp=3.0, sigma_v=80.0)
    W = (rng.random(len(pos), dtype=np.float32) + np.float32(0.1)) if weighted else None
    rho = C.deposit_f64(pos, dims, L, scheme, W)
    mass = float(len(pos)) if W is None else float(np.sum(W, dtype=np.float64))
    d = rho * (dims ** 3 / mass) - 1.0
    p = torch.from_numpy(pos).cuda()
    w = None if W is None else torch.from_numpy(W).cuda()
    s = _lib.SCHEMES[scheme]
    for flags in (_lib.ATOMIC, _lib.BINNED, _lib.SORTED, 0):
        for f64 in (False, True):
            g = torch.full((dims,) * 3, 7.0, dtype=torch.float64 if f64 else torch.float32, device="cuda")
            _lib.check(_lib.lib().pyl_mas_overdensity(_lib.ptr(p), _lib.ptr(w), len(pos), dims, L, s,
                                                      flags | (_lib.GRID_F64 if f64 else 0), _lib.ptr(g),
                                                      _lib.stream()))
            h = g.cpu().numpy()
            if f64:
                assert np.allclose(h, d, rtol=1e-12, atol=1e-12), flags
            else:
                err = np.abs(h.astype(np.float64) - d)
                assert np.all(err <= 6.5e-8 * np.abs(d) + 1e-12), (flags, err.max())


def test_overdensity_errors(gpu):
    from pylians_amd import _lib
    g = torch.zeros((40,) * 3, device="cuda")
    p = torch.rand((100, 3), device="cuda") * 10
    L = _lib.lib()
    assert L.pyl_mas_overdensity(_lib.ptr(p), None, 100, 40, 10.0, 9, 0, _lib.ptr(g), _lib.stream()) == _lib.E_ARG
    assert L.pyl_mas_overdensity(_lib.ptr(p), None, 0, 40, 10.0, 2, 0, _lib.ptr(g), _lib.stream()) == _lib.E_ARG
    # dims 40 is not a multiple of 16: the binned deposit is refused, the atomic one runs
    assert L.pyl_mas_overdensity(_lib.ptr(p), None, 100, 40, 10.0, 2, _lib.BINNED, _lib.ptr(g),
                                 _lib.stream()) == _lib.E_UNSUPPORTED
    _lib.check(L.pyl_mas_overdensity(_lib.ptr(p), None, 100, 40, 10.0, 2, 0, _lib.ptr(g), _lib.stream()))
    assert abs(float(g.double().mean())) < 1e-6


def test_overdensity_python_wrapper(gpu):
    """pipeline.overdensity from host arrays equals the C entry point's default path."""
    from pylians_amd import pipeline
    pos = synth.zeldovich(32, 80.0, seed=3, rms_disp=2.0, sigma_v=50.0)
    rho = C.deposit_f64(pos, 32, 80.0, "TSC", None)
    d = rho * (32 ** 3 / len(pos)) - 1.0
    g = pipeline.overdensity(pos, 80.0, 32, "TSC")
    h = np.asarray(g.cpu().numpy() if hasattr(g, "cpu") else g.numpy(), np.float64)
    assert np.all(np.abs(h - d) <= 6.5e-8 * np.abs(d) + 1e-12)
    g64 = pipeline.overdensity(pos, 80.0, 32, "TSC", precision="fp64")
    h64 = g64.cpu().numpy() if hasattr(g64, "cpu") else g64.numpy()
    assert np.allclose(h64, d, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("scheme", SCHEMES)
@pytest.mark.parametrize("weighted", [False, True])
def test_binned_run_counting(gpu, scheme, weighted):
    """From 2 particles per cell the bucketing pass counts runs of equal (tile, class) keys in
    consecutive lanes once (lattice-ordered particles: runs of tens of lanes).  Ordered particles at 8
    per cell with a particle count that leaves the last wave ragged (lanes past the end form runs of
    their own), plain and presorted, deposit and overdensity: every cell the single rounding of its
    float64 sum, as without the run counting (MAS_library.pyx:123-545)."""
    from pylians_amd import _lib
    rng = np.random.default_rng(53)
    dims, L = 32, 90.0
    pos = synth.zeldovich(64, L, seed=14, rms_disp=1.5, sigma_v=40.0)[:-777]
    assert len(pos) >= 2 * dims ** 3 and len(pos) % 2048 != 0
    W = (rng.random(len(pos), dtype=np.float32) + np.float32(0.1)) if weighted else None
    o = C.deposit_f64(pos, dims, L, scheme, W)
    s = _lib.SCHEMES[scheme]
    for flags in (_lib.BINNED, _lib.SORTED):
        g = _dev_deposit(pos, W, dims, L, s, flags | _lib.ZERO)
        assert np.max(np.abs(g - o) / np.maximum(np.abs(o), 1e-30)) <= 6.5e-8, flags
        g64 = _dev_deposit(pos, W, dims, L, s, flags | _lib.ZERO, f64=True)
        assert np.allclose(g64, o, rtol=1e-13, atol=1e-13), flags
    mass = float(len(pos)) if W is None else float(np.sum(W, dtype=np.float64))
    d = o * (dims ** 3 / mass) - 1.0
    p = torch.from_numpy(pos).cuda()
    w = None if W is None else torch.from_numpy(W).cuda()
    g = torch.empty((dims,) * 3, device="cuda")
    _lib.check(_lib.lib().pyl_mas_overdensity(_lib.ptr(p), _lib.ptr(w), len(pos), dims, L, s, _lib.BINNED,
                                              _lib.ptr(g), _lib.stream()))
    h = g.cpu().numpy().astype(np.float64)
    assert np.all(np.abs(h - d) <= 6.5e-8 * np.abs(d) + 1e-12)
