"""Pins the mass-assignment oracle (oracle/mas.py) against the reference's own test cases,
closed forms, a brute-force kernel evaluation and the reference's MAS_c.c compiled from source."""
import numpy as np
import pytest

from oracle import cpu as C
from oracle import mas as M
from tests._cases import SCHEMES, places_ok, ref_mas_case


@pytest.mark.parametrize("scheme", SCHEMES)
@pytest.mark.parametrize("weighted", [False, True])
@pytest.mark.parametrize("axes", [3, 2])
def test_reference_mass_conservation(scheme, weighted, axes):
    """Test/test_MAS.py:10-238: sum(delta)/(W*particles) == 1 to places 20 (NGP) or 8."""
    pos = ref_mas_case(axes)
    W = np.full(1000, 3.0, np.float32) if weighted else None
    grid = M.MA(pos, np.zeros((64,) * axes), 1.0, scheme, W)
    tot = grid.sum() / (1000 * (3.0 if weighted else 1.0))
    assert places_ok(tot, 20 if scheme == "NGP" else 8)
    gc = M.MAS_c(pos, np.zeros((64,) * axes), 1.0, scheme, W)
    assert places_ok(gc.sum() / (1000 * (3.0 if weighted else 1.0)), 20 if scheme == "NGP" else 8)


@pytest.mark.parametrize("scheme", SCHEMES)
def test_script_mass_conservation_2d_3d(scheme):
    """library/tests/test.py:8-33 (reduced to 64^3 / 64^2 particles): |sum/elements - 1| < 1e-5."""
    rng = np.random.default_rng(5)
    dims = 64
    for axes in (3, 2):
        pos = rng.random((dims ** axes, axes), dtype=np.float32)
        for W in (None, np.ones(dims ** axes, np.float32)):
            g = M.MA(pos, np.zeros((dims,) * axes), 1.0, scheme, W)
            assert abs(g.sum() / dims ** axes - 1.0) < 1e-5


def test_cic_single_particle_closed_form():
    # dist = (1.25, 2.5, 3.75) cells on an 8^3 grid of BoxSize 8
    g = M.deposit(np.array([[1.25, 2.5, 3.75]], np.float32), 8, 8.0, "CIC")
    w = {(1, 2, 3): .75 * .5 * .25, (2, 2, 3): .25 * .5 * .25, (1, 3, 3): .75 * .5 * .25, (1, 2, 4): .75 * .5 * .75,
         (2, 3, 3): .25 * .5 * .25, (2, 2, 4): .25 * .5 * .75, (1, 3, 4): .75 * .5 * .75, (2, 3, 4): .25 * .5 * .75}
    exp = np.zeros((8, 8, 8))
    for k, v in w.items():
        exp[k] = v
    assert np.array_equal(g, exp)


def test_tsc_pcs_at_grid_point():
    # a particle exactly on grid point 3 along each axis
    p = np.array([[3.0, 3.0, 3.0]], np.float32)
    t = M.deposit(p, 8, 8.0, "TSC")
    w1 = {2: 0.125, 3: 0.75, 4: 0.125}
    for (i, j, k), v in np.ndenumerate(t):
        assert v == pytest.approx(w1.get(i, 0) * w1.get(j, 0) * w1.get(k, 0), abs=1e-15)
    q = M.deposit(p, 8, 8.0, "PCS")
    w2 = {2: 1 / 6, 3: 4 / 6, 4: 1 / 6}
    for (i, j, k), v in np.ndenumerate(q):
        assert v == pytest.approx(w2.get(i, 0) * w2.get(j, 0) * w2.get(k, 0), abs=1e-15)


@pytest.mark.parametrize("scheme", ["NGP", "CIC"])
def test_lattice_gives_one_per_cell(scheme):
    n, L = 16, 100.0
    q = np.arange(n) * (L / n)
    pos = np.stack(np.meshgrid(q, q, q, indexing="ij"), -1).reshape(-1, 3).astype(np.float32)
    g = M.deposit(pos, n, L, scheme)
    assert np.allclose(g, 1.0, rtol=0, atol=1e-6)


def test_periodic_wrap_edges():
    L, n = 10.0, 10
    g = M.deposit(np.array([[0.0, 0.0, 0.0], [10.0, 10.0, 10.0], [9.5, 9.5, 9.5]], np.float32), n, L, "CIC")
    assert g[0, 0, 0] == pytest.approx(2.0 + 0.125)
    assert g[9, 9, 9] == pytest.approx(0.125)
    assert g.sum() == pytest.approx(3.0)


@pytest.mark.parametrize("scheme", SCHEMES)
def test_brute_force_kernel_all_cells(scheme):
    """Every cell evaluated with the B-spline kernel of the periodic distance (no index arithmetic)."""
    rng = np.random.default_rng(11)
    n, L = 8, 3.0
    pos = rng.random((40, 3), dtype=np.float32) * np.float32(L)
    g = M.deposit(pos, n, L, scheme)
    d = pos.astype(np.float64) * (n / L)
    cells = np.arange(n, dtype=np.float64)
    bf = np.zeros((n, n, n))
    for p in d:
        ks = []
        for a in range(3):
            s = p[a] - cells
            s = s - n * np.round(s / n)
            ks.append(M.kernel_1d(s, scheme))
        bf += ks[0][:, None, None] * ks[1][None, :, None] * ks[2][None, None, :]
    assert np.allclose(g, bf, rtol=1e-12, atol=1e-13)


@pytest.mark.parametrize("scheme", SCHEMES)
@pytest.mark.parametrize("axes", [3, 2])
@pytest.mark.parametrize("weighted", [False, True])
def test_against_compiled_reference_mas_c(scheme, axes, weighted):
    """oracle/_ref/libmas_c_ref.so is the reference's MAS_c.c built from its own source."""
    if not C.ref_available():
        pytest.skip("reference MAS_c.c not built (reference tree absent)")
    rng = np.random.default_rng(3)
    n, dims, L = 20000, 32, 500.0
    pos = rng.random((n, axes), dtype=np.float32) * np.float32(L)
    W = rng.random(n, dtype=np.float32) + np.float32(0.5) if weighted else None
    ref = C.ref_mas_c(pos, np.zeros((dims,) * axes, np.float32), L, scheme, W, threads=2)
    o = M.MAS_c(pos, np.zeros((dims,) * axes), L, scheme, W)
    scale = np.abs(o).max()
    # float32 reference arithmetic (dist ~ 32 cells: ulp 4e-6) against float64
    assert np.max(np.abs(ref - o)) <= 2e-5 * scale


def test_2d_cython_duplication_and_renormalisation():
    """MA 2-D deposits STENCIL copies then divides the whole grid (MAS_library.pyx:84-110)."""
    rng = np.random.default_rng(4)
    pos = rng.random((500, 2), dtype=np.float32)
    prev = rng.random((16, 16))
    once = M.deposit(pos, 16, 1.0, "TSC")
    assert np.allclose(M.MA(pos, prev, 1.0, "TSC", renormalize_2D=False), prev + 3 * once)
    assert np.allclose(M.MA(pos, prev, 1.0, "TSC", renormalize_2D=True), (prev + 3 * once) / 3)
    assert np.allclose(M.MA(pos, prev, 1.0, "NGP"), prev + M.deposit(pos, 16, 1.0, "NGP"))


def test_errors_follow_reference():
    with pytest.raises(SystemExit):
        M.MA(np.zeros((3, 3), np.float32), np.zeros((4, 4)), 1.0, "CIC")
    with pytest.raises(SystemExit):
        M.MA(np.zeros((3, 3), np.float32), np.zeros((4, 4, 4)), 1.0, "XYZ")


def test_c_oracle_matches_numpy_oracle():
    rng = np.random.default_rng(8)
    pos = rng.random((3000, 3), dtype=np.float32) * np.float32(7.0)
    W = rng.random(3000, dtype=np.float32)
    for s in SCHEMES:
        a = M.deposit(pos, 12, 7.0, s, W)
        b = C.deposit_f64(pos, 12, 7.0, s, W)
        assert np.allclose(a, b, rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize("MAS", ["NGP", "CIC", "TSC", "PCS"])
def test_parallel_oracle_deposit_is_bit_identical(MAS):
    """The plane-split threaded oracle deposit (used at the benchmark sizes) gives the serial
    loop's bits: every cell gets the same terms in the same order."""
    from oracle import cpu as C
    rng = np.random.default_rng(11)
    n, dims, L = (1 << 20) + 4097, 40, 100.0
    pos = (rng.random((n, 3)) * L).astype(np.float32)
    pos[:7] = np.float32([0.0, L - 1e-4, L, 0.5 * L / dims, 99.99, 1e-7, 50.0])[:, None]
    W = rng.random(n).astype(np.float32)
    ser = C.deposit_f64(pos, dims, L, MAS, W, threads=1)
    for t in (3, 8):
        par = C.deposit_f64(pos, dims, L, MAS, W, threads=t)
        assert np.array_equal(par, ser), (MAS, t)
    assert ser.sum() == pytest.approx(W.astype(np.float64).sum(), rel=1e-12)
