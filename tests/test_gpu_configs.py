"""GPU parity at the benchmark workloads' real sizes, against the FP64 oracle.

* 512^3 Zel'dovich particles (seed 2, rms displacement 5 Mpc/h, P(k) ~ k^-2 exp(-(k/0.5)^2)) moved
  to redshift space along axis 0 by v ~ N(0, 300 km/s); CIC, TSC and PCS onto a 512^3 grid,
  L = 1000 Mpc/h, every Pk output (Pk_library.pyx:266-425; MAS_library.pyx:123-166, 369-545).
* 64 uneven realizations (N_p log-uniform in [64^3, 256^3], N in {128, 192, 256, 384}, L in
  [250, 1000]), CIC + P0 in one captured graph: every replay equals the eager result, and one
  realization of each grid size equals the oracle.
* 256^3 Zel'dovich particles, CIC onto 256^3, L = 500: gradient of sum_k w_k P0(k) (64 bins)
  against the FP64 analytic gradient (itself pinned by finite differences).
* 512^3 Zel'dovich particles, CIC onto 256^3, L = 1000: Bk(k1 = 0.1, k2 = 0.2, 20 theta in [0, pi])
  (bispectrum_library.pyx:32-196).

Per-bin bar: P(k) and B(k) within 1e-5 relative (P2, P4 relative to P0 of the bin), Nmodes (counted by
the binning kernel) equal to the oracle's count.  At 512^3 a bin of the 2-D table, or of TSC/PCS's
window-suppressed corner, may hold one or a few modes whose power is 1e-4 of their neighbours'; there
no float32 grid and FFT reaches 1e-5, and the bar for a bin is max(1e-5, 6 sigma of the float32
representation noise of that bin) (``oracle.cpu.f32_sigma_pk``).  For CIC the multipoles P0/P2/P4,
the phase and the 1-D spectrum are held to 1e-5 outright.  The errors are printed with the test (-s)."""
import numpy as np
import pytest
import torch

from oracle import bk as Bo
from oracle import cpu as C
from pylians_amd import synth

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _rel(a, b, den=None):
    den = np.abs(b) if den is None else np.abs(den)
    return np.abs(a - b) / den


@pytest.fixture(scope="module")
def zeldovich512(gpu):
    pos = synth.zeldovich_device(512, 1000.0, seed=2, rms_disp=5.0, kc=0.5, sigma_v=300.0, axis=0)
    torch.cuda.synchronize()
    host = pos.cpu().numpy()
    yield pos, host
    del pos
    torch.cuda.empty_cache()


@pytest.mark.parametrize("MAS", ["CIC", "TSC", "PCS"])
def test_zeldovich512_rsd_grid512(zeldovich512, MAS):
    from pylians_amd import pipeline
    from tests._gpuref import assert_pk_close
    pos, host = zeldovich512
    N, L = 512, 1000.0
    g = pipeline.PkParticles(pos, L, N, MAS=MAS, axis=0)
    o, nz = C.fp64_oracle_pipeline(host, N, L, 0, MAS=MAS, workers=C.host_threads(), with_sigma=True)
    # every output per bin within max(1e-5, 6 sigma of the float32 representation noise of the bin)
    res = assert_pk_close(g, o, TOL, noise=nz)
    P0 = o.Pk[:, 0]
    e0 = _rel(g.Pk[:, 0], P0)
    e2 = _rel(g.Pk[:, 1], o.Pk[:, 1], P0)
    e4 = _rel(g.Pk[:, 2], o.Pk[:, 2], P0)
    m2 = np.arange(len(o.Pk2D)) > 0   # (0, 0) holds only the DC mode
    e2d = _rel(g.Pk2D[m2], o.Pk2D[m2])
    print("\n%s 512^3: max rel err P0 %.2e  P2/P0 %.2e  P4/P0 %.2e  phase %.2e  Pk1D %.2e  Pk2D %.2e; "
          "3-D bins above 1e-5: %d of %d; 2-D bins above 1e-5: %d of %d; float32 sigma_rel %.2e"
          % (MAS, e0.max(), e2.max(), e4.max(), res["phase"], res["Pk1D"], res["Pk2D"],
             int(np.sum(np.maximum(e0, np.maximum(e2, e4)) > TOL)), len(e0), int(np.sum(e2d > TOL)), len(e2d),
             nz.sigma_rel))
    if MAS == "CIC":
        # the benchmark's configuration: the multipoles, the phase and the 1-D spectrum hold 1e-5
        # outright in every bin
        assert e0.max() <= TOL and e2.max() <= TOL and e4.max() <= TOL
        assert res["phase"] <= TOL and res["Pk1D"] <= TOL
        # the float64 path (float64 grid, complex128 transform; its binning blocks take one CU each
        # with 8 lines of 512 points) against the same oracle
        g64 = pipeline.PkParticles(pos, L, N, MAS=MAS, axis=0, precision="fp64")
        r64 = assert_pk_close(g64, o, 1e-10)
        print("float64 path 512^3: max rel err P0 %.2e  phase %.2e  Pk2D %.2e" % (r64["P0"], r64["phase"], r64["Pk2D"]))


def test_batch64_uneven_graph(gpu):
    from pylians_amd import batch, pipeline
    from tests._gpuref import oracle_pk_particles
    jobs = synth.batch_jobs(64, seed=100)
    b = batch.PkBatch(jobs, flags=0)
    eager = [pipeline.PkParticles(j["pos"], j["BoxSize"], j["dims"], MAS="CIC", axis=0, flags=0) for j in jobs]
    # a replay and the eager run differ only in the order of the float64 LDS additions of the cell
    # sums: a cell whose exact sum lies within float64 rounding of a float32 rounding boundary (about
    # one in 10^9) can round the other way, which moves a bin by up to ~1e-9 relative
    for rep in range(2):
        b.run()
        res = b.results()
        worst = 0.0
        for g, e in zip(res, eager):
            assert np.array_equal(g.Nmodes3D, e.Nmodes3D)
            worst = max(worst, float(np.max(np.abs(g.Pk[:, 0] - e.Pk[:, 0]) / np.abs(e.Pk[:, 0]))))
        print("\nreplay %d: max rel difference to the eager runs %.2e" % (rep, worst))
        assert worst <= 1e-8
    checked = set()
    for g, j in zip(res, jobs):
        if j["dims"] in checked:
            continue
        checked.add(j["dims"])
        host = j["pos"].cpu().numpy()
        o, nz = C.fp64_oracle_pipeline(host, j["dims"], j["BoxSize"], 0, MAS="CIC", workers=C.host_threads(),
                                       with_sigma=True)
        assert np.array_equal(g.Nmodes3D, o.Nmodes3D)
        assert np.allclose(g.k3D, o.k3D, rtol=1e-12, atol=0)
        e0 = _rel(g.Pk[:, 0], o.Pk[:, 0])
        # white (shot-noise) spectra: the window-suppressed corner bins hold a few modes, where the bar is
        # max(1e-5, 6 sigma of the float32 representation noise of the bin)
        from tests._gpuref import NOISE_Z, _envelope
        allowed = np.maximum(TOL, NOISE_Z * _envelope(nz.Pk[:, 0] / np.abs(o.Pk[:, 0])))
        print("\nbatch job N=%d n=%d L=%.1f: P0 max rel err %.2e (%d of %d bins above 1e-5, max %.1f sigma)"
              % (j["dims"], len(host), j["BoxSize"], e0.max(), int(np.sum(e0 > TOL)), len(e0),
                 float(np.max(np.abs(g.Pk[:, 0] - o.Pk[:, 0]) / nz.Pk[:, 0]))))
        assert np.all(e0 <= allowed)
    assert checked == {128, 192, 256, 384}
    b.close()


def test_adjoint_zeldovich256(gpu):
    from oracle import adjoint as A
    from pylians_amd import autograd
    N, L = 256, 500.0
    pos = synth.zeldovich_device(256, L, seed=3, rms_disp=5.0)
    w = np.random.default_rng(3).random(64)
    g, P0 = autograd.grad_loss(pos, L, N, w)
    g = g.cpu().numpy().astype(np.float64)
    host = pos.cpu().numpy()
    del pos
    ref = A.grad(host, N, L, w)
    rel = np.linalg.norm(g - ref) / np.linalg.norm(ref)
    mx = np.abs(g - ref).max() / np.abs(ref).max()
    print("\nadjoint 256^3: relative L2 error %.2e, max abs / max %.2e" % (rel, mx))
    assert rel <= TOL
    assert mx <= TOL


def test_bk_zeldovich512_grid256_20theta(gpu):
    from pylians_amd import Pk_library, _lib
    N, L = 256, 1000.0
    pos = synth.zeldovich_device(512, L, seed=200, rms_disp=5.0)
    host = pos.cpu().numpy()
    grid = torch.zeros((N, N, N), dtype=torch.float32, device="cuda")
    _lib.check(_lib.lib().pyl_mas_deposit(_lib.ptr(pos), None, int(pos.shape[0]), 3, N, L, 2, _lib.ZERO,
                                          _lib.ptr(grid), _lib.stream()))
    # the benchmark's deposit: delta straight from pyl_mas_overdensity (8 particles per cell: the
    # bucketing pass counts runs of equal keys once)
    dgrid = torch.empty((N, N, N), dtype=torch.float32, device="cuda")
    _lib.check(_lib.lib().pyl_mas_overdensity(_lib.ptr(pos), None, int(pos.shape[0]), N, L, 2, 0, _lib.ptr(dgrid),
                                              _lib.stream()))
    n_p = int(pos.shape[0])
    del pos
    dover = dgrid.cpu().numpy().astype(np.float64)
    del dgrid
    grid /= grid.double().mean().float()
    grid -= 1.0
    theta = np.linspace(0, np.pi, 20)
    g = Pk_library.Bk(grid, L, 0.1, 0.2, theta, MAS="CIC")
    del grid
    rho = C.deposit_f64(host, N, L, "CIC")
    exact = rho * (N ** 3 / n_p) - 1.0
    assert np.all(np.abs(dover - exact) <= 6.5e-8 * np.abs(exact) + 1e-12)
    del dover, exact
    delta = rho / rho.mean() - 1.0
    del rho, host
    o = Bo.Bk(delta, L, 0.1, 0.2, theta, MAS="CIC")
    eB, eQ, eP = _rel(g.B, o.B), _rel(g.Q, o.Q), _rel(g.Pk, o.Pk)
    print("\nBk 256^3, 20 theta: max rel err B %.2e  Q %.2e  Pk %.2e" % (eB.max(), eQ.max(), eP.max()))
    assert np.allclose(g.k, o.k, rtol=1e-14)
    assert eB.max() <= TOL and eQ.max() <= TOL and eP.max() <= TOL
