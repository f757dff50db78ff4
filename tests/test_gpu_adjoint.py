"""GPU adjoint of loss = sum_b w_b P0_b against the FP64 analytic gradient (itself pinned by finite
differences in tests/test_oracle_adjoint.py), central differences of the GPU forward, autograd, and
CIC_interp against the reference formula."""
import numpy as np
import pytest
import torch

from oracle import adjoint as A
from oracle import mas as M
from pylians_amd import synth

pytestmark = pytest.mark.gpu


def test_adjoint_matches_fp64_gradient(gpu):
    from pylians_amd import autograd
    n_side, N, L, nb = 48, 48, 500.0 * 48 / 256, 12
    pos = synth.zeldovich(n_side, L, seed=3, rms_disp=5.0 * 48 / 256 * 2)
    w = np.random.default_rng(1).random(nb)
    g, P0 = autograd.grad_loss(pos, L, N, w)
    g = g.cpu().numpy().astype(np.float64)
    ref = A.grad(pos, N, L, w)
    rel = np.linalg.norm(g - ref) / np.linalg.norm(ref)
    print("adjoint: relative L2 error vs FP64 gradient %.2e, max abs / max %.2e"
          % (rel, np.abs(g - ref).max() / np.abs(ref).max()))
    assert rel <= 1e-5
    assert np.abs(g - ref).max() <= 1e-5 * np.abs(ref).max()


def test_autograd_and_finite_differences(gpu):
    from pylians_amd import autograd
    n_side, N, L, nb = 32, 32, 100.0, 8
    pos = synth.zeldovich(n_side, L, seed=9, rms_disp=1.5)
    w = torch.tensor(np.random.default_rng(2).random(nb), device="cuda")
    x = torch.from_numpy(pos).cuda().requires_grad_(True)
    P0 = autograd.pk0(x, L, N, "CIC", nbins=nb)
    loss = (w * P0).sum()
    loss.backward()
    ref = A.grad(pos, N, L, w.cpu().numpy())
    assert np.linalg.norm(x.grad.cpu().numpy() - ref) <= 1e-5 * np.linalg.norm(ref)
    # the forward value against the FP64 oracle loss
    assert float(loss) == pytest.approx(A.loss(pos, N, L, w.cpu().numpy()), rel=1e-5)


def test_cic_interp(gpu):
    from pylians_amd import MAS_library as MASL
    rng = np.random.default_rng(4)
    N, L = 32, 10.0
    dens = rng.random((N, N, N)).astype(np.float32)
    pos = rng.random((5000, 3), dtype=np.float32) * np.float32(L)
    den = np.zeros(5000, np.float32)
    MASL.CIC_interp(dens, L, pos, den)
    # reference formula (MAS_library.pyx:576-599) in float64 through the oracle's stencil
    idx, wts = [], []
    for a in range(3):
        i, ww = M.axis_weights(pos[:, a].astype(np.float64) * (N / L), "CIC", N)
        idx.append(i)
        wts.append(ww)
    ref = np.zeros(5000)
    for l in range(2):
        for m in range(2):
            for q in range(2):
                ref += dens[idx[0][:, l], idx[1][:, m], idx[2][:, q]] * wts[0][:, l] * wts[1][:, m] * wts[2][:, q]
    assert np.allclose(den, ref, rtol=1e-6, atol=1e-7)
