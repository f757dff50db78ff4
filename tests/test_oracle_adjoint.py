"""Pins the FP64 analytic gradient of loss = sum_b w_b P0_b (oracle/adjoint.py) with central
finite differences (float64) on particles away from cell planes (CIC is piecewise linear)."""
import numpy as np
import pytest

from oracle import adjoint as A
from pylians_amd import synth


def test_gradient_matches_central_differences():
    N, L, nb = 16, 50.0, 6
    pos = synth.zeldovich(12, L, seed=3, rms_disp=1.5).astype(np.float64)
    w = np.random.default_rng(1).random(nb)
    g = A.grad(pos, N, L, w)
    h = 1e-3
    rng = np.random.default_rng(2)
    cell = L / N
    picked = 0
    for p in rng.permutation(len(pos)):
        frac = (pos[p] / cell) % 1.0
        if np.any(np.minimum(frac, 1 - frac) < 3 * h / cell):
            continue
        for a in range(3):
            pp, pm = pos.copy(), pos.copy()
            pp[p, a] += h
            pm[p, a] -= h
            fd = (A.loss(pp, N, L, w) - A.loss(pm, N, L, w)) / (2 * h)
            assert fd == pytest.approx(g[p, a], rel=2e-4, abs=1e-6 * np.abs(g).max())
        picked += 1
        if picked == 6:
            break
    assert picked == 6
