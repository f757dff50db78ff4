"""Input cases shared by the CPU (oracle) and GPU (parity) tests."""
import numpy as np

SCHEMES = ["NGP", "CIC", "TSC", "PCS"]


def ref_mas_case(axes=3, seed=1, particles=1000):
    """The reference's unit-test input, Test/test_MAS.py:11-13 (legacy RNG, seed 1, BoxSize 1, dims 64)."""
    np.random.seed(seed)
    return np.random.random((particles, axes)).astype(np.float32)


def places_ok(x, places):
    """unittest.assertAlmostEqual(x, 1.0, places) semantics."""
    return round(abs(x - 1.0), places) == 0


def rel_err(a, b, floor=0.0):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return np.max(np.abs(a - b) / np.maximum(np.abs(b), floor)) if a.size else 0.0
