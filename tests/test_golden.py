"""Golden fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py): the oracle must
reproduce them (CPU) and the GPU must match them to the parity bar (gpu)."""
import os

import numpy as np
import pytest

from oracle import bk as Bo
from oracle import mas as M
from oracle import pk as P

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
SCHEMES = ["NGP", "CIC", "TSC", "PCS"]
PK_ATTRS = ("k1D", "Pk1D", "Nmodes1D", "kpar", "kper", "Pk2D", "Nmodes2D", "k3D", "Pk", "Pkphase", "Nmodes3D")


def load(name):
    return np.load(os.path.join(G, name))


def test_oracle_reproduces_mas_fixture():
    f = load("mas16.npz")
    for s in SCHEMES:
        assert np.allclose(M.MA(f["pos3"], np.zeros((16,) * 3), 1.0, s), f["g3_" + s], rtol=1e-14, atol=1e-15)
        assert np.allclose(M.MA(f["pos3"], np.zeros((16,) * 3), 1.0, s, f["W"]), f["g3W_" + s], rtol=1e-14, atol=1e-15)
        assert np.allclose(M.MA(f["pos2"], np.zeros((16,) * 2), 1.0, s), f["g2_" + s], rtol=1e-14, atol=1e-15)
    assert f["ref_case_sums"][0] == 1000.0
    assert np.allclose(f["ref_case_sums"], 1000.0, rtol=1e-11)


def test_oracle_reproduces_pk_and_modes_fixtures():
    f = load("pk16.npz")
    for axis in range(3):
        p = P.Pk(f["delta"], 250.0, axis=axis, MAS="CIC")
        for name in PK_ATTRS:
            assert np.allclose(getattr(p, name), f["%s_%d" % (name, axis)], rtol=1e-12, atol=0), name
    m = load("modes.npz")
    for N in (16, 32):
        for axis in range(3):
            s, _ = P.pk_sums(np.zeros((N, N, N // 2 + 1), complex), N, axis, 0)
            for name in ("Nmodes3D", "Nmodes1D", "Nmodes2D"):
                assert np.array_equal(s[name].astype(np.int64), m["%s_%d_%d" % (name, N, axis)])


def test_oracle_reproduces_bk_fixture():
    f = load("bk16.npz")
    b = Bo.Bk(f["delta"].astype(np.float64), 100.0, float(f["k1"]), float(f["k2"]), f["theta"], MAS="CIC")
    for name in ("B", "Q", "k", "Pk"):
        assert np.allclose(getattr(b, name), f[name], rtol=1e-12), name


@pytest.mark.gpu
def test_gpu_matches_fixtures(gpu):
    from pylians_amd import MAS_library as MASL
    from pylians_amd import Pk_library as PKL
    f = load("mas16.npz")
    for s in SCHEMES:
        for key, pos, W, shape in (("g3_", f["pos3"], None, (16,) * 3), ("g3W_", f["pos3"], f["W"], (16,) * 3),
                                   ("g2_", f["pos2"], None, (16,) * 2)):
            g = np.zeros(shape, np.float32)
            MASL.MA(pos, g, 1.0, s, W=W)
            o = f[key + s]
            assert np.all(np.abs(g - o) <= 1e-5 * np.abs(o) + 1e-12), (key, s)
    p = load("pk16.npz")
    m = load("modes.npz")
    for axis in range(3):
        g = PKL.Pk(p["delta"], 250.0, axis=axis, MAS="CIC")
        assert np.array_equal(g.Nmodes3D, p["Nmodes3D_%d" % axis])
        assert np.array_equal(g.Nmodes2D.astype(np.int64), m["Nmodes2D_16_%d" % axis])
        assert np.allclose(g.Pk[:, 0], p["Pk_%d" % axis][:, 0], rtol=1e-5)
        assert np.allclose(g.k3D, p["k3D_%d" % axis], rtol=1e-12)
    b = load("bk16.npz")
    gb = PKL.Bk(b["delta"], 100.0, float(b["k1"]), float(b["k2"]), b["theta"], MAS="CIC")
    assert np.allclose(gb.B, b["B"], rtol=1e-5) and np.allclose(gb.Pk, b["Pk"], rtol=1e-5)
