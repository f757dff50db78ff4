"""Writes the small golden fixtures under tests/golden/ from the FP64 oracle (run from the repo root:
``python tests/golden/make_golden.py``).  Inputs are seeded; outputs are the oracle's, which the
CPU tests pin independently (brute force, closed forms, the compiled reference MAS_c.c).

* mas16.npz   : 200 particles (seeded uniform, BoxSize 1) on a 16^3 and a 16^2 grid, every scheme,
                weighted and not; the reference unit-test input (Test/test_MAS.py:11-13) sums.
* pk16.npz    : a seeded 16^3 float32 field and every Pk attribute for axes 0/1/2, MAS CIC.
* modes.npz   : Nmodes3D/1D/2D for N = 16, 32 and axes 0/1/2 (exact lattice counts).
* bk16.npz    : a seeded 16^3 field and Bk(k1, k2, theta) outputs.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import bk as Bo  # noqa: E402
from oracle import mas as M  # noqa: E402
from oracle import pk as P  # noqa: E402

SCHEMES = ["NGP", "CIC", "TSC", "PCS"]


def main():
    rng = np.random.default_rng(2024)
    out = {}
    pos3 = rng.random((200, 3), dtype=np.float32)
    pos2 = rng.random((200, 2), dtype=np.float32)
    W = (rng.random(200, dtype=np.float32) + np.float32(0.5)).astype(np.float32)
    out["pos3"], out["pos2"], out["W"] = pos3, pos2, W
    for s in SCHEMES:
        out["g3_" + s] = M.MA(pos3, np.zeros((16,) * 3), 1.0, s)
        out["g3W_" + s] = M.MA(pos3, np.zeros((16,) * 3), 1.0, s, W)
        out["g2_" + s] = M.MA(pos2, np.zeros((16,) * 2), 1.0, s)
    np.random.seed(1)
    ref_pos = np.random.random((1000, 3)).astype(np.float32)
    out["ref_case_sums"] = np.array([M.MA(ref_pos, np.zeros((64,) * 3), 1.0, s).sum() for s in SCHEMES])
    np.savez_compressed(os.path.join(HERE, "mas16.npz"), **out)

    delta = rng.standard_normal((16, 16, 16)).astype(np.float32)
    pk = {"delta": delta}
    for axis in range(3):
        p = P.Pk(delta, 250.0, axis=axis, MAS="CIC")
        for name in ("k1D", "Pk1D", "Nmodes1D", "kpar", "kper", "Pk2D", "Nmodes2D", "k3D", "Pk", "Pkphase",
                     "Nmodes3D"):
            pk["%s_%d" % (name, axis)] = getattr(p, name)
    np.savez_compressed(os.path.join(HERE, "pk16.npz"), **pk)

    modes = {}
    for N in (16, 32):
        for axis in range(3):
            s, _ = P.pk_sums(np.zeros((N, N, N // 2 + 1), complex), N, axis, 0)
            for name in ("Nmodes3D", "Nmodes1D", "Nmodes2D"):
                modes["%s_%d_%d" % (name, N, axis)] = s[name].astype(np.int64)
    np.savez_compressed(os.path.join(HERE, "modes.npz"), **modes)

    field = np.exp(rng.standard_normal((16, 16, 16))).astype(np.float32)
    field = field / field.mean() - 1.0
    theta = np.linspace(0.0, np.pi, 5)
    kF = 2 * np.pi / 100.0
    b = Bo.Bk(field.astype(np.float64), 100.0, 2.2 * kF, 3.1 * kF, theta, MAS="CIC")
    np.savez_compressed(os.path.join(HERE, "bk16.npz"), delta=field.astype(np.float32), theta=theta,
                        k1=2.2 * kF, k2=3.1 * kF, B=b.B, Q=b.Q, k=b.k, Pk=b.Pk)


if __name__ == "__main__":
    main()
