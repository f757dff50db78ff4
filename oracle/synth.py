"""Seeded synthetic inputs for the tests and the benchmark (TEST INFRASTRUCTURE ONLY).

* ``uniform``   -- positions uniform in [0, L), float32 (the CPU-runnable 128^3 case).
* ``zeldovich`` -- a particle lattice q = (i+1/2) L/n displaced by the Zel'dovich
  field Psi_k = i k delta_k / k^2 of a Gaussian field with P(k) = A k^-2
  exp(-(k/kc)^2), A fixed so that the 3-D rms |Psi| equals ``rms_disp``;
  optional redshift-space shift s = x + v (1+z)/H along ``axis`` with
  v ~ N(0, sigma_v), wrapped as in redshift_space_library.pyx:29-43.
  Particles are stored in lattice order (x slowest, z fastest).
"""
import os

import numpy as np
import scipy.fft


def _workers():
    try:
        return max(1, min(16, len(os.sched_getaffinity(0))))
    except AttributeError:
        return 1


def uniform(n, BoxSize, seed, dtype=np.float32):
    rng = np.random.default_rng(seed)
    return (rng.random((n, 3), dtype=np.float32) * np.float32(BoxSize)).astype(dtype)


def zeldovich(n_side, BoxSize, seed, rms_disp=5.0, kc=0.5, sigma_v=0.0, axis=0,
              redshift=0.0, Hubble=100.0):
    """(n_side^3, 3) float32 positions in lattice order."""
    n = int(n_side)
    L = float(BoxSize)
    rng = np.random.default_rng(seed)
    white = rng.standard_normal((n, n, n), dtype=np.float32)
    wk = scipy.fft.rfftn(white, workers=_workers())
    del white
    kF = 2.0 * np.pi / L
    f = np.fft.fftfreq(n, 1.0 / n).astype(np.float32) * np.float32(kF)
    fz = np.arange(n // 2 + 1, dtype=np.float32) * np.float32(kF)
    k2 = f[:, None, None] ** 2 + f[None, :, None] ** 2 + fz[None, None, :] ** 2
    k2[0, 0, 0] = 1.0
    amp = np.sqrt(np.exp(-k2 / np.float32(kc * kc)) / k2) / k2   # sqrt(P)/k^2
    amp[0, 0, 0] = 0.0
    wk *= amp.astype(np.complex64) * np.complex64(1j)
    del amp, k2
    disp = np.empty((n ** 3, 3), dtype=np.float32)
    for a, ka in enumerate((f[:, None, None], f[None, :, None], fz[None, None, :])):
        disp[:, a] = scipy.fft.irfftn(wk * ka, s=(n, n, n), workers=_workers()).ravel()
    del wk
    scale = rms_disp / np.sqrt(np.mean(np.sum(disp.astype(np.float64) ** 2, axis=1)))
    q = (np.arange(n, dtype=np.float64) + 0.5) * (L / n)
    pos = np.empty((n ** 3, 3), dtype=np.float32)
    grid = np.meshgrid(q, q, q, indexing="ij")
    for a in range(3):
        pos[:, a] = np.mod(grid[a].ravel() + scale * disp[:, a].astype(np.float64), L)
    del disp, grid
    if sigma_v > 0.0:
        v = rng.normal(0.0, sigma_v, n ** 3).astype(np.float32)
        factor = np.float32((1.0 + redshift) / Hubble)
        x = pos[:, axis] + v * factor
        out = (x > np.float32(L)) | (x < np.float32(0.0))
        x[out] = np.mod(x[out] + np.float32(L), np.float32(L))
        pos[:, axis] = x
    return pos
