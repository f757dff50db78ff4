"""Seeded synthetic particle sets (uniform; Zel'dovich lattice with redshift-space distortions).

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
    for a in range(3):
        shape = [1, 1, 1]
        shape[a] = n
        x = scale * disp[:, a].reshape(n, n, n).astype(np.float64)
        x += q.reshape(shape)
        pos[:, a] = np.mod(x, L).ravel()
    del disp
    if sigma_v > 0.0:
        v = rng.normal(0.0, sigma_v, n ** 3).astype(np.float32)
        factor = np.float32((1.0 + redshift) / Hubble)
        x = pos[:, axis] + v * factor
        out = (x > np.float32(L)) | (x < np.float32(0.0))
        x[out] = np.mod(x[out] + np.float32(L), np.float32(L))
        pos[:, axis] = x
    return pos


def zeldovich_device(n_side, BoxSize, seed, rms_disp=5.0, kc=0.5, sigma_v=0.0, axis=0, redshift=0.0,
                     Hubble=100.0):
    """Same construction as ``zeldovich`` generated on the GPU with torch (seeded Philox streams),
    for the large benchmark sizes.  Returns a (n_side^3, 3) float32 device tensor in lattice order."""
    import torch

    n = int(n_side)
    L = float(BoxSize)
    dev = torch.device("cuda")
    g = torch.Generator(device=dev)
    g.manual_seed(int(seed))
    white = torch.randn((n, n, n), generator=g, device=dev, dtype=torch.float32)
    wk = torch.fft.rfftn(white)
    del white
    kF = 2.0 * np.pi / L
    f = torch.fft.fftfreq(n, 1.0 / n, device=dev, dtype=torch.float32) * kF
    fz = torch.arange(n // 2 + 1, device=dev, dtype=torch.float32) * kF
    k2 = f[:, None, None] ** 2 + f[None, :, None] ** 2 + fz[None, None, :] ** 2
    k2[0, 0, 0] = 1.0
    amp = torch.sqrt(torch.exp(-k2 / (kc * kc)) / k2) / k2
    amp[0, 0, 0] = 0.0
    wk = wk * (amp * 1j)
    del amp, k2
    pos = torch.empty((n ** 3, 3), device=dev, dtype=torch.float32)
    disp = []
    for ka in (f[:, None, None], f[None, :, None], fz[None, None, :]):
        disp.append(torch.fft.irfftn(wk * ka, s=(n, n, n)))
    del wk
    ms = sum(torch.sum(d.double() ** 2) for d in disp) / n ** 3
    scale = rms_disp / float(torch.sqrt(ms))
    q = (torch.arange(n, device=dev, dtype=torch.float64) + 0.5) * (L / n)
    for a in range(3):
        shape = [1, 1, 1]
        shape[a] = n
        x = disp[a].double() * scale + q.reshape(shape)
        pos[:, a] = torch.remainder(x, L).reshape(-1).float()
        disp[a] = None
    if sigma_v > 0.0:
        v = torch.randn(n ** 3, generator=g, device=dev, dtype=torch.float32) * sigma_v
        x = pos[:, axis] + v * np.float32((1.0 + redshift) / Hubble)
        out = (x > L) | (x < 0.0)
        x = torch.where(out, torch.remainder(x + L, L), x)
        pos[:, axis] = x
    return pos


def batch_jobs(count=64, seed=100, MAS="CIC", axis=0):
    """The uneven batch workload: `count` realizations, realization r drawn from a PCG64 stream seeded
    with `seed`: N_p = round(exp(U(ln 64^3, ln 256^3))), grid N uniform over {128, 192, 256, 384},
    L ~ U(250, 1000) Mpc/h, positions uniform in [0, L) generated on the GPU (Philox stream seeded
    with seed + r).  Returns PkBatch job dicts with device positions."""
    import torch

    rng = np.random.default_rng(seed)
    jobs = []
    for r in range(count):
        n_p = int(round(np.exp(rng.uniform(np.log(64 ** 3), np.log(256 ** 3)))))
        N = int(rng.choice([128, 192, 256, 384]))
        L = float(rng.uniform(250.0, 1000.0))
        g = torch.Generator(device="cuda")
        g.manual_seed(seed + r)
        pos = torch.rand((n_p, 3), generator=g, device="cuda", dtype=torch.float32) * np.float32(L)
        jobs.append({"pos": pos, "BoxSize": L, "dims": N, "MAS": MAS, "axis": axis})
    return jobs
