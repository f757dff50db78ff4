"""Test infrastructure: float32 restatement of pos_redshift_space (redshift_space_library.pyx:29-43).

Pinned by tests/test_oracle_rsd.py: exactly representable shifts, both wrap directions, the x = L
edge (not wrapped: the reference wraps only x > L or x < 0), several box crossings (fmod) and the
float32 rounding of the factor.
"""
import numpy as np


def factor(Hubble, redshift):
    """(1.0 + redshift)/Hubble: float32 inputs, double arithmetic, stored in a C float (:36)."""
    return np.float32((1.0 + float(np.float32(redshift))) / float(np.float32(Hubble)))


def pos_redshift_space(pos, vel, BoxSize, Hubble, redshift, axis):
    """Returns a shifted float32 copy of pos (the reference shifts in place, :38-43)."""
    pos = np.array(pos, dtype=np.float32, copy=True)
    vel = np.asarray(vel, dtype=np.float32)
    L = np.float32(BoxSize)
    f = factor(Hubble, redshift)
    s = pos[:, axis] + vel[:, axis] * f          # float32 product, then float32 sum (:40)
    out = (s > L) | (s < np.float32(0.0))
    s[out] = np.fmod(s[out] + L, L)              # (pos + BoxSize) % BoxSize, fmod semantics (:42-43)
    pos[:, axis] = s
    return pos
