"""Closed form of Pk_library.Bk (MAS='None') for a field made of plane waves -- a check with no FFT
in it, used at every grid size the GPU FFT supports and pinned against the FP64 oracle on small grids
(tests/test_oracle_bk.py)."""
import numpy as np

# three waves closing a triangle, n1 + n2 + n3 = 0: |n| = 3, 2, sqrt(5)
TRIANGLE_WAVES = [(1, 2, 2), (0, -2, 0), (-1, 0, -2)]


def plane_wave_theta():
    """theta with k3 = sqrt(5) kF for k1 = 3 kF, k2 = 2 kF (k3^2 = 13 + 12 cos theta), and a wider one."""
    return np.array([np.arccos((5.0 - 13.0) / 12.0), 2.2])


def plane_wave_field(N, waves=TRIANGLE_WAVES):
    """sum_j cos(2 pi n_j.x / N) on the N^3 grid, float64 (numpy)."""
    x = np.arange(N)
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    return sum(np.cos(2 * np.pi * np.mod(a * X + b * Y + c * Z, N) / N) for a, b, c in waves)


def _kf(i, N):
    return i - N if i > N // 2 else i


def _shell(kmin, kmax, N):
    """Full-spectrum lattice vectors (components as kfreq) with kmin <= |k| < kmax."""
    r = int(np.ceil(kmax)) + 1
    c = np.arange(-r, r + 1)
    k = np.stack(np.meshgrid(c, c, c, indexing="ij"), -1).reshape(-1, 3)
    k = k[np.all((k > -((N + 1) // 2)) & (k <= N // 2), axis=1)]   # one representative per residue
    a = np.sqrt(np.sum(k.astype(np.float64) ** 2, axis=1))
    return k[(a >= kmin) & (a < kmax)]


def closed_form_bk(waves, N, L, k_all):
    """B, Pk of Bk (MAS='None') for delta = sum_j cos(2 pi n_j.x / N), from the mathematics alone:
    d_i = sum of the waves with |n_j| in shell i, so sum_x d_a d_b = N^3/4 #{s: s1 n + s2 m = 0 mod N}
    and sum_x d1 d2 d3 = N^3/8 #{s: s1 n + s2 m + s3 p = 0 mod N}; sum_x I_i^2 = |shell i| / N^3 and
    sum_x I1 I2 I3 = #{q1 + q2 + q3 = 0 mod N, q_i in shell i} / N^6 (1/N^3-normalised transforms)."""
    kF = 2 * np.pi / L
    kmin, kmax = (k_all - kF) / kF, (k_all + kF) / kF
    nrm = np.sqrt(np.sum(np.asarray(waves, dtype=np.float64) ** 2, axis=1))
    inside = [[w for w, a in zip(waves, nrm) if lo <= a < hi] for lo, hi in zip(kmin, kmax)]
    shells = [_shell(lo, hi, N) for lo, hi in zip(kmin, kmax)]
    zero = lambda v: np.all(np.mod(v, N) == 0)
    sgn = (1, -1)

    def pair(S):
        return sum(zero(s * np.array(a) + t * np.array(b)) for a in S for b in S for s in sgn for t in sgn)

    P = np.array([(N ** 3 / 4.0) * pair(S) / (len(sh) / N ** 3) * (L / N ** 2) ** 3
                  for S, sh in zip(inside, shells)])
    key = lambda v: tuple(np.mod(v, N))
    B = []
    for j in range(2, len(k_all)):
        tri = sum(zero(s * np.array(a) + t * np.array(b) + u * np.array(c))
                  for a in inside[0] for b in inside[1] for c in inside[j] for s in sgn for t in sgn for u in sgn)
        s3 = {key(q) for q in shells[j]}
        T = sum(key(-(q1 + q2)) in s3 for q1 in shells[0] for q2 in shells[1])
        B.append((N ** 3 / 8.0) * tri / (T / float(N) ** 6) * (L ** 2 / N ** 3) ** 3)
    return np.array(B), P
