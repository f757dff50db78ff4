"""Pins the power-spectrum oracle (oracle/pk.py, oracle/c/oracle_c.c) with the reference's mode-count
invariant, a brute-force conjugate-class enumeration of the full k-grid, plane-wave closed forms,
the window closed form and Parseval."""
import itertools

import numpy as np
import pytest

from oracle import cpu as C
from oracle import pk as P


def full_classes(N):
    """One representative per conjugate class {k, -k} of the full N^3 k-grid, components in (-N/2, N/2]."""
    rep = lambda i: i - N if i > N // 2 else i
    seen, out = set(), []
    for a, b, c in itertools.product(range(N), repeat=3):
        if (a, b, c) in seen:
            continue
        conj = ((-a) % N, (-b) % N, (-c) % N)
        seen.add((a, b, c))
        seen.add(conj)
        out.append(((a, b, c), (rep(a), rep(b), rep(c))))
    return out


@pytest.mark.parametrize("N", [2, 4, 7, 8, 9, 12, 16, 20, 33])
def test_mode_count_invariant(N):
    s, _ = P.pk_sums(np.fft.rfftn(np.random.default_rng(N).standard_normal((N, N, N))), N, 0, 0)
    assert int(s["Nmodes3D"].sum()) == P.indep_modes_expected(N)
    P.check_number_modes(s["Nmodes3D"], N)
    with pytest.raises(SystemExit):
        P.check_number_modes(s["Nmodes3D"][1:], N)


@pytest.mark.parametrize("N,axis", [(8, 0), (8, 2), (12, 1), (9, 2), (10, 0)])
def test_brute_force_histograms(N, axis):
    """Every raw histogram of the mode loop, recomputed from a full-grid conjugate-class enumeration."""
    rng = np.random.default_rng(100 + N + axis)
    delta = rng.standard_normal((N, N, N))
    full = np.fft.fftn(delta)
    s, _ = P.pk_sums(np.fft.rfftn(delta), N, axis, 0)
    mid = N // 2
    _, _, kmax_par, kmax_per, kmax = P.frequencies(1.0, N)
    ref = {k: np.zeros_like(v) for k, v in s.items()}
    for idx, k in full_classes(N):
        d2 = abs(full[idx]) ** 2
        k2 = sum(x * x for x in k)
        kk = np.sqrt(k2)
        b = P.isqrt_array(np.array([k2]))[0]
        kpar = k[axis]
        perp = [k[i] for i in range(3) if i != axis]
        kper = P.isqrt_array(np.array([perp[0] ** 2 + perp[1] ** 2]))[0]
        mu2 = 0.0 if k2 == 0 else kpar * kpar / k2
        ref["Nmodes3D"][b] += 1
        ref["k3D"][b] += kk
        ref["P0"][b] += d2
        ref["P2"][b] += d2 * (3 * mu2 - 1) / 2
        ref["P4"][b] += d2 * (35 * mu2 * mu2 - 30 * mu2 + 3) / 8
        re = full[idx].real
        ref["Pkphase"][b] += np.arctan2(re, np.sqrt(d2)) ** 2
        if k2 <= mid * mid:
            ref["Nmodes1D"][abs(kpar)] += 1
            ref["k1D"][abs(kpar)] += abs(kpar)
            ref["Pk1D"][abs(kpar)] += d2
        i2 = (kmax_par + 1) * kper + abs(kpar)
        ref["Nmodes2D"][i2] += 1
        ref["Pk2D"][i2] += d2
    for key in s:
        if key.startswith("Nmodes"):
            assert np.array_equal(s[key], ref[key]), key
        elif key == "Pkphase":
            # atan2(Re, |d|) of k and -k are equal (Re is even): any class member gives the same phase
            assert np.allclose(s[key], ref[key], rtol=1e-10, atol=1e-10), key
        else:
            assert np.allclose(s[key], ref[key], rtol=1e-10, atol=1e-9), key


@pytest.mark.parametrize("los", [True, False])
@pytest.mark.parametrize("MAS", ["None", "CIC", "PCS"])
def test_plane_wave_multipoles(los, MAS):
    N, L, n = 16, 300.0, 3
    axis = 1
    wave_axis = axis if los else 2
    x = np.arange(N)
    shape = [1, 1, 1]
    shape[wave_axis] = N
    delta = np.broadcast_to(np.cos(2 * np.pi * n * x / N).reshape(shape), (N, N, N)).copy()
    p = P.Pk(delta, L, axis=axis, MAS=MAS)
    b = n - 1                     # DC bin dropped
    Nm = p.Nmodes3D[b]
    corr = P.MAS_correction(np.pi * n / N, P.MAS_function(MAS)) ** 2
    assert p.Pk[b, 0] == pytest.approx(L ** 3 / (4 * Nm) * corr, rel=1e-12)
    if los:
        assert p.Pk[b, 1] == pytest.approx(5 * p.Pk[b, 0], rel=1e-12)
        assert p.Pk[b, 2] == pytest.approx(9 * p.Pk[b, 0], rel=1e-12)
    else:
        assert p.Pk[b, 1] == pytest.approx(-2.5 * p.Pk[b, 0], rel=1e-12)
        assert p.Pk[b, 2] == pytest.approx(3.375 * p.Pk[b, 0], rel=1e-12)
    others = np.delete(p.Pk[:, 0], b)
    assert np.all(np.abs(others) < 1e-20 * L ** 3)


def test_window_closed_form():
    x = np.linspace(-1.5, 1.5, 31)
    for p in range(5):
        w = P.MAS_correction(x, p)
        with np.errstate(invalid="ignore"):
            exp = np.where(x == 0, 1.0, (x / np.sin(x)) ** p)
        assert np.allclose(w, exp)
    assert P.MAS_function("CIC") == 2 and P.MAS_function("None") == 0 and P.MAS_function("XYZ") == 0


@pytest.mark.parametrize("N", [8, 16, 15])
def test_parseval(N):
    rng = np.random.default_rng(N)
    delta = rng.standard_normal((N, N, N))
    s, _ = P.pk_sums(np.fft.rfftn(delta), N, 0, 0)
    full = np.fft.fftn(delta)
    own = [(a, b, c) for a in (0, N // 2) for b in (0, N // 2) for c in (0, N // 2)] if N % 2 == 0 else [(0, 0, 0)]
    own = set(own)
    self_sum = sum(abs(full[i]) ** 2 for i in own)
    total = 2 * s["P0"].sum() - self_sum
    assert total == pytest.approx(N ** 3 * np.sum(delta ** 2), rel=1e-11)


def test_outputs_normalisation_and_keep_deltak():
    N, L = 16, 123.0
    rng = np.random.default_rng(2)
    delta = rng.standard_normal((N, N, N))
    p = P.Pk(delta, L, axis=0, MAS="TSC", keep_deltak=True)
    kF = 2 * np.pi / L
    assert np.allclose(p.k1D, np.arange(1, N // 2 + 1) * kF)
    assert p.kpar.shape == p.kper.shape == p.Pk2D.shape
    assert p.kpar[0] == pytest.approx(0.5 * kF) and p.kper[N // 2 + 1] == pytest.approx(1.5 * kF)
    # keep_deltak: independent modes corrected, skipped modes left as in the FFT (Pk_library.pyx:327-355)
    raw = P.FFT3Dr(delta)
    _, _, _, indep = P.mode_grid(N)
    indep = np.broadcast_to(indep, raw.shape)
    assert np.array_equal(p.delta_k[~indep], raw[~indep])
    assert not np.allclose(p.delta_k[indep], raw[indep])


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_c_mode_loop_matches_numpy(axis):
    N = 16
    rng = np.random.default_rng(axis)
    dk = np.fft.rfftn(rng.standard_normal((N, N, N)))
    a, _ = P.pk_sums(dk, N, axis, 3)
    b = C.pk_sums_c(dk, N, axis, "TSC")
    for k in a:
        assert np.allclose(a[k], b[k], rtol=1e-12, atol=1e-12), k
    c = C.pk_sums_c(dk, N, axis, "TSC", faithful_f32=True)
    assert np.allclose(c["P0"], a["P0"], rtol=1e-5)


def test_xpk_oracle_consistency():
    """XPk autos equal Pk of each field (pinned above); the cross of a field with a scaled copy of
    itself is the scaled auto; cross of independent fields is symmetric in the pair order."""
    N, L = 16, 200.0
    rng = np.random.default_rng(9)
    a = rng.standard_normal((N, N, N))
    b = rng.standard_normal((N, N, N))
    x = P.XPk([a, 2.0 * a, b], L, axis=1, MAS=["CIC", "CIC", "TSC"])
    pa = P.Pk(a, L, axis=1, MAS="CIC")
    pb = P.Pk(b, L, axis=1, MAS="TSC")
    assert np.allclose(x.Pk[:, :, 0], pa.Pk, rtol=1e-12)
    assert np.allclose(x.Pk[:, :, 1], 4 * pa.Pk, rtol=1e-12)
    assert np.allclose(x.Pk[:, :, 2], pb.Pk, rtol=1e-12)
    assert np.allclose(x.XPk[:, :, 0], 2 * pa.Pk, rtol=1e-12)          # (0, 1)
    assert np.allclose(x.Pk1D[:, 0], pa.Pk1D) and np.allclose(x.Pk2D[:, 2], pb.Pk2D)
    assert np.allclose(x.PkX2D[:, 0], 2 * pa.Pk2D)
    assert np.array_equal(x.Nmodes3D, pa.Nmodes3D) and np.allclose(x.k3D, pa.k3D)
    y = P.XPk([b, a], L, axis=1, MAS=["TSC", "CIC"])
    assert np.allclose(y.XPk[:, :, 0], x.XPk[:, :, 1], rtol=1e-12)      # (b,a) == (a,b): pair (0,2) of x
    assert np.allclose(x.XPk[:, :, 2], 2 * y.XPk[:, :, 0], rtol=1e-12)  # (2a, b)
    with pytest.raises(SystemExit):
        P.XPk([a, np.zeros((8, 8, 8))], L, MAS=["CIC", "CIC"])


@pytest.mark.parametrize("axis", [0, 2])
def test_pk1d_area_factor_white_noise(axis):
    """Pk1D is the integral of P(k) over the disk |k_perp|^2 <= kN^2 - k_par^2 (Pk_library.pyx:387-392):
    P1D(k_par) = int d^2k_perp / (2 pi)^2 P(k).  For white noise of variance s^2 per cell, P(k) is flat,
    P = s^2 (L/N)^3 (and so is the measured P0), and E[P1D(k_par)] = P pi (kN^2 - k_par^2) / (2 pi)^2.
    The mean over seeded fields matches this closed form within its sampling error; a planted area
    factor of 2 pi kmaxper^2 / (2 pi)^2 (or pi kmaxper / (2 pi)^2) misses it."""
    N, L, s2, R = 32, 250.0, 1.0, 24
    rng = np.random.default_rng(77 + axis)
    acc1, acc0, n1 = 0.0, 0.0, None
    for _ in range(R):
        d = rng.standard_normal((N, N, N)) * np.sqrt(s2)
        o = P.Pk(d, L, axis=axis, MAS="None")
        acc1 = acc1 + o.Pk1D
        acc0 = acc0 + o.Pk[:, 0]
        n1 = o.Nmodes1D
        kF, kN = 2 * np.pi / L, (N // 2) * 2 * np.pi / L
        k1 = o.k1D
    Pflat = s2 * (L / N) ** 3
    mean0 = acc0 / R
    assert abs(np.mean(mean0) / Pflat - 1.0) < 0.01
    want = Pflat * np.pi * (kN ** 2 - k1 ** 2) / (2 * np.pi) ** 2
    mean1 = acc1 / R
    # per bin: the mean of R * n1 exponential-like |delta_k|^2 draws, relative sd 1/sqrt(R n1)
    ok = want > 0
    z = (mean1[ok] - want[ok]) / (want[ok] / np.sqrt(R * n1[ok]))
    assert np.all(np.abs(z) < 5.0), np.abs(z).max()
    assert np.all(mean1[~ok] == 0.0)
    wrong = Pflat * 2 * np.pi * (kN ** 2 - k1 ** 2) / (2 * np.pi) ** 2
    assert np.max(np.abs(mean1[ok] - wrong[ok]) / (want[ok] / np.sqrt(R * n1[ok]))) > 50
