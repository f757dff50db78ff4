"""Fused particles -> P(k) (deposit, rho/rhobar - 1, r2c, deconvolution, binning) against the FP64 oracle."""
import numpy as np
import pytest

from pylians_amd import synth
from tests._gpuref import assert_pk_close, oracle_pk_particles

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pipe(gpu):
    from pylians_amd import pipeline
    return pipeline


def test_uniform_128_cic_axis0(pipe):
    """The CPU-runnable case: 128^3 uniform particles, CIC onto 128^3, axis 0."""
    N, L = 128, 1000.0
    pos = synth.uniform(N ** 3, L, seed=1)
    g = pipe.PkParticles(pos, L, N, MAS="CIC", axis=0)
    o, noise = oracle_pk_particles(pos, L, N, "CIC", 0, with_noise=True)
    errs = assert_pk_close(g, o, rtol=1e-5, noise=noise)
    print("uniform 128^3 CIC: max rel err vs FP64 oracle", errs)


@pytest.mark.parametrize("MAS", ["CIC", "TSC", "PCS"])
def test_zeldovich_rsd(pipe, MAS):
    n, N, L = 64, 64, 250.0
    pos = synth.zeldovich(n, L, seed=2, rms_disp=1.25, sigma_v=75.0, axis=0)
    g = pipe.PkParticles(pos, L, N, MAS=MAS, axis=0)
    o, noise = oracle_pk_particles(pos, L, N, MAS, 0, with_noise=True)
    errs = assert_pk_close(g, o, rtol=1e-5, noise=noise)
    print("zeldovich 64^3 %s: max rel err vs FP64 oracle" % MAS, errs)


def test_float64_path_and_weights(pipe):
    n, N, L = 48, 48, 200.0
    pos = synth.zeldovich(n, L, seed=3, rms_disp=1.0)
    W = np.random.default_rng(3).random(len(pos), dtype=np.float32) + np.float32(0.5)
    g = pipe.PkParticles(pos, L, N, MAS="CIC", axis=2, W=W, precision="fp64")
    o = oracle_pk_particles(pos, L, N, "CIC", 2, W=W)
    assert_pk_close(g, o, rtol=1e-10)
    g32 = pipe.PkParticles(pos, L, N, MAS="CIC", axis=2, W=W)
    _, noise = oracle_pk_particles(pos, L, N, "CIC", 2, W=W, with_noise=True)
    assert_pk_close(g32, o, rtol=1e-5, noise=noise)
