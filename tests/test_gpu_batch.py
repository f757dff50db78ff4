"""Batched realizations of uneven sizes in one hipGraph: every replay equals the eager per-job
result and the FP64 oracle."""
import numpy as np
import pytest
import torch

from pylians_amd import synth
from tests._gpuref import assert_pk_close, oracle_pk_particles

pytestmark = pytest.mark.gpu

# Replays and eager runs differ only in the order of the float64 LDS additions of the cell sums: a
# cell whose exact sum lies within float64 rounding of a float32 rounding boundary (about one in 10^9)
# can round the other way, which moves a bin by up to ~1e-9 relative on these grids.
RT = 1e-8


def _jobs():
    rng = np.random.default_rng(100)
    jobs = []
    for r, (dims, n_side) in enumerate([(32, 20), (48, 40), (64, 33), (32, 48), (80, 30)]):
        L = float(rng.uniform(250.0, 1000.0))
        pos = synth.uniform(n_side ** 3, L, seed=100 + r)
        jobs.append({"pos": pos, "BoxSize": L, "dims": dims, "MAS": "CIC", "axis": r % 3})
    return jobs


def test_graph_batch_matches_eager_and_oracle(gpu):
    from pylians_amd import batch, pipeline
    jobs = _jobs()
    b = batch.PkBatch(jobs)
    eager = [pipeline.PkParticles(j["pos"], j["BoxSize"], j["dims"], MAS="CIC", axis=j["axis"]) for j in jobs]
    for rep in range(3):
        b.run()
        res = b.results()
        # the program's own eager work between replays
        x = torch.randn(1 << 20, device="cuda")
        torch.cuda.synchronize()
        assert float((x * 0).sum()) == 0.0
        for g, e in zip(res, eager):
            assert np.array_equal(g.Nmodes3D, e.Nmodes3D)
            assert np.allclose(g.Pk, e.Pk, rtol=RT, atol=0)
            assert np.allclose(g.Pk2D, e.Pk2D, rtol=RT, atol=1e-300)
    for g, j in zip(res, jobs):
        o, fl = oracle_pk_particles(j["pos"], j["BoxSize"], j["dims"], "CIC", j["axis"], with_noise=True)
        assert_pk_close(g, o, rtol=1e-5, noise=fl)
    b.close()


def test_batch_without_graph(gpu):
    from pylians_amd import batch
    jobs = _jobs()[:2]
    b = batch.PkBatch(jobs, graph=False)
    b.run()
    for g, j in zip(b.results(), jobs):
        o, fl = oracle_pk_particles(j["pos"], j["BoxSize"], j["dims"], "CIC", j["axis"], with_noise=True)
        assert_pk_close(g, o, rtol=1e-5, noise=fl)


def test_autotuner_cache_roundtrip(gpu, tmp_path):
    from pylians_amd import pipeline, tuning
    pos = synth.uniform(40 ** 3, 300.0, seed=5)
    pipeline.PkParticles(pos, 300.0, 64, MAS="TSC", axis=1)
    ent = tuning.entries()
    assert any(e["dims"] == 64 and e["scheme"] == 3 for e in ent)
    path = tuning.save(str(tmp_path / "tune.json"))
    assert tuning.load(path) == len(ent)
    assert tuning.entries() == ent


def test_graph_batch_presorted_weighted(gpu):
    """Weighted jobs through the presorted deposit inside the captured graph (its count, scan and
    scatter kernels and the zero fill are graph nodes): every replay equals the eager result, and
    the result the FP64 oracle."""
    from pylians_amd import _lib, batch, pipeline
    jobs = _jobs()[:3]
    rng = np.random.default_rng(7)
    for j in jobs:
        j["W"] = (rng.random(len(j["pos"])) + 0.25).astype(np.float32)
    b = batch.PkBatch(jobs, flags=_lib.PK_ALL | _lib.SORTED)
    eager = [pipeline.PkParticles(j["pos"], j["BoxSize"], j["dims"], MAS="CIC", axis=j["axis"], W=j["W"],
                                  deposit="presorted") for j in jobs]
    for rep in range(3):
        b.run()
        res = b.results()
        x = torch.randn(1 << 20, device="cuda")
        torch.cuda.synchronize()
        assert float((x * 0).sum()) == 0.0
        for g, e in zip(res, eager):
            assert np.array_equal(g.Nmodes3D, e.Nmodes3D)
            assert np.allclose(g.Pk, e.Pk, rtol=RT, atol=0)
    for g, j in zip(res, jobs):
        o, fl = oracle_pk_particles(j["pos"], j["BoxSize"], j["dims"], "CIC", j["axis"], W=j["W"], with_noise=True)
        assert_pk_close(g, o, rtol=1e-5, noise=fl)
    b.close()


def test_graph_owns_its_workspaces(gpu):
    """A captured batch keeps working after eager calls that grow (and so free and reallocate) the
    library's shared workspaces, also when such a call runs on another stream while the graph
    replays: the graph's buffers are its own."""
    from pylians_amd import Pk_library, batch, pipeline
    jobs = _jobs()[:3]
    b = batch.PkBatch(jobs)
    eager = [pipeline.PkParticles(j["pos"], j["BoxSize"], j["dims"], MAS="CIC", axis=j["axis"]) for j in jobs]
    big_pos = synth.uniform(64 ** 3, 700.0, seed=31)
    for rep, dims in enumerate((96, 128, 160)):
        # larger than every job: the shared slots are freed and allocated again
        big = pipeline.PkParticles(big_pos, 700.0, dims, MAS="PCS", axis=rep)
        b.run()
        # overlapping the replay (the graph runs on the batch's stream, this on the current one)
        f = torch.randn((dims, dims, dims), device="cuda", dtype=torch.float32)
        Pk_library.Pk(f, 700.0, axis=0, MAS="None")
        res = b.results()
        for g, e in zip(res, eager):
            assert np.array_equal(g.Nmodes3D, e.Nmodes3D)
            assert np.allclose(g.Pk, e.Pk, rtol=RT, atol=0)
            assert np.allclose(g.Pk1D, e.Pk1D, rtol=RT, atol=1e-300)
        assert int(big.Nmodes3D.sum()) + 1 == (dims ** 3 - 8) // 2 + 8
    for g, j in zip(res, jobs):
        o, fl = oracle_pk_particles(j["pos"], j["BoxSize"], j["dims"], "CIC", j["axis"], with_noise=True)
        assert_pk_close(g, o, rtol=1e-5, noise=fl)
    b.close()
