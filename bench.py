#!/usr/bin/env python
"""Benchmark of the density-assignment -> P(k) path on one MI355X.

Workload (default): 512^3 particles from a Zel'dovich-displaced 512^3 lattice (Gaussian field
P(k) = A k^-2 exp(-(k/0.5)^2), rms displacement 5 Mpc/h, L = 1000 Mpc/h) shifted to redshift space
along axis 0 by v ~ N(0, 300 km/s); CIC onto a 512^3 grid, delta = rho/rhobar - 1, r2c, CIC
deconvolution and P0/P2/P4 (+ 1-D, 2-D, phase) binning, all in one fused library call
(pyl_pk_particles).  A step is one pass over the particle set with positions resident in HBM.

Prints ONE JSON line with the metric, the float64 GPU path's time (speedup), the roofline of the
dominant kernel (HIP events recorded by the library on the stream it launches on) and the
CPU baseline (the reference's own MAS_c.c + pocketfft + the restated serial mode loop) timed on
this host, plus the error of the float32-compensated result against the FP64 oracle.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

METRIC = "Mparticles/s CIC+Pk end-to-end at 512^3; FP32 vs FP64 speedup at 1e-5 tol"
HBM_PEAK_GBS = 8000.0
ROOT = os.path.dirname(os.path.abspath(__file__))
# HBM bytes per launch of each stage's kernels at this workload, from rocprofv3 FETCH_SIZE /
# WRITE_SIZE passes scaled by what the counters read for a plain device copy of known size
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic_zeldovich512_cic_f32.txt")


def measured_traffic(workload):
    """{stage: bytes per launch} from the committed PMC summary (lines 'workload stage bytes')."""
    out = {}
    if not os.path.exists(TRAFFIC_FILE):
        return out
    for line in open(TRAFFIC_FILE):
        f = line.split()
        if len(f) == 3 and not line.startswith("#") and f[0] == workload:
            out[f[1]] = int(float(f[2]))
    return out


def e2e_bytes(n_p, N, weights=False):
    """Algorithmic bytes of one particles -> P(k) job: each stage's compulsory input read once and its
    materialised output written once; fusion is credited through time, never through fewer bytes.
      B_MA  = 12 N_p (+4 N_p with weights) + 4 N_g    (positions in, float32 overdensity grid out)
      B_FFT = 4 N_g + 8 N_h                           (real grid in, complex64 half spectrum out)
      B_BIN = 8 N_h                                   (half spectrum in; histograms negligible)"""
    Ng = N ** 3
    Nh = N * N * (N // 2 + 1)
    return (12 + (4 if weights else 0)) * n_p + 4 * Ng + (4 * Ng + 8 * Nh) + 8 * Nh


def algorithmic_bytes(n_p, N):
    """Per-kernel algorithmic bytes (float32 path, no weights), the split of e2e_bytes over the
    kernels that run it.  The deposit's compulsory traffic is its position read (k_bin: the bucket
    records it writes and k_tile reads back are not algorithmic) and its grid write (k_tile); the
    FFT's is the real grid read and the half spectrum written (pass Z); the binning reads the half
    spectrum once (the last FFT pass, fused with it).  The middle FFT pass re-reads and re-writes
    the half spectrum: an intermediate of the transform, counted with no algorithmic bytes of its
    own in the end-to-end figure; its own line uses the 16 N_h it moves."""
    Ng = N ** 3
    Nh = N * N * (N // 2 + 1)
    return {
        "deposit_atomic": 12 * n_p + 4 * Ng,     # read positions, write the grid
        "deposit_bin": 12 * n_p,                 # read positions
        "deposit_tiles": 4 * Ng,                 # write the overdensity grid
        "fft_z_r2c": 4 * Ng + 8 * Nh,            # read the real grid, write the half spectrum
        "fft_mid": 16 * Nh,                      # (intermediate: read + write the half spectrum)
        "fft_last_bin": 8 * Nh,                  # read the half spectrum
        "e2e": e2e_bytes(n_p, N),
    }


def adjoint_bytes(n_p, N):
    """Forward pass (e2e_bytes) plus the adjoint: gradient spectrum in and real gradient field out of
    the c2r (8 N_h + 4 N_g), then positions and the field read, dL/dpos written by the gather
    (12 N_p + 4 N_g + 12 N_p)."""
    Ng = N ** 3
    Nh = N * N * (N // 2 + 1)
    return e2e_bytes(n_p, N) + (8 * Nh + 4 * Ng) + (24 * n_p + 4 * Ng)


def bk_bytes(n_p, N, ntheta):
    """One Bk realization with the indicator-field sums cached: the deposit (12 N_p + 4 N_g), the r2c
    (4 N_g + 8 N_h), one c2r per shell (ntheta + 2 of them, 8 N_h + 4 N_g each) and one triple
    product per theta bin (three float32 fields read, 12 N_g)."""
    Ng = N ** 3
    Nh = N * N * (N // 2 + 1)
    return 12 * n_p + 4 * Ng + (4 * Ng + 8 * Nh) + (ntheta + 2) * (8 * Nh + 4 * Ng) + ntheta * 12 * Ng


def bk_bytes_parseval(n_p, N, K):
    """The compulsory traffic of this implementation's Bk (indicator sums cached): the deposit
    (12 N_p + 4 N_g), the r2c of delta (4 N_g + 8 N_h), the c2r of the k1 and k2 shells (8 N_h + 4 N_g
    each, the second writing d1 d2), the r2c of d1 d2 (4 N_g + 8 N_h) and one read of the two spectra's
    columns kz < K by the theta-shell mode sums (2 x 8 N^2 K)."""
    Ng = N ** 3
    Nh = N * N * (N // 2 + 1)
    return 12 * n_p + 4 * Ng + 2 * (4 * Ng + 8 * Nh) + 2 * (8 * Nh + 4 * Ng) + 16 * N * N * K


def cpu_leg(pos_host, N, L, axis, gpu_pk, threads):
    """Reference CPU path timed on this host + FP64 oracle error of the GPU result (checker only)."""
    from oracle import cpu as C

    out = {}
    if C.ref_available():
        res, t = C.cpu_reference_pipeline(pos_host, N, L, axis, threads)
        out["cpu_baseline"] = {
            "value": round(len(pos_host) / t["total"] / 1e6, 3),
            "unit": "Mparticles/s",
            "cores": threads,
            "kind": "port",
            "sample": ("full workload once: reference MAS_c.c CIC (OpenMP, %d threads) %.2fs + float32 "
                       "normalise + scipy.fft r2c (%d workers; FFTW absent) %.2fs + serial C restatement of "
                       "the Pk mode loop %.2fs" % (threads, t["deposit"], threads, t["normalise_fft"], t["modes"])),
        }
        # the reference's scatter on one thread (its Cython MA loop is serial), on the first eighth
        import time
        sub = np.ascontiguousarray(pos_host[: len(pos_host) // 8])
        grid = np.zeros((N, N, N), dtype=np.float32)
        t0 = time.perf_counter()
        C.ref_mas_c(sub, grid, L, "CIC", None, 1)
        ts = time.perf_counter() - t0
        out["cpu_serial_deposit"] = {"value": round(len(sub) / ts / 1e6, 3), "unit": "Mparticles/s", "cores": 1,
                                     "kind": "reference",
                                     "sample": "reference MAS_c.c CIC with threads=1 on the first %d particles "
                                               "(deposit only)" % len(sub)}
    o, fl = C.fp64_oracle_pipeline(pos_host, N, L, axis, workers=threads, with_floor=True)
    P0 = o.Pk[:, 0]
    err = {
        "P0": float(np.max(np.abs(gpu_pk.Pk[:, 0] - P0) / np.abs(P0))),
        "P2_over_P0": float(np.max(np.abs(gpu_pk.Pk[:, 1] - o.Pk[:, 1]) / np.abs(P0))),
        "P4_over_P0": float(np.max(np.abs(gpu_pk.Pk[:, 2] - o.Pk[:, 2]) / np.abs(P0))),
        "Nmodes3D_exact": bool(np.array_equal(gpu_pk.Nmodes3D, o.Nmodes3D)),
        "k3D": float(np.max(np.abs(gpu_pk.k3D - o.k3D) / o.k3D)),
        "bins_P0_within_1e-5": float(np.mean(np.abs(gpu_pk.Pk[:, 0] - P0) / np.abs(P0) <= 1e-5)),
        "bins_P2P4_within_1e-5_of_P0": float(np.mean(np.abs(gpu_pk.Pk[:, 1:] - o.Pk[:, 1:]) / np.abs(P0)[:, None] <= 1e-5)),
        "f32_floor_P0": float(np.max(np.abs(fl.Pk[:, 0] - P0) / np.abs(P0))),
        "f32_floor_P4_over_P0": float(np.max(np.abs(fl.Pk[:, 2] - o.Pk[:, 2]) / np.abs(P0))),
    }
    if "cpu_baseline" in out:
        err["cpu_reference_f32_P0"] = float(np.max(np.abs(res.Pk[:, 0] - P0) / np.abs(P0)))
    out["err_vs_fp64_oracle"] = err
    return out


def _timed(fn, steps, warmup):
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(steps):
        fn()
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / steps


def workload_batch(args):
    """64 seeded realizations: N_p log-uniform in [64^3, 256^3], N in {128,192,256,384},
    L in [250, 1000], uniform positions, CIC + P0, one hipGraph launch per step."""
    import torch

    from pylians_amd import batch, synth
    jobs = synth.batch_jobs(args.batch, seed=100)
    b = batch.PkBatch(jobs, flags=0)
    for _ in range(args.warmup):
        b.run()
    b.stream.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    with torch.cuda.stream(b.stream):
        e0.record()
        for _ in range(args.steps):
            b.run()
        e1.record()
    b.stream.synchronize()
    ms = e0.elapsed_time(e1) / args.steps
    b.results()   # mode check of every job on the last replay
    n_tot = b.n_particles
    bytes_e2e = sum(e2e_bytes(int(j["pos"].shape[0]), j["dims"]) for j in jobs)
    b.close()
    return {"metric": "Mparticles/s batched CIC+P0, %d uneven realizations in one hipGraph launch" % args.batch,
            "value": round(n_tot / (ms * 1e-3) / 1e6, 1), "unit": "Mparticles/s", "ms_per_step": round(ms, 4),
            "grids_per_s": round(args.batch / (ms * 1e-3), 1), "bytes": bytes_e2e,
            "e2e_hbm_frac": round(bytes_e2e / (ms * 1e-3) / (HBM_PEAK_GBS * 1e9), 4),
            "config": {"workload": "batch%d_uniform_cic_p0" % args.batch, "n_particles_total": n_tot,
                       "grids": sorted(set(j["dims"] for j in jobs))}, "dtype": "f32", "steps": args.steps}


def workload_adjoint(args):
    """256^3 Zel'dovich particles (no RSD), CIC onto 256^3, L = 500: forward P0 + dLoss/dpos."""
    import torch

    from pylians_amd import autograd, synth
    N, L = 256, 500.0
    pos = synth.zeldovich_device(256, L, seed=3, rms_disp=5.0)
    w = torch.tensor(np.random.default_rng(3).random(64), device="cuda")
    ms = _timed(lambda: autograd.grad_loss(pos, L, N, w), args.steps, args.warmup)
    n = int(pos.shape[0])
    nbytes = adjoint_bytes(n, N)
    del pos
    return {"metric": "Mparticles/s forward P0 + adjoint dL/dpos, 256^3 CIC", "value": round(n / (ms * 1e-3) / 1e6, 1),
            "unit": "Mparticles/s", "ms_per_step": round(ms, 4), "dtype": "f32", "steps": args.steps,
            "bytes": nbytes, "e2e_hbm_frac": round(nbytes / (ms * 1e-3) / (HBM_PEAK_GBS * 1e9), 4),
            "config": {"workload": "zeldovich256_cic_adjoint_64bins", "n_particles": n, "grid": N, "BoxSize": L}}


def workload_xpk(args):
    """XPk of two 512^3 float32 fields resident in HBM (L = 1000, axis 0, CIC windows): both auto and the
    cross spectrum with every output (Pk_library.pyx:534-798), one whole call each."""
    import torch

    from pylians_amd import Pk_library
    N, L = 512, 1000.0
    g = torch.Generator(device="cuda")
    g.manual_seed(5)
    a = torch.randn((N, N, N), device="cuda", dtype=torch.float32, generator=g)
    b = (0.5 * a + torch.randn((N, N, N), device="cuda", dtype=torch.float32, generator=g)).contiguous()
    ms = _timed(lambda: Pk_library.XPk([a, b], L, axis=0, MAS=["CIC", "CIC"]), args.steps, args.warmup)
    Ng, Nh = N ** 3, N * N * (N // 2 + 1)
    nbytes = 2 * (4 * Ng + 8 * Nh + 8 * Nh)   # per field: r2c (read grid, write spectrum), binning read
    del a, b
    return {"metric": "XPk of two 512^3 fields (auto + cross, all outputs)", "value": round(ms, 3),
            "unit": "ms/call", "ms_per_step": round(ms, 4), "higher_is_better": False, "dtype": "f32",
            "bytes": nbytes, "e2e_hbm_frac": round(nbytes / (ms * 1e-3) / (HBM_PEAK_GBS * 1e9), 4),
            "config": {"workload": "xpk_two_fields_grid512", "fields": 2, "grid": N, "BoxSize": L}}


def workload_bispectrum(args):
    """Realizations of 512^3 Zel'dovich particles (resident in HBM), each: CIC onto a 256^3 grid,
    delta = rho/mean - 1, Bk(k1=0.1, k2=0.2, 20 theta) -- the deposit is inside the timed region."""
    import torch

    from pylians_amd import Pk_library, _lib, synth
    N, L, nth = 256, 1000.0, 20
    theta = np.linspace(0, np.pi, nth)
    grid = torch.empty((N, N, N), dtype=torch.float32, device="cuda")
    parts = [synth.zeldovich_device(512, L, seed=200 + r, rms_disp=5.0) for r in range(args.realizations)]

    def one(pos):
        _lib.check(_lib.lib().pyl_mas_overdensity(_lib.ptr(pos), None, int(pos.shape[0]), N, L, 2, 0,
                                                  _lib.ptr(grid), _lib.stream()))
        return Pk_library.Bk(grid, L, 0.1, 0.2, theta)

    one(parts[0])       # builds the cached indicator sums
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for pos in parts:
        one(pos)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / len(parts)
    n = int(parts[0].shape[0])
    nbytes = bk_bytes(n, N, nth)
    # the theta shells reach |k| < k1 + k2 + kF (kF units): columns kz < K of the spectra
    kF = 2.0 * np.pi / L
    K = min(N // 2 + 1, int(np.floor((0.1 + 0.2 + kF) / kF)) + 2)
    own = bk_bytes_parseval(n, N, K)
    del parts
    return {"metric": "Bk per realization (512^3 particles CIC onto 256^3, 20 theta bins, indicator sums cached)",
            "value": round(ms, 3), "unit": "ms/realization", "higher_is_better": False, "dtype": "f32",
            "bytes": nbytes, "e2e_hbm_frac": round(nbytes / (ms * 1e-3) / (HBM_PEAK_GBS * 1e9), 4),
            "bytes_parseval": own, "hbm_frac_parseval": round(own / (ms * 1e-3) / (HBM_PEAK_GBS * 1e9), 4),
            "note": "bytes: the estimator's 22 shell c2r and 20 triple products (the reference algorithm); this "
                    "build sums the theta shells over their modes by Parseval after one r2c of d1 d2, and "
                    "bytes_parseval is its own compulsory traffic",
            "config": {"workload": "bk_zeldovich512_grid256_20theta", "realizations": args.realizations}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n-side", type=int, default=512, help="particles per side of the lattice")
    ap.add_argument("--grid", type=int, default=512)
    ap.add_argument("--mas", default="CIC")
    ap.add_argument("--deposit", default=None, choices=[None, "atomic", "binned"])
    ap.add_argument("--fp64-steps", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline and the FP64 oracle check")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="before the timed region, copy a 512 MiB device buffer twice (a known-size kernel for "
                         "scaling rocprofv3 FETCH_SIZE / WRITE_SIZE)")
    ap.add_argument("--workload", default="zeldovich512", choices=["zeldovich512", "batch", "adjoint", "bispectrum",
                                                                   "xpk"])
    ap.add_argument("--batch", type=int, default=64)
    ap.add_argument("--realizations", type=int, default=16)
    ap.add_argument("--no-extra", action="store_true",
                    help="default workload only: skip the batch, adjoint and bispectrum figures")
    ap.add_argument("--only-fp32", action="store_true",
                    help="profiling runs: time the float32 step only (no float64 path, host drop-in, other "
                         "workloads or CPU leg), so that the last --steps launches of each kernel are the timed steps")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one MI355X: this build has no multi-GPU path (run with --gpus 1)", file=sys.stderr)
        sys.exit(2)
    if args.workload != "zeldovich512":
        fn = {"batch": workload_batch, "adjoint": workload_adjoint, "bispectrum": workload_bispectrum,
              "xpk": workload_xpk}[args.workload]
        print(json.dumps(fn(args)), flush=True)
        return

    import torch

    from pylians_amd import _lib, synth
    from pylians_amd.Pk_library import pk_layout, unpack_pk
    from pylians_amd.pipeline import run_device

    _lib.require_gpu()
    if args.calibrate_copy:
        src = torch.ones(128 << 20, device="cuda", dtype=torch.float32)
        dst = torch.empty_like(src)
        for _ in range(2):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    n_side, N, L, axis = args.n_side, args.grid, 1000.0, 0
    scheme = _lib.SCHEMES[args.mas]
    t0 = time.time()
    pos = synth.zeldovich_device(n_side, L, seed=2, rms_disp=5.0, kc=0.5, sigma_v=300.0, axis=axis)
    torch.cuda.synchronize()
    n_p = int(pos.shape[0])
    print("# generated %d particles in %.1fs" % (n_p, time.time() - t0), file=sys.stderr, flush=True)
    _, _, _, nout = pk_layout(N)
    out = torch.empty(nout, dtype=torch.float64, device="cuda")
    # PK_ASYNC: no host synchronisation inside the timed loop; the mode counts the kernel returns are
    # checked after it (check_number_modes, as the reference does on its Nmodes3D)
    flags = _lib.PK_ALL | _lib.PK_ASYNC | {None: 0, "atomic": _lib.ATOMIC, "binned": _lib.BINNED}[args.deposit]

    def step(prec):
        run_device(pos, None, out, L, N, scheme, axis, prec, flags)

    # ---- float32-compensated path: warmup, then the timed region with stage events
    for _ in range(args.warmup):
        step(_lib.F32)
    torch.cuda.synchronize()
    _lib.profile(True)
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    w0 = time.perf_counter()
    ev0.record()
    for _ in range(args.steps):
        step(_lib.F32)
    ev1.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - w0
    stages = _lib.profile_read()
    _lib.profile(False)
    ms = ev0.elapsed_time(ev1) / args.steps
    res32 = out.cpu().numpy().copy()

    class R:
        pass

    gpu_pk = R()
    unpack_pk(gpu_pk, res32, N)
    from pylians_amd.Pk_library import check_number_modes
    check_number_modes(np.concatenate([[1.0], gpu_pk.Nmodes3D]), N)
    print("# fp32c: %.3f ms/step (wall %.3f ms/step)" % (ms, wall * 1e3 / args.steps), file=sys.stderr, flush=True)

    if args.only_fp32:
        print(json.dumps({"metric": METRIC, "value": round(n_p / (ms * 1e-3) / 1e6, 1), "unit": "Mparticles/s",
                          "ms_per_step": round(ms, 4), "steps": args.steps,
                          "stage_ms": {k: round(v[1] / v[0], 4) for k, v in stages.items() if v[0] > 0}}), flush=True)
        return

    # per-step times (a second pass of max(K, 10) steps, one event pair each): the median step
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
           for _ in range(max(args.steps, 10))]
    for a, b in evs:
        a.record()
        step(_lib.F32)
        b.record()
    torch.cuda.synchronize()
    ms_med = float(np.median([a.elapsed_time(b) for a, b in evs]))

    # ---- plain float64 GPU path (same particles, float64 grid, FFT and sums)
    step(_lib.F64)
    torch.cuda.synchronize()
    ev0.record()
    for _ in range(args.fp64_steps):
        step(_lib.F64)
    ev1.record()
    torch.cuda.synchronize()
    ms64 = ev0.elapsed_time(ev1) / args.fp64_steps
    res64 = R()
    unpack_pk(res64, out.cpu().numpy(), N)
    check_number_modes(np.concatenate([[1.0], res64.Nmodes3D]), N)
    print("# fp64: %.3f ms/step" % ms64, file=sys.stderr, flush=True)

    # ---- host drop-in (numpy positions in, Pk attributes out: PCIe both ways); reported beside the
    # value, never as it
    from pylians_amd.pipeline import PkParticles
    pos_host = pos.cpu().numpy()
    PkParticles(pos_host, L, N, MAS=args.mas, axis=axis)
    torch.cuda.synchronize()
    h0 = time.perf_counter()
    for _ in range(3):
        PkParticles(pos_host, L, N, MAS=args.mas, axis=axis)
    torch.cuda.synchronize()
    host_ms = (time.perf_counter() - h0) * 1e3 / 3
    print("# host drop-in: %.2f ms/step" % host_ms, file=sys.stderr, flush=True)

    # ---- roofline of the dominant kernel (the stage with the longest average launch)
    ab = algorithmic_bytes(n_p, N)
    per = {k: (v[1] / v[0]) for k, v in stages.items() if v[0] > 0}
    dom = max((k for k in per if k in ab), key=lambda k: per[k])
    achieved = ab[dom] / (per[dom] * 1e-3) / 1e9
    workload = "zeldovich%d_rsd_%s_pk_grid%d" % (n_side, args.mas.lower(), N)
    traffic = measured_traffic(workload) if args.deposit is None else {}
    roofline = {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic.get(dom),
                "bytes_per_launch": ab[dom], "avg_ms": round(per[dom], 4),
                "traffic_by_stage": traffic or None,
                "traffic_over_algorithmic": {k: round(traffic[k] / ab[k], 3) for k in traffic if k in ab} or None}
    stage_ms = {k: round(v, 4) for k, v in per.items()}
    stage_frac = {k: round(ab[k] / (per[k] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4) for k in per if k in ab}

    line = {
        "metric": METRIC, "value": round(n_p / (ms * 1e-3) / 1e6, 1), "unit": "Mparticles/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 4),
        "median_ms_per_step": round(ms_med, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (Zel'dovich lattice + RSD, seeded, generated on the GPU)",
        "config": {"workload": workload, "n_particles": n_p,
                   "grid": N, "BoxSize": L, "MAS": args.mas, "axis": axis, "outputs": "P0/P2/P4, 1-D, 2-D, phase",
                   "deposit": args.deposit or "auto"},
        "fp64_ms_per_step": round(ms64, 4), "speedup_fp32c_vs_fp64": round(ms64 / ms, 3),
        "grids_per_s": round(1e3 / ms, 3),
        "e2e_bytes": ab["e2e"], "e2e_hbm_frac": round(ab["e2e"] / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
        "stage_ms": stage_ms, "stage_hbm_frac": stage_frac, "roofline": roofline,
        "fp64_vs_fp32c_P0_maxrel": float(np.max(np.abs(res64.Pk[:, 0] - gpu_pk.Pk[:, 0]) / np.abs(res64.Pk[:, 0]))),
        "host_dropin": {"ms_per_step": round(host_ms, 2), "Mparticles_s": round(n_p / (host_ms * 1e-3) / 1e6, 1),
                        "note": "PkParticles on a pageable numpy array: positions host->device and results back"},
    }
    del pos
    torch.cuda.empty_cache()
    if not args.no_extra:
        # the other workloads of the benchmark set, each with its algorithmic bytes and HBM fraction
        import copy
        ex = copy.copy(args)
        ex.steps, ex.warmup, ex.realizations = 5, 2, 4
        extra = {}
        for key, fn in (("batch64", workload_batch), ("adjoint256", workload_adjoint),
                        ("bispectrum256", workload_bispectrum), ("xpk512", workload_xpk)):
            t1 = time.time()
            r = fn(ex)
            extra[key] = {k: r[k] for k in ("value", "unit", "ms_per_step", "bytes", "e2e_hbm_frac", "bytes_parseval",
                                            "hbm_frac_parseval", "note", "config") if k in r}
            print("# %s %.1fs" % (key, time.time() - t1), file=sys.stderr, flush=True)
            torch.cuda.empty_cache()
        line["workloads"] = extra
    if not args.no_cpu:
        threads = len(os.sched_getaffinity(0))
        t1 = time.time()
        line.update(cpu_leg(pos_host, N, L, axis, gpu_pk, threads))
        print("# cpu leg %.1fs" % (time.time() - t1), file=sys.stderr, flush=True)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
