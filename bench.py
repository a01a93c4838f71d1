#!/usr/bin/env python3
"""Headline benchmark: posterior cells/s of the FFT-domain Bayesian AVO inversion on one MI355X.

Workload (one step = one inversion, inputs resident in HBM when the timed region starts):
  single cube 512 x 512 x 256 (padded 576 x 576 x 320), dx = dy = 25 m, dt = 4 ms, 5 angles
  0-40 deg, 30 Hz Ricker with per-angle scaling U(0.8, 1.2), S/N per angle U(3, 8); prior
  ln(3000, 1500, 2.3), std (0.10, 0.15, 0.05), correlations (0.7, 0.4, 0.3), spherical 1000 m
  lateral and exponential 40 ms temporal correlation; a seeded Gaussian elastic field drawn from
  the prior and forward-modelled to seismic on the GPU.
  Step: one cg_invert call -- 3 + 5 forward FFTs, per-wavenumber posterior (FP32, FP64 re-solve
  where the pivot ratio of the Cholesky factor exceeds 1e6), 3 inverse FFTs of the posterior
  means, at this geometry with the z transform split into two stages carried by the x passes and
  the posterior pass -- then 2 posterior realisations (counter-based noise, colouring, 3 inverse
  FFTs + the real-domain mean each).

Prints one JSON line.  Also reported: per-kernel HIP-event times, the roofline of the dominant
kernel, p50/p99 step latency, and the reference's own CPU code (FFTW 2.1.2, lib_matr,
RandomGen compiled from the CRAVA sources) timed on a bounded sample on this host.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import crava_amd as C  # noqa: E402
from crava_amd import synthetic as S  # noqa: E402

HBM_PEAK = 8.0e12  # MI355X HBM3E, bytes/s


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def kernel_bytes(w, nt, R, n_inv_planes, launches, split):
    """Algorithmic HBM bytes per kernel class for one step (what each kernel must read + write).
    `launches`: launches per step of each FFT class, as counted on the context stream -- every FFT
    launch transforms one grid (the realisations' inverse z passes run inside the fused colouring
    and are not z launches).  split: the request ran the split-z passes (cg_invert) -- the x
    passes carry the z stages, the posterior pass reads and writes whole grids, and the
    colouring adds the means in their split layout; otherwise the drop-in sequence (the
    posterior reads the spectra of inverted planes only, the colouring adds the spectral
    means).  Either way a pair of realisations reads the means once."""
    G = (w["nxp"] // 2 + 1) * w["nyp"] * w["nzp"] * 8  # one complex grid
    Nr = w["nxp"] * w["nyp"] * w["nzp"] * 4  # one real cube (without row padding)
    plane = (w["nxp"] // 2 + 1) * w["nyp"]
    if split:
        post = (3 + nt + 3 + 6) * G  # z stage 2 needs every plane of the inputs and of the means
    else:
        post = n_inv_planes * plane * (3 + nt + 9) * 8 + (w["nzp"] - n_inv_planes) * plane * 6 * 8
    return {
        # x: forward reads the real cube and writes the spectrum, inverse the reverse
        C.K_FFT_X: launches[C.K_FFT_X] * (Nr + G),
        # y / z: one grid read and written per launch
        C.K_FFT_Y: launches[C.K_FFT_Y] * 2 * G,
        C.K_FFT_Z: launches[C.K_FFT_Z] * 2 * G,
        # split: read 3 + nt grids, write 3 means + 6 covariance grids; drop-in: inverted cells read
        # 3 + nt spectra and write 3 means + 6 covariance entries, pass-through cells write the 6
        # prior covariance entries
        C.K_POSTERIOR: post,
        # colouring: read 6 covariance and 3 mean spectra, write 3 per realisation (noise
        # generated in registers); pairs read their inputs once
        C.K_NOISE: (R // 2) * 15 * G + (R % 2) * 12 * G,
    }


NAMES = {C.K_FFT_X: "fft_x", C.K_FFT_Y: "fft_y", C.K_FFT_Z: "fft_z", C.K_POSTERIOR: "posterior",
         C.K_NOISE: "colour", C.K_ELEMENTWISE: "elementwise"}


def _on_threads(fn, T):
    """Runs fn(i), i < T, on T threads at once and returns the wall time (the reference's code is
    called through ctypes, which releases the interpreter lock, so the threads run in parallel)."""
    import threading
    ths = [threading.Thread(target=fn, args=(i,)) for i in range(T)]
    t = time.perf_counter()
    for th in ths:
        th.start()
    for th in ths:
        th.join()
    return time.perf_counter() - t


def cpu_baseline(w, inv, R, threads=1):
    """The reference's own code (oracle/_ref: FFTW 2.1.2, lib_matr, RandomGen) on `threads` host
    cores, on a bounded sample, extrapolated to the reference path for this workload: 15 forward
    FFTs (3 means, 5 seismic, 6 prior covariance, errCorr), 9 inverse (3 means, 6 covariance), the
    posterior loop, and per realisation 3 noise fills, the colouring loop and 3 inverse FFTs.  With
    threads > 1 (the reference itself is single-threaded; this is the all-cores figure), every
    sampled piece runs once per thread at the same time on its own data and counts as that many
    pieces done; the noise fill stays serial (the reference's RandomGen is one process-wide
    sequential stream)."""
    import oracle as O
    if not O.ref_available():
        return None
    R_ = O.ref()
    T = max(1, threads)
    nxp, nyp, nzp = w["nxp"], w["nyp"], w["nzp"]
    cnxp = nxp // 2 + 1
    plane = cnxp * nyp
    rng = np.random.default_rng(1)
    xs = [rng.standard_normal((nzp, nyp, 2 * cnxp)).astype(np.float32) for _ in range(T)]
    t_fwd = _on_threads(lambda i: R_.ref_grid_fft(nxp, nyp, nzp, O.DATA, O._ptr(xs[i], O.ctypes.c_float)), T) / T
    t_inv = _on_threads(lambda i: R_.ref_grid_ifft(nxp, nyp, nzp, O.PARAMETER, O._ptr(xs[i], O.ctypes.c_float)),
                        T) / T
    del xs
    # posterior loop on NPL planes of inverted and NPL planes of pass-through wavenumbers, with the
    # workload's tables, in the reference's covariance-grid form
    NPL = 4  # planes sampled per kind
    tb = inv.tables
    inv_k = np.where(tb["invert"] != 0)[0]
    pass_k = np.where(tb["invert"] == 0)[0]
    times = {}
    cov_post = None
    for label, ks in (("inv", inv_k[np.linspace(0, len(inv_k) - 1, NPL).astype(int)]),
                      ("pass", pass_k[np.linspace(0, len(pass_k) - 1, NPL).astype(int)])):
        p = dict(nxp=nxp, nyp=nyp, nzp=NPL, ntheta=inv.nt, K=tb["K"][ks], e1=tb["e1"][ks], e2=tb["e2"][ks],
                 e3=tb["e3"][ks], invert=tb["invert"][ks], sigma0e=tb["sigma0e"], wnc=tb["wnc"],
                 prior_var0=tb["prior_var0"], prior_mode=0, rxy=None, rt=None, lam2d=None)
        shape = (NPL, nyp, cnxp)
        r = np.abs(inv.rxy[None] * inv.rt[ks][:, None, None])
        S0 = tb["prior_var0"]
        covs = np.stack([(r * S0[a, b]).astype(np.complex64) for a, b in
                         ((0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2))])
        p["lam_grid"] = np.broadcast_to(inv.rxy[None], shape).astype(np.complex64)
        mean = (rng.standard_normal((3,) + shape) + 1j * rng.standard_normal((3,) + shape)).astype(np.complex64)
        data = (rng.standard_normal((inv.nt,) + shape) + 1j * rng.standard_normal((inv.nt,) + shape)).astype(np.complex64)
        out = [None] * T
        wall = _on_threads(lambda i: out.__setitem__(i, O.ref_posterior(p, mean, data, covs)), T)
        times[label] = wall / (T * NPL * plane)
        if label == "inv":
            cov_post = out[0][2]
    seeds = [(rng.standard_normal((3,) + cov_post.shape[1:]) + 0j).astype(np.complex64) for _ in range(T)]
    t_col = _on_threads(lambda i: O.ref_colour(cov_post, seeds[i]), T) / (T * NPL * plane)
    R_.ref_rng_seed(23665)
    noise = np.zeros((NPL, nyp, cnxp), dtype=np.complex64)
    t = time.perf_counter()
    R_.ref_fill_complex_noise(nxp, nyp, NPL, O._ptr(noise.view(np.float32), O.ctypes.c_float))
    t_noise = (time.perf_counter() - t) / (NPL * plane)
    n_inv = inv.n_inverted_planes * plane
    n_all = nzp * plane
    total = (15 * t_fwd + 9 * t_inv + n_inv * times["inv"] + (n_all - n_inv) * times["pass"]
             + R * (3 * n_all * t_noise + n_all * t_col + 3 * t_inv))
    cells = w["nx"] * w["ny"] * w["nz"]
    each = "" if T == 1 else ", %d at a time on %d threads (noise fill serial)" % (T, T)
    return {
        "value": cells / total, "unit": "cells/s", "cores": T, "kind": "reference",
        "seconds_per_inversion": total,
        "sample": ("reference FFTW 2.1.2 rfftw3d forward (%.2f s) and inverse (%.2f s) of one 576x576x320 grid; "
                   "reference lib_matr posterior loop on %d inverted + %d pass-through k planes (%.2f / %.3f us "
                   "per wavenumber); lib_matr colouring and RandomGen noise fill on %d planes%s; extrapolated to "
                   "15 forward + 9 inverse FFTs, the full loop and %d realisations" %
                   (t_fwd, t_inv, NPL, NPL, 1e6 * times["inv"], 1e6 * times["pass"], NPL, each, R)),
    }


TRAFFIC = os.path.join(ROOT, "profiles", "single_cube_576x576x320_traffic.json")


def measured_traffic(kernel):
    """HBM bytes per launch of a kernel class on this workload, from the rocprofv3 FETCH_SIZE /
    WRITE_SIZE runs summarised in profiles/ (scaled by a copy kernel of known traffic)."""
    try:
        with open(TRAFFIC) as f:
            return json.load(f)["bytes_per_launch"].get(kernel)
    except (OSError, ValueError, KeyError):
        return None


ACCURACY = os.path.join(ROOT, "profiles", "accuracy_configs.json")


def accuracy_summary():
    """Per-output relative L2 of the GPU results against the FP64 oracle (and the reference's own FP32
    pipeline for the plain case) for each workload, as measured by tests/test_gpu_configs.py on the
    GPU with CG_ACCURACY_OUT set (summary committed under profiles/)."""
    try:
        with open(ACCURACY) as f:
            d = json.load(f)
        return {"source": d.get("command"), "unit": d.get("unit"), "workloads": d.get("workloads")}
    except (OSError, ValueError):
        return None


def rel_l2(a, b):
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-30))


def bench_mix(ctx, n=64, reps=3):
    """64 independent inversions of uneven sizes (posterior mean + covariance each): the plain
    per-request path (each request's launches, then a synchronisation) against one batch of all
    of them (every pass of every cube in shared launches, replayed as a captured graph)."""
    specs = S.uneven_mix(n)
    t0 = time.perf_counter()
    probs = [S.Problem(ctx, s) for s in specs]
    for shape in sorted({(s["nxp"], s["nyp"], s["nzp"]) for s in specs}):
        ctx.autotune(*shape)  # both paths use the tuned tiles
    log("mix setup %.1f s (%d cubes, %.1f GB)" % (time.perf_counter() - t0, n,
                                                 sum(S.footprint(s) for s in specs) / 1e9))
    cells = sum(S.cells(s) for s in specs)
    for p in probs:  # warm: twiddle tables and plans of every size
        p.step(realisations=0)
    ctx.sync()
    plain = []
    for _ in range(reps):
        t = time.perf_counter()
        for p in probs:
            p.step(realisations=0)
            ctx.sync()
        plain.append(time.perf_counter() - t)
    plain_means = [np.stack([g.download_real() for g in p.mean_w]) for p in probs]
    b = C.Batch(ctx, [p.item(realisations=0) for p in probs])
    b.use_graph(True)
    b.run()  # captures the graph
    b.wait()
    batched = []
    for _ in range(reps):
        t = time.perf_counter()
        b.run()
        b.wait()
        batched.append(time.perf_counter() - t)
    err = 0.0
    for p, pm in zip(probs, plain_means):
        bm = np.stack([g.download_real() for g in p.mean_w])
        bg = np.asarray(S.BACKGROUND, dtype=np.float32)[:, None, None, None]
        err = max(err, rel_l2(bm - bg, pm - bg))
    b.close()
    for p in probs:
        p.close()
    tp, tb = min(plain), min(batched)
    sizes = sorted((s["nx"] * s["ny"] * s["nz"], s["nxp"], s["nyp"], s["nzp"]) for s in specs)
    return {
        "workload": "%d inversions, sizes log-uniform 32x32x64..256x256x192 (padded %dx%dx%d..%dx%dx%d by the "
                    "reference rule), 3-6 angles, peak U(20,40) Hz, std U(0.05,0.2), lateral range "
                    "U(300,2000) m; posterior mean + covariance, FP32 with FP64 re-solve where cond > 1e6"
                    % ((n,) + sizes[0][1:] + sizes[-1][1:]),
        "cells": cells,
        "plain_ms": round(1000 * tp, 3), "batched_ms": round(1000 * tb, 3),
        "plain_cells_per_s": cells / tp, "batched_cells_per_s": cells / tb, "speedup": tp / tb,
        "max_rel_l2_mean_batched_vs_plain": err,
        "plain": "each request: its FFT/posterior launches on one stream, then a synchronisation",
        "batched": "one cg_batch of all requests, replayed as a captured HIP graph",
    }


def bench_input_padding(ctx, inv, w, reps=5):
    """The padding fills on the input side at the workload's size, on scratch grids of `inv`
    (its realisation outputs, rewritten by every step): the background cube fill (nx*ny*nz
    floats read from HBM, the padded grid written) and the seismic trace interpolation with 5%
    of the traces zeroed (energies, host walk, trace averages, tapered extrapolation)."""
    src, dst, seis = inv.out[0]
    nx, ny, nz = w["nx"], w["ny"], w["nz"]
    dst.fill_padded(src.device_ptr())  # src's memory as a device cube of nz*ny*nx floats; warm-up
    ctx.sync()
    ctx.set_profiling(True)
    ctx.reset_kernel_time()
    for _ in range(reps):
        dst.fill_padded(src.device_ptr())
    n, ms = ctx.kernel_time(C.K_ELEMENTWISE)
    ctx.set_profiling(False)
    fill_bytes = 4 * (nx * ny * nz + dst.rnxp * w["nyp"] * w["nzp"])
    fill_gbps = fill_bytes / (ms * 1e-3 / n) / 1e9
    rv = inv.seis_in[0].download_real(padded=True)
    rng = np.random.default_rng(3)
    bad = rng.random((ny, nx)) < 0.05
    rv[:, :ny, :nx][:, bad] = 0.0
    seis.setType(C.DATA)
    seis.upload_real(rv)
    seis.interpolateSeismic(0.01)  # warm-up, then the same input again
    seis.upload_real(rv)
    ctx.set_profiling(True)
    ctx.reset_kernel_time()
    t = time.perf_counter()
    nbad, _ = seis.interpolateSeismic(0.01)
    ctx.sync()
    wall = 1000.0 * (time.perf_counter() - t)
    kn, kms = ctx.kernel_time(C.K_ELEMENTWISE)
    ctx.set_profiling(False)
    seis.setType(C.PARAMETER)
    # one grid through a CRAVA binary file and back (byte order swapped on the device)
    import tempfile
    geo = dict(x0=0.0, y0=0.0, dx=w["dx"], dy=w["dy"], nx=nx, ny=ny, il0=1.0, xl0=1.0, il_step_x=1.0,
               il_step_y=0.0, xl_step_x=0.0, xl_step_y=1.0, angle=0.0)
    with tempfile.TemporaryDirectory() as tmp:
        name = os.path.join(tmp, "grid")
        t = time.perf_counter()
        dst.writeCravaFile(name, geo)
        t_write = time.perf_counter() - t
        t = time.perf_counter()
        src.readCravaFile(name + ".crava")
        t_read = time.perf_counter() - t
    file_bytes = 4 * dst.rnxp * w["nyp"] * w["nzp"]
    # prior covariance grid fill (fillInParamCorr + multiplyByScalar) with a correlation gradient
    rng = np.random.default_rng(4)
    cxy = rng.random((w["nyp"], w["nxp"])).astype(np.float32)
    ct = rng.random(w["nzp"]).astype(np.float32)
    corr_us = {}
    for grad in ((0.0, 0.0), (0.1, 0.05)):
        dst.fillInParamCorr(cxy, ct, grad[0], grad[1], 0.01)  # warm-up
        ctx.sync()
        ctx.set_profiling(True)
        ctx.reset_kernel_time()
        for _ in range(reps):
            dst.fillInParamCorr(cxy, ct, grad[0], grad[1], 0.01)
        pn, pms = ctx.kernel_time(C.K_ELEMENTWISE)
        ctx.set_profiling(False)
        corr_us[grad] = 1000.0 * pms / pn
    # CommonData::FillInData's seismic branch: one trace per box column (nz + 2 x 25 samples of
    # 4 ms, a 100 ms guard zone), the padding columns from their fill cells, onto the padded grid
    from crava_amd import seismic as SE
    ns_tr = nz + 50
    tr = rng.standard_normal((nx * ny, ns_tr)).astype(np.float32)
    ri = np.array([-1 if SE.fill_number(i, nx, w["nxp"]) is None else SE.fill_number(i, nx, w["nxp"])
                   for i in range(dst.rnxp)])
    rj = np.array([SE.fill_number(j, ny, w["nyp"]) for j in range(w["nyp"])])
    col = np.where(ri[None, :] >= 0, rj[:, None] * nx + ri[None, :], -1).astype(np.int32)
    z0g = np.full(col.shape, 100.0 + 2.0, np.float32)
    dzg = np.full(col.shape, 4.0, np.float32)
    z0d = np.zeros(nx * ny, np.float32)
    seis.setType(C.DATA)
    seis.fillInData(list(tr), z0d, 4.0, col, z0g, dzg, 100.0)  # warm-up
    ctx.sync()
    ctx.set_profiling(True)
    ctx.reset_kernel_time()
    t = time.perf_counter()
    seis.fillInData(list(tr), z0d, 4.0, col, z0g, dzg, 100.0)
    wall_tr = 1000.0 * (time.perf_counter() - t)
    tk = sum(ctx.kernel_time(k)[1] for k in (C.K_ELEMENTWISE, C.K_FFT_X))
    ctx.set_profiling(False)
    seis.setType(C.PARAMETER)
    nt_tr = C.lib().cg_factorable(ns_tr)
    return {
        "fill_from_traces": {"what": "CommonData::FillInData (SEG-Y seismic branch): %d traces of %d samples "
                                     "(guard-zone taper, ten-fold spectral resampling: R2C of %d, C2R of %d) "
                                     "onto %dx%dx%d" % (nx * ny, ns_tr, nt_tr, 10 * nt_tr, w["nxp"], w["nyp"], w["nzp"]),
                             "ms_synchronised_incl_host_copies": round(wall_tr, 2), "kernel_ms": round(tk, 3)},
        "fill_padded": {"what": "FFTGrid(NRLib::Grid) background fill, %dx%dx%d cube into %dx%dx%d"
                                % (nx, ny, nz, w["nxp"], w["nyp"], w["nzp"]),
                        "avg_launch_us": round(1000.0 * ms / n, 2), "bytes_per_launch": fill_bytes,
                        "gbps": round(fill_gbps, 1), "frac_of_peak": fill_gbps * 1e9 / HBM_PEAK},
        "interpolate_seismic": {"what": "FFTGrid::interpolateSeismic, 5%% of %d traces zeroed" % (nx * ny),
                                "bad_traces": nbad, "ms_synchronised": round(wall, 2),
                                "kernel_ms": round(kms, 3), "launches": kn},
        "fill_param_corr": {"what": "FFTGrid::fillInParamCorr + multiplyByScalar, %dx%dx%d (one grid written)"
                                    % (w["nxp"], w["nyp"], w["nzp"]), "bytes_per_launch": file_bytes,
                            "no_gradient": {"avg_launch_us": round(corr_us[(0.0, 0.0)], 2),
                                            "gbps": round(file_bytes / corr_us[(0.0, 0.0)] / 1e3, 1)},
                            "gradients_0.1_0.05": {"avg_launch_us": round(corr_us[(0.1, 0.05)], 2),
                                                   "gbps": round(file_bytes / corr_us[(0.1, 0.05)] / 1e3, 1)}},
        "crava_file": {"what": "FFTGrid::writeCravaFile then readCravaFile of one grid (%.0f MB), local disk"
                               % (file_bytes / 1e6),
                       "write_ms": round(1000 * t_write, 1), "read_ms": round(1000 * t_read, 1),
                       "write_gbps": file_bytes / t_write / 1e9, "read_gbps": file_bytes / t_read / 1e9},
    }


def bench_large(ctx, steps=2):
    """The large single cube with simulation: 1024 x 1024 x 320 padded to 1152 x 1152 x 384, 6
    angles 0-50 deg, Ormsby 5-10-60-80 Hz wavelet, 1500 m lateral range, 20 posterior realisations
    (all resident, coloured in pairs inside cg_invert), covariance kept as the 6 Hermitian
    spectra."""
    sp = S.large_cube_spec(realisations=20)
    t0 = time.perf_counter()
    prob = S.Problem(ctx, sp)
    ctx.autotune(1152, 1152, 384)
    log("large cube setup %.1f s (%.1f GB)" % (time.perf_counter() - t0, S.footprint(sp) / 1e9))
    prob.step()
    ctx.sync()
    t = time.perf_counter()
    for _ in range(steps):
        prob.step()
    ctx.sync()
    ms = 1000.0 * (time.perf_counter() - t) / steps
    t = time.perf_counter()
    prob.step(realisations=0)
    ctx.sync()
    ms0 = 1000.0 * (time.perf_counter() - t)
    # which colouring the realisation pairs took (the autotuned choice at this length): the fused
    # kernel runs no z launches of its own
    ctx.set_profiling(True)
    ctx.reset_kernel_time()
    prob.realisations(2)
    ctx.sync()
    nz_launch = ctx.kernel_time(C.K_FFT_Z)[0]
    n_col, ms_col = ctx.kernel_time(C.K_NOISE)
    ctx.set_profiling(False)
    G = (sp["nxp"] // 2 + 1) * sp["nyp"] * sp["nzp"] * 8
    nt = len(sp["angles"])
    cells = S.cells(sp)
    prob.close()
    return {
        "workload": "single cube 1024x1024x320 padded 1152x1152x384, 6 angles 0-50 deg, Ormsby 5-10-60-80 Hz, "
                    "1500 m lateral range, posterior mean + covariance + 20 realisations",
        "ms_per_step": ms, "cells_per_s": cells / (ms * 1e-3), "mean_cov_only_ms": ms0,
        "model_v1_frac": G * (5 * (3 + nt) + 21 + 24 * 20) / (ms * 1e-3) / HBM_PEAK,
        "mean_cov_model_v1_frac": G * (5 * (3 + nt) + 21) / (ms0 * 1e-3) / HBM_PEAK,
        "pair_colouring": {"fused": nz_launch == 0, "colour_ms": round(ms_col / max(n_col, 1), 3)},
        "steps": steps,
    }


def bench_serving(n=48, lanes=2, load=0.7, ingest_workers=2):
    """Continuous batching under Poisson arrivals (crava_amd.serve): requests of log-uniform sizes
    64^2 x 64 .. 384^2 x 256, 3-8 angles, 0-10 realisations, admitted against a 200 GB HBM
    budget.  A first pass with every request arriving at once measures the server's capacity
    (and builds every FFT table); the reported pass offers Poisson arrivals at `load` times that
    capacity.  Latency = completion - arrival (queueing, admission, input generation on the
    device, batched compute)."""
    from crava_amd import serve
    specs = S.serving_requests(n)
    ex = serve.GpuExecutor(lanes=lanes, ingest_workers=ingest_workers)
    try:
        srv = serve.Server(ex, hbm_budget=200e9, lanes=lanes, max_batch=16)
        cap = serve.run_open_loop(srv, specs, np.zeros(n))
        rate = load * cap["requests"] / cap["seconds"]
        srv = serve.Server(ex, hbm_budget=200e9, lanes=lanes, max_batch=16)
        rep = serve.run_open_loop(srv, specs, serve.poisson_arrivals(n, rate, seed=7))
        ingest = np.mean([r.ready - r.admitted for r in srv.done]) * 1e3
    finally:
        ex.close()
    return {
        "workload": "%d requests, sizes log-uniform 64x64x64..384x384x256 (padded by the reference rule), 3-8 "
                    "angles, 0-10 realisations, Poisson arrivals at %.1f req/s (%.0f%% of capacity), 200 GB HBM "
                    "budget, %d lanes, %d ingest threads" % (n, rate, 100 * load, lanes, ingest_workers),
        "cells_per_s": rep["cells_per_s"], "latency_ms": rep["latency_ms"], "requests": rep["requests"],
        "batches": rep["batches"], "mean_batch": rep["mean_batch"], "peak_admitted_gb": rep["peak_admitted_gb"],
        "mean_ingest_ms": float(ingest),
        "capacity": {"cells_per_s": cap["cells_per_s"], "requests_per_s": cap["requests"] / cap["seconds"],
                     "latency_ms": cap["latency_ms"], "batches": cap["batches"], "mean_batch": cap["mean_batch"]},
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--realisations", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-mix", action="store_true", help="skip the batched-vs-plain uneven mix")
    ap.add_argument("--no-serve", action="store_true", help="skip the continuous-batching serving run")
    ap.add_argument("--no-large", action="store_true", help="skip the 1152x1152x384 cube with 20 realisations")
    ap.add_argument("--no-pad", action="store_true", help="skip the input padding fills")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a copy-like kernel of known traffic (for scaling hardware byte counters)")
    ap.add_argument("--profile-steps", type=int, default=2)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: there is no multi-GPU path (independent requests would "
              "run one per GPU, a single cube as z slabs with an all-to-all transpose); run with "
              "--gpus 1", file=sys.stderr)
        sys.exit(2)
    R = args.realisations
    w = S.single_cube_spec(realisations=R)
    ctx = C.Context(0)
    if args.copy_calibration:
        # a copy-like kernel of known traffic for scaling hardware byte counters: FFTGrid::add
        # on two grids of the workload's size reads 2 and writes 1 grid (k_add), 5 launches
        a = C.FFTGrid(ctx, w["nx"], w["ny"], w["nz"], w["nxp"], w["nyp"], w["nzp"], C.PARAMETER)
        b = C.FFTGrid(ctx, w["nx"], w["ny"], w["nz"], w["nxp"], w["nyp"], w["nzp"], C.PARAMETER)
        for _ in range(5):
            a.add(b)
        ctx.sync()
        log("copy calibration: 5 x k_add over %d floats" % (2 * (w["nxp"] // 2 + 1) * w["nyp"] * w["nzp"]))
        ctx.close()
        return
    t0 = time.perf_counter()
    inv = S.Problem(ctx, w)
    ctx.autotune(w["nxp"], w["nyp"], w["nzp"])
    log("setup %.1f s (incl. autotuning the FFT tiles of the grid)" % (time.perf_counter() - t0))
    for _ in range(args.warmup):
        inv.step()
    ctx.sync()
    stats = inv.step(stats=True)
    ctx.sync()
    log("posterior stats: inverted %d, cholesky failures %d, fp64 re-solves %d" % stats)
    # throughput: exactly K steps between two synchronisations
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        inv.step()
    ctx.sync()
    dt = time.perf_counter() - t0
    ms_per_step = 1000.0 * dt / args.steps
    cells = w["nx"] * w["ny"] * w["nz"]
    value = cells * args.steps / dt
    log("timed %d steps: %.3f ms/step" % (args.steps, ms_per_step))
    # latency of individual requests (each step synchronised)
    lat = []
    for _ in range(args.steps):
        t = time.perf_counter()
        inv.step()
        ctx.sync()
        lat.append(1000.0 * (time.perf_counter() - t))
    # mean + covariance only
    t = time.perf_counter()
    for _ in range(args.steps):
        inv.step(realisations=0)
    ctx.sync()
    ms_meancov = 1000.0 * (time.perf_counter() - t) / args.steps
    # per-kernel HIP-event times over profiled steps
    ctx.set_profiling(True)
    ctx.reset_kernel_time()
    for _ in range(args.profile_steps):
        inv.step()
    kt = {k: ctx.kernel_time(k) for k in NAMES}
    # streaming ceiling of this GPU in the same run: FFTGrid::add (k_add) on two grids of the
    # workload's size reads 2 and writes 1 grid, nothing else
    ctx.reset_kernel_time()
    for _ in range(5):
        inv.mean_in[0].add(inv.mean_in[1])
    n_add, ms_add = ctx.kernel_time(C.K_ELEMENTWISE)
    ctx.set_profiling(False)
    for g, v in zip(inv.mean_in, S.BACKGROUND):
        g.fillInConstant(v)  # restore the prior means
    stream_gbps = 3 * inv.mean_in[0].rnxp * w["nyp"] * w["nzp"] * 4 / (ms_add * 1e-3 / n_add) / 1e9
    # the reference's own noise stream (RandomGen + fillInComplexNoise), bit for bit, on one grid
    seed = 23665
    seed = inv.out[0][0].fillInComplexNoiseReference(seed)  # warm
    t = time.perf_counter()
    for g in inv.out[0]:
        seed = g.fillInComplexNoiseReference(seed)
    ms_refnoise = 1000.0 * (time.perf_counter() - t) / 3
    split = kt[C.K_FFT_Z][0] == 0  # the split-z request launches no z pass of its own
    kb = kernel_bytes(w, inv.nt, R, inv.n_inverted_planes,
                      {k: kt[k][0] // args.profile_steps for k in (C.K_FFT_X, C.K_FFT_Y, C.K_FFT_Z)}, split)
    kernels = {}
    for k, (n, ms) in kt.items():
        if n == 0:
            continue
        per_step_ms = ms / args.profile_steps
        e = {"launches_per_step": n // args.profile_steps, "ms_per_step": round(per_step_ms, 4),
             "avg_launch_us": round(1000.0 * ms / n, 2)}
        if k in kb:
            e["gbps"] = round(kb[k] / (per_step_ms * 1e-3) / 1e9, 1)
        kernels[NAMES[k]] = e
    dom = max((k for k in kb if kt[k][0] > 0), key=lambda k: kt[k][1])
    n_l, ms_l = kt[dom]
    per_launch_bytes = kb[dom] / (n_l / args.profile_steps)
    avg_s = ms_l * 1e-3 / n_l
    achieved = per_launch_bytes / avg_s
    G = (w["nxp"] // 2 + 1) * w["nyp"] * w["nzp"] * 8
    model_bytes = G * (5 * (3 + inv.nt) + 21 + 24 * R)
    model_bytes_r0 = G * (5 * (3 + inv.nt) + 21)
    model_min = G * ((3 + inv.nt) + 9 + 3 * R)
    result = {
        "metric": "posterior cells/s, 512²×256×5 angles, batched vs plain; p50/p99 latency",
        "value": value,
        "unit": "cells/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "single cube 512x512x256 padded 576x576x320, 5 angles, posterior mean + "
                               "covariance + %d realisations, FP32 with FP64 re-solve where cond > 1e6" % R,
                   "grid": [w["nx"], w["ny"], w["nz"]], "padded": [w["nxp"], w["nyp"], w["nzp"]],
                   "angles": list(w["angles"]), "realisations": R, "inputs": "resident in HBM"},
        "roofline": {"bound": "hbm", "kernel": NAMES[dom], "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK, "traffic": measured_traffic(NAMES[dom]),
                     "bytes_per_launch": per_launch_bytes, "avg_launch_us": avg_s * 1e6},
        "latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
                       "samples": len(lat), "kind": "plain single request, synchronised"},
        "mean_cov_only_ms": ms_meancov,
        "mean_cov_model_v1": {"bytes": model_bytes_r0,
                              "frac_of_roofline": model_bytes_r0 / (ms_meancov * 1e-3) / HBM_PEAK},
        "model_v1": {"bytes": model_bytes, "frac_of_roofline": model_bytes / (ms_per_step * 1e-3) / HBM_PEAK,
                     "bytes_min": model_min},
        "kernels": kernels,
        "reference_noise_fill": {"ms_per_grid": ms_refnoise, "normals_per_grid": 2 * (w["nxp"] // 2 + 1) * w["nyp"] * w["nzp"],
                                 "what": "RandomGen + fillInComplexNoise stream, bit-identical, three consecutive "
                                         "grids (one realisation's seeds), synchronised"},
        "stream_ceiling": {"kernel": "k_add (read 2 grids, write 1)", "gbps": round(stream_gbps, 1),
                           "frac_of_peak": stream_gbps * 1e9 / HBM_PEAK},
        "posterior_stats": {"inverted": stats[0], "cholesky_failures": stats[1], "fp64_resolves": stats[2]},
        "accuracy": accuracy_summary(),
        "cpu_baseline": None,
    }
    if not args.no_pad:
        result["input_padding"] = bench_input_padding(ctx, inv, w)
        log("input padding: %s" % result["input_padding"])
    inv.close()
    if not args.no_mix:
        log("uneven mix: plain vs batched ...")
        result["batched"] = bench_mix(ctx)
    if not args.no_serve:
        log("serving ...")
        result["serving"] = bench_serving()
    if not args.no_large:
        log("large cube ...")
        result["large_cube"] = bench_large(ctx)
    if not args.no_cpu:
        log("cpu baseline ...")
        try:
            result["cpu_baseline"] = cpu_baseline(w, inv, R)
            ncores = min(16, int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or os.cpu_count() or 1)
            if ncores > 1:
                result["cpu_baseline_all_cores"] = cpu_baseline(w, inv, R, threads=ncores)
        except Exception as e:  # reported, not fatal: the GPU figure stands on its own
            result["cpu_baseline"] = {"error": repr(e)}
    ctx.close()
    print(json.dumps(result))


if __name__ == "__main__":
    main()
