#!/usr/bin/env python3
"""Headline benchmark: posterior cells/s of the FFT-domain Bayesian AVO inversion on one MI355X.

Workload (one step = one inversion, inputs resident in HBM when the timed region starts):
  single cube 512 x 512 x 256 (padded 576 x 576 x 320), dx = dy = 25 m, dt = 4 ms, 5 angles
  0-40 deg, 30 Hz Ricker with per-angle scaling U(0.8, 1.2), S/N per angle U(3, 8); prior
  ln(3000, 1500, 2.3), std (0.10, 0.15, 0.05), correlations (0.7, 0.4, 0.3), spherical 1000 m
  lateral and exponential 40 ms temporal correlation; a seeded Gaussian elastic field drawn from
  the prior and forward-modelled to seismic on the GPU.
  Step: 3 + 5 forward FFTs, per-wavenumber posterior (FP32, FP64 re-solve where the pivot ratio of
  the Cholesky factor exceeds 1e6), 3 inverse FFTs of the posterior means, 2 posterior
  realisations (counter-based noise, colouring, 3 inverse FFTs + mean each).

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
from crava_amd import model as M  # noqa: E402

HBM_PEAK = 8.0e12  # MI355X HBM3E, bytes/s

WORKLOAD = dict(nx=512, ny=512, nz=256, nxp=576, nyp=576, nzp=320, dx=25.0, dy=25.0, dt=4.0,
                angles=(0.0, 10.0, 20.0, 30.0, 40.0), peak_hz=30.0, std=(0.10, 0.15, 0.05),
                corr=(0.7, 0.4, 0.3), lateral_range=1000.0, temporal_range=40.0, wnc=0.1, vsvp=0.5,
                low_cut=5.0, high_cut=55.0, seed=20261202)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


class Inversion:
    """Device-resident inputs, work grids and the per-step call sequence."""

    def __init__(self, ctx, w, realisations):
        self.ctx, self.w, self.R = ctx, w, realisations
        rng = np.random.default_rng(w["seed"])
        nx, ny, nz, nxp, nyp, nzp = (w[k] for k in ("nx", "ny", "nz", "nxp", "nyp", "nzp"))
        self.nt = nt = len(w["angles"])
        self.cnxp = nxp // 2 + 1
        scales = rng.uniform(0.8, 1.2, nt)
        sn = rng.uniform(3.0, 8.0, nt)
        tab = M.inversion_tables(nz, nzp, w["dt"], w["angles"], w["peak_hz"], scales, w["vsvp"], w["low_cut"],
                                 w["high_cut"])
        cxy = M.spherical_corr_xy(nxp, nyp, w["dx"], w["dy"], w["lateral_range"])
        self.rxy = M.lateral_spectrum(cxy)
        self.rt = M.temporal_spectrum(M.circ_corr_t(w["dt"], w["temporal_range"], nzp, w["low_cut"]))
        S0 = M.sigma0(w["std"], w["corr"])
        mk = lambda ct: C.FFTGrid(ctx, nx, ny, nz, nxp, nyp, nzp, ct)
        self.mean_in = [mk(C.PARAMETER) for _ in range(3)]
        self.seis_in = [mk(C.DATA) for _ in range(nt)]
        self.mean_w = [mk(C.PARAMETER) for _ in range(3)]
        self.data_w = [mk(C.DATA) for _ in range(nt)]
        self.cov = [mk(C.COVARIANCE) for _ in range(6)]
        self.out = [[mk(C.PARAMETER) for _ in range(3)] for _ in range(max(realisations, 1))]
        rnxp = 2 * self.cnxp
        for g, v in zip(self.mean_in, (math.log(3000.0), math.log(1500.0), math.log(2.3))):
            g.upload_real(np.full((nzp, nyp, rnxp), v, dtype=np.float32))
        # prior covariance spectra: a posterior with no inverted plane writes the prior
        for g, s in zip(self.mean_w, self.mean_in):
            g.fft_from(s)
        for g, s in zip(self.data_w, self.seis_in):
            g.fft_from(s)
        base = dict(K=tab["K"], e1=tab["e1"], e2=tab["e2"], e3=tab["e3"], sigma0e=np.eye(nt), wnc=w["wnc"],
                    prior_var0=S0)
        C.posterior(ctx, dict(base, invert=np.zeros(nzp, np.int32)), self.mean_w, self.data_w, self.cov,
                    rxy=self.rxy, rt=self.rt, lam2d=self.rxy, want_residual=False)
        # elastic field from the prior, forward-modelled seismic, white noise at S/N
        truth = self.out[0]
        C.simulate(ctx, self.cov, self.mean_in, truth, seed=w["seed"], realisation=1000)
        for g, s in zip(self.mean_w, truth):
            g.fft_from(s)
        C.forward_model(ctx, tab["K"], self.mean_w, self.data_w)
        sig_e = np.zeros(nt)
        for l in range(nt):
            self.data_w[l].setType(C.DATA)
            self.data_w[l].invFFTInPlace()
            var_s = self.data_w[l].sum_squares() / (nx * ny * nz)
            sig_e[l] = math.sqrt(var_s / (sn[l] - 1.0))  # S/N = (|s|^2 + |e|^2) / |e|^2
            self.seis_in[l].add(self.data_w[l])
            self.seis_in[l].add_white_noise(float(sig_e[l]), w["seed"], 500 + l)
        self.tables = dict(base, invert=tab["invert"], sigma0e=M.angular_error_covariance(sig_e, w["angles"]))
        self.n_inverted_planes = int(np.sum(tab["invert"] != 0))
        ctx.sync()

    def step(self, realisations=None, stats=False):
        R = self.R if realisations is None else realisations
        for g, s in zip(self.mean_w, self.mean_in):
            g.fft_from(s)
        for g, s in zip(self.data_w, self.seis_in):
            g.fft_from(s)
        st = C.posterior(self.ctx, self.tables, self.mean_w, self.data_w, self.cov, rxy=self.rxy, rt=self.rt,
                         lam2d=self.rxy, want_residual=False, precision=C.PREC_MIXED, cond_limit=1e6, stats=stats)
        for g in self.mean_w:
            g.invFFTInPlace()
        for r in range(R):
            C.simulate(self.ctx, self.cov, self.mean_w, self.out[r], seed=self.w["seed"], realisation=r)
        return st


def kernel_bytes(w, nt, R, n_inv_planes):
    """Algorithmic HBM bytes per kernel class for one step (what each kernel must read + write)."""
    G = (w["nxp"] // 2 + 1) * w["nyp"] * w["nzp"] * 8  # one complex grid
    Nr = w["nxp"] * w["nyp"] * w["nzp"] * 4  # one real cube (without row padding)
    plane = (w["nxp"] // 2 + 1) * w["nyp"]
    nfwd, ninv = 3 + nt, 3 + 3 * R
    return {
        C.K_FFT_X: nfwd * (Nr + G) + 3 * (G + Nr) + 3 * R * (G + 2 * Nr),
        C.K_FFT_Y: (nfwd + ninv) * 2 * G,
        C.K_FFT_Z: (nfwd + ninv) * 2 * G,
        # inverted cells: read 3 + nt spectra, write 3 means + 6 covariance entries;
        # pass-through cells: write the 6 prior covariance entries
        C.K_POSTERIOR: n_inv_planes * plane * (3 + nt + 9) * 8 + (w["nzp"] - n_inv_planes) * plane * 6 * 8,
        C.K_NOISE: R * 9 * G,  # colouring: read 6 covariance, write 3 (noise generated in registers)
    }


NAMES = {C.K_FFT_X: "fft_x", C.K_FFT_Y: "fft_y", C.K_FFT_Z: "fft_z", C.K_POSTERIOR: "posterior",
         C.K_NOISE: "colour", C.K_ELEMENTWISE: "elementwise"}


def cpu_baseline(w, inv, R, seconds_hint=20.0):
    """The reference's own code (oracle/_ref: FFTW 2.1.2, lib_matr, RandomGen) on one host core, on a
    bounded sample, extrapolated to the reference path for this workload: 15 forward FFTs (3 means,
    5 seismic, 6 prior covariance, errCorr), 9 inverse (3 means, 6 covariance), the posterior loop,
    and per realisation 3 noise fills, the colouring loop and 3 inverse FFTs."""
    import oracle as O
    if not O.ref_available():
        return None
    R_ = O.ref()
    nxp, nyp, nzp = w["nxp"], w["nyp"], w["nzp"]
    cnxp = nxp // 2 + 1
    plane = cnxp * nyp
    rng = np.random.default_rng(1)
    x = rng.standard_normal((nzp, nyp, 2 * cnxp)).astype(np.float32)
    t = time.perf_counter()
    R_.ref_grid_fft(nxp, nyp, nzp, O.DATA, O._ptr(x, O.ctypes.c_float))
    t_fwd = time.perf_counter() - t
    t = time.perf_counter()
    R_.ref_grid_ifft(nxp, nyp, nzp, O.PARAMETER, O._ptr(x, O.ctypes.c_float))
    t_inv = time.perf_counter() - t
    del x
    # posterior loop on S planes of inverted and S planes of pass-through wavenumbers, with the
    # workload's tables, in the reference's covariance-grid form
    S = 4
    tb = inv.tables
    inv_k = np.where(tb["invert"] != 0)[0]
    pass_k = np.where(tb["invert"] == 0)[0]
    times = {}
    cov_post = None
    for label, ks in (("inv", inv_k[np.linspace(0, len(inv_k) - 1, S).astype(int)]),
                      ("pass", pass_k[np.linspace(0, len(pass_k) - 1, S).astype(int)])):
        p = dict(nxp=nxp, nyp=nyp, nzp=S, ntheta=inv.nt, K=tb["K"][ks], e1=tb["e1"][ks], e2=tb["e2"][ks],
                 e3=tb["e3"][ks], invert=tb["invert"][ks], sigma0e=tb["sigma0e"], wnc=tb["wnc"],
                 prior_var0=tb["prior_var0"], prior_mode=0, rxy=None, rt=None, lam2d=None)
        shape = (S, nyp, cnxp)
        r = np.abs(inv.rxy[None] * inv.rt[ks][:, None, None])
        S0 = tb["prior_var0"]
        covs = np.stack([(r * S0[a, b]).astype(np.complex64) for a, b in
                         ((0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2))])
        p["lam_grid"] = np.broadcast_to(inv.rxy[None], shape).astype(np.complex64)
        mean = (rng.standard_normal((3,) + shape) + 1j * rng.standard_normal((3,) + shape)).astype(np.complex64)
        data = (rng.standard_normal((inv.nt,) + shape) + 1j * rng.standard_normal((inv.nt,) + shape)).astype(np.complex64)
        t = time.perf_counter()
        _, _, c, _, _ = O.ref_posterior(p, mean, data, covs)
        times[label] = (time.perf_counter() - t) / (S * plane)
        if label == "inv":
            cov_post = c
    seed = (rng.standard_normal((3,) + cov_post.shape[1:]) + 0j).astype(np.complex64)
    t = time.perf_counter()
    O.ref_colour(cov_post, seed)
    t_col = (time.perf_counter() - t) / (S * plane)
    R_.ref_rng_seed(23665)
    noise = np.zeros((S, nyp, cnxp), dtype=np.complex64)
    t = time.perf_counter()
    R_.ref_fill_complex_noise(nxp, nyp, S, O._ptr(noise.view(np.float32), O.ctypes.c_float))
    t_noise = (time.perf_counter() - t) / (S * plane)
    n_inv = inv.n_inverted_planes * plane
    n_all = nzp * plane
    total = (15 * t_fwd + 9 * t_inv + n_inv * times["inv"] + (n_all - n_inv) * times["pass"]
             + R * (3 * n_all * t_noise + n_all * t_col + 3 * t_inv))
    cells = w["nx"] * w["ny"] * w["nz"]
    return {
        "value": cells / total, "unit": "cells/s", "cores": 1, "kind": "reference",
        "seconds_per_inversion": total,
        "sample": ("reference FFTW 2.1.2 rfftw3d forward (%.2f s) and inverse (%.2f s) of one 576x576x320 grid; "
                   "reference lib_matr posterior loop on %d inverted + %d pass-through k planes (%.2f / %.3f us "
                   "per wavenumber); lib_matr colouring and RandomGen noise fill on %d planes; extrapolated to "
                   "15 forward + 9 inverse FFTs, the full loop and %d realisations" %
                   (t_fwd, t_inv, S, S, 1e6 * times["inv"], 1e6 * times["pass"], S, R)),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--realisations", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--profile-steps", type=int, default=2)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (the slab-decomposed "
              "all-to-all design is described in DESIGN.md); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    w = dict(WORKLOAD)
    R = args.realisations
    ctx = C.Context(0)
    t0 = time.perf_counter()
    inv = Inversion(ctx, w, R)
    log("setup %.1f s" % (time.perf_counter() - t0))
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
    ctx.set_profiling(False)
    kb = kernel_bytes(w, inv.nt, R, inv.n_inverted_planes)
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
                     "unit": "GB/s", "frac": achieved / HBM_PEAK, "traffic": None,
                     "bytes_per_launch": per_launch_bytes, "avg_launch_us": avg_s * 1e6},
        "latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
                       "samples": len(lat), "kind": "plain single request, synchronised"},
        "mean_cov_only_ms": ms_meancov,
        "model_v1": {"bytes": model_bytes, "frac_of_roofline": model_bytes / (ms_per_step * 1e-3) / HBM_PEAK,
                     "bytes_min": model_min},
        "kernels": kernels,
        "posterior_stats": {"inverted": stats[0], "cholesky_failures": stats[1], "fp64_resolves": stats[2]},
        "cpu_baseline": None,
    }
    if not args.no_cpu:
        log("cpu baseline ...")
        try:
            result["cpu_baseline"] = cpu_baseline(w, inv, R)
        except Exception as e:  # reported, not fatal: the GPU figure stands on its own
            result["cpu_baseline"] = {"error": repr(e)}
    ctx.close()
    print(json.dumps(result))


if __name__ == "__main__":
    main()
