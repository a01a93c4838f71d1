"""Batched inversions (cg_batch): every cube of a mix shares the launches of each pass.
Results must equal the single-request path and not depend on the other cubes in the batch."""
import numpy as np
import pytest

import crava_amd as C
import oracle as O
import problem as P
from conftest import rel_l2

pytestmark = pytest.mark.gpu

MIX = [dict(nx=16, ny=12, nz=20, seed=31),
       dict(nx=20, ny=10, nz=30, seed=32, angles_deg=(0.0, 10.0, 20.0, 30.0, 40.0), peak_hz=30.0),
       dict(nx=12, ny=12, nz=24, seed=33, angles_deg=(0.0, 12.0, 24.0, 36.0)),
       dict(nx=30, ny=20, nz=16, seed=34, angles_deg=(5.0, 25.0, 45.0, 10.0, 20.0, 30.0))]


def _item(ctx, p, nreal, seed):
    mk = lambda ct: C.FFTGrid(ctx, p["nx"], p["ny"], p["nz"], p["nxp"], p["nyp"], p["nzp"], ct)
    it = dict(tables=p, rxy=p["rxy"], rt=p["rt"], lam2d=p["lam2d"], seed=seed,
              mean_in=[mk(C.PARAMETER) for _ in range(3)], seis_in=[mk(C.DATA) for _ in range(p["ntheta"])],
              mean=[mk(C.PARAMETER) for _ in range(3)], data=[mk(C.DATA) for _ in range(p["ntheta"])],
              cov=[mk(C.COVARIANCE) for _ in range(6)], out=[[mk(C.PARAMETER) for _ in range(3)] for _ in range(nreal)])
    for g, r in zip(it["mean_in"], p["mean_real"]):
        g.upload_real(r)
    for g, r in zip(it["seis_in"], p["seis_real"]):
        g.upload_real(r)
    return it


def _results(it):
    return (np.stack([g.download_real() for g in it["mean"]]), np.stack([g.download_complex() for g in it["cov"]]),
            [np.stack([g.download_real() for g in r]) for r in it["out"]])


def test_batch_matches_oracle_and_is_independent_of_the_mix(ctx):
    probs = [P.make_problem(**kw) for kw in MIX]
    items = [_item(ctx, p, 1, 100 + i) for i, p in enumerate(probs)]
    b = C.Batch(ctx, items)
    b.run()
    b.wait()
    res = [_results(it) for it in items]
    for i, (p, (mean, cov, outs)) in enumerate(zip(probs, res)):
        ms, ds = P.forward_spectra(p, np.float64)
        om, _, oc, _, _ = O.posterior(p, ms, ds, None, "f64")
        want = np.stack([O.grid_ifft(m, p["nxp"], O.PARAMETER) for m in om])
        assert rel_l2(mean - p["mean_real"], want - p["mean_real"]) < 1e-4
        assert rel_l2(cov, oc) < 1e-5
        # realisation = posterior mean + coloured counter noise
        noise = np.stack([O.fill_complex_noise_cb(p["nxp"], p["nyp"], p["nzp"], 100 + i, g)
                          for g in range(3)])
        z, _ = O.colour(oc, noise, "f64")
        real = np.stack([O.grid_ifft(z[g], p["nxp"], O.PARAMETER) for g in range(3)])
        assert rel_l2(outs[0] - mean, real) < 1e-4
    # the same inversion alone in a batch of one: bit-identical
    alone = _item(ctx, probs[1], 1, 101)
    b1 = C.Batch(ctx, [alone])
    b1.run()
    b1.wait()
    a_mean, a_cov, a_outs = _results(alone)
    assert np.array_equal(a_mean, res[1][0]) and np.array_equal(a_cov, res[1][1])
    assert np.array_equal(a_outs[0], res[1][2][0])
    b.close()
    b1.close()


def test_batch_graph_replays_match_eager_runs(ctx):
    probs = [P.make_problem(**kw) for kw in MIX[:3]]
    items = [_item(ctx, p, 1, 7 + i) for i, p in enumerate(probs)]
    b = C.Batch(ctx, items)
    b.run()
    b.wait()
    eager = [_results(it) for it in items]
    b.use_graph(True)
    for rep in range(3):
        # eager work on the same stream between replays
        for it in items:
            for g in it["mean"]:
                g.multiplyByScalar(0.0)
        b.run()
        b.wait()
        for it, ref in zip(items, eager):
            got = _results(it)
            assert np.array_equal(got[0], ref[0]) and np.array_equal(got[1], ref[1]), rep
            assert np.array_equal(got[2][0], ref[2][0]), rep
    b.close()


def test_batch_matches_single_request_api(ctx):
    p = P.make_problem(**MIX[1])
    it = _item(ctx, p, 0, 5)
    b = C.Batch(ctx, [it], precision=C.PREC_MIXED)
    b.run()
    b.wait()
    mean_b = np.stack([g.download_real() for g in it["mean"]])
    # single-request path on the same inputs
    mean = [C.FFTGrid(ctx, p["nx"], p["ny"], p["nz"], p["nxp"], p["nyp"], p["nzp"], C.PARAMETER) for _ in range(3)]
    data = [C.FFTGrid(ctx, p["nx"], p["ny"], p["nz"], p["nxp"], p["nyp"], p["nzp"], C.DATA)
            for _ in range(p["ntheta"])]
    cov = [C.FFTGrid(ctx, p["nx"], p["ny"], p["nz"], p["nxp"], p["nyp"], p["nzp"], C.COVARIANCE) for _ in range(6)]
    for g, s in zip(mean, it["mean_in"]):
        g.fft_from(s)
    for g, s in zip(data, it["seis_in"]):
        g.fft_from(s)
    C.posterior(ctx, p, mean, data, cov, rxy=p["rxy"], rt=p["rt"], lam2d=p["lam2d"], want_residual=False,
                precision=C.PREC_MIXED)
    for g in mean:
        g.invFFTInPlace()
    mean_s = np.stack([g.download_real() for g in mean])
    assert rel_l2(mean_b - p["mean_real"], mean_s - p["mean_real"]) < 1e-5
    b.close()


def test_batch_compiled_plans_match_runtime_plans(ctx):
    """The batch groups grids by compiled FFT length; the runtime-plan kernels (every grid in one
    launch per pass) must give the same inversion."""
    probs = [P.make_problem(**kw) for kw in MIX]
    res = {}
    for eng in (C.FFT_AUTO, C.FFT_GENERIC):
        ctx.set_fft_engine(eng)
        items = [_item(ctx, p, 1, 40 + i) for i, p in enumerate(probs)]
        b = C.Batch(ctx, items)
        b.run()
        b.wait()
        res[eng] = [_results(it) for it in items]
        b.close()
    ctx.set_fft_engine(C.FFT_AUTO)
    # two FP32 paths (different radix orders): each is within 1e-4 of the FP64 oracle (the test
    # above); their difference, ~1e-5 here, is held to 5e-5
    for (ma, ca, oa), (mg, cg_, og), p in zip(res[C.FFT_AUTO], res[C.FFT_GENERIC], probs):
        assert rel_l2(ma - p["mean_real"], mg - p["mean_real"]) < 5e-5
        assert rel_l2(ca, cg_) < 5e-5
        assert rel_l2(oa[0] - ma, og[0] - mg) < 5e-5


def test_batch_fused_colouring_matches_unfused_and_oracle(ctx):
    """Realisations of a batch with the colouring fused into the inverse z pass (one launch per z
    length over every inversion of that length) and with the colouring kernel then z passes: both
    match the oracle, and the fused schedule really skips the z passes it absorbed.  The mix holds
    two z lengths of one plan each plus the single-cube depth (256 of 320 planes); with three
    realisations the fused schedule colours the first two in one paired launch per z length."""
    probs = [P.make_problem(**MIX[0]), P.make_problem(**MIX[1]), P.make_problem(11, 7, 256, nzp=320, seed=5),
             P.make_problem(9, 6, 256, nzp=320, seed=6)]
    nreal = 3  # fused: realisations 0 and 1 by the paired kernel, 2 alone
    items = [_item(ctx, p, nreal, 60 + i) for i, p in enumerate(probs)]
    res, zlaunches = {}, {}
    for fused in (0, 1):
        print("batch fused colour", fused, flush=True)
        ctx.set_fused_colour(fused)
        b = C.Batch(ctx, items)
        ctx.set_profiling(True)
        ctx.reset_kernel_time()
        b.run()
        b.wait()
        zlaunches[fused] = ctx.kernel_time(C.K_FFT_Z)[0]
        ctx.set_profiling(False)
        res[fused] = [_results(it) for it in items]
        b.close()
    ctx.set_fused_colour(-1)
    assert zlaunches[1] < zlaunches[0]
    for i, p in enumerate(probs):
        mean, cov, outs = res[1][i]
        _, _, outs0 = res[0][i]
        ms, ds = P.forward_spectra(p, np.float64)
        _, _, oc, _, _ = O.posterior(p, ms, ds, None, "f64")
        for r in range(nreal):
            noise = np.stack([O.fill_complex_noise_cb(p["nxp"], p["nyp"], p["nzp"], 60 + i, 3 * r + g)
                              for g in range(3)])
            z, _ = O.colour(oc, noise, "f64")
            real = np.stack([O.grid_ifft(z[g], p["nxp"], O.PARAMETER) for g in range(3)])
            assert rel_l2(outs[r] - mean, real) < 1e-4, (i, r)
            assert rel_l2(outs[r] - mean, outs0[r] - mean) < 1e-5, (i, r)
