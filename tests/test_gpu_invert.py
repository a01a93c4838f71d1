"""One whole inversion in one call (cg_invert).  Grids with a compiled split of the z transform
(x length 576 with nzp 320 = 16 x 20, x length 1152 with nzp 384 = 16 x 24) run the split-z
passes: the x passes carry the z transform's first stage, the posterior pass its second stage
and the inverse's first.  The results must match the FP64 oracle as the unsplit path does, and
the unsplit path on the same inputs to FP32 rounding."""
import numpy as np
import pytest

import crava_amd as C
import oracle as O
import problem as P
from conftest import rel_l2

pytestmark = pytest.mark.gpu

# thin slabs of the production geometries: x and z at their full padded lengths, few y rows
CASES = [dict(nx=512, ny=3, nz=256, nxp=576, nyp=4, nzp=320, seed=41, angles_deg=(0.0, 10.0, 20.0, 30.0, 40.0)),
         dict(nx=500, ny=1, nz=250, nxp=576, nyp=1, nzp=320, seed=42),
         dict(nx=1024, ny=2, nz=320, nxp=1152, nyp=3, nzp=384, seed=43,
              angles_deg=(0.0, 10.0, 20.0, 30.0, 40.0, 50.0), peak_hz=30.0)]


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


def _close(it):
    for k in ("mean_in", "seis_in", "mean", "data", "cov"):
        for g in it[k]:
            g.close()
    for r in it["out"]:
        for g in r:
            g.close()


def _run(ctx, p, nreal, seed, engine=C.FFT_AUTO, precision=C.PREC_MIXED):
    it = _item(ctx, p, nreal, seed)
    ctx.set_fft_engine(engine)
    ctx.set_profiling(True)
    ctx.reset_kernel_time()
    try:
        stats = C.invert(ctx, it, precision=precision, stats=True)
        ctx.sync()
        z_launches = ctx.kernel_time(C.K_FFT_Z)[0]
    finally:
        ctx.set_profiling(False)
        ctx.set_fft_engine(C.FFT_AUTO)
    res = _results(it)
    _close(it)
    return res, stats, z_launches


@pytest.mark.parametrize("case", range(len(CASES)))
def test_split_z_inversion_matches_fp64_oracle(ctx, case):
    kw = CASES[case]
    print("split-z case", kw, flush=True)
    p = P.make_problem(**kw)
    (mean, cov, outs), (ninv, nfail, n64), zl = _run(ctx, p, 2, 500 + case)
    assert zl == 0  # the split passes ran: no z pass of its own
    assert ninv == int(p["invert"].sum()) * p["nyp"] * p["cnxp"] and nfail == 0
    ms, ds = P.forward_spectra(p, np.float64)
    om, _, oc, _, _ = O.posterior(p, ms, ds, None, "f64")
    want = np.stack([O.grid_ifft(m, p["nxp"], O.PARAMETER) for m in om])
    e_mean = rel_l2(mean - p["mean_real"], want - p["mean_real"])
    e_cov = rel_l2(cov, oc)
    # realisations = posterior mean + coloured counter noise (streams 3r, 3r + 1, 3r + 2).  The
    # colouring adds the means in the spectral domain, so a realisation carries the FP32 rounding
    # of the mean (~1e-7 of ln Vp ~ 8) through its inverse transform, well above 1e-5 of the
    # coloured noise: the noise part is held to 5e-5, the realisation itself to a few units in
    # the last place of its values
    e_real, e_abs = [], []
    for r in range(2):
        noise = np.stack([O.fill_complex_noise_cb(p["nxp"], p["nyp"], p["nzp"], 500 + case, 3 * r + g)
                          for g in range(3)])
        z, _ = O.colour(oc, noise, "f64")
        real = np.stack([O.grid_ifft(z[g], p["nxp"], O.PARAMETER) for g in range(3)])
        e_real.append(rel_l2(outs[r] - mean, real))
        e_abs.append(float(np.max(np.abs(outs[r] - mean - real)) / np.max(np.abs(outs[r]))))
    print("split-z %dx%dx%d: mean update %.2e, covariance %.2e, realisations (noise part) %.2e %.2e, "
          "max |realisation - mean - colouring| / max |realisation| %.1e %.1e, FP64 re-solves %d"
          % (p["nxp"], p["nyp"], p["nzp"], e_mean, e_cov, e_real[0], e_real[1], e_abs[0], e_abs[1], n64), flush=True)
    assert e_mean < 1e-4 and e_cov < 1e-5 and max(e_real) < 5e-5 and max(e_abs) < 3e-6


def test_split_z_equals_unsplit_to_fp32_rounding(ctx):
    p = P.make_problem(**CASES[0])
    (m1, c1, o1), s1, z1 = _run(ctx, p, 2, 77)
    (m2, c2, o2), s2, z2 = _run(ctx, p, 2, 77, engine=C.FFT_GENERIC)
    assert z1 == 0 and z2 > 0
    assert s1[0] == s2[0] and s1[1] == s2[1] == 0
    e = (rel_l2(m1 - p["mean_real"], m2 - p["mean_real"]), rel_l2(c1, c2), rel_l2(o1[0], o2[0]), rel_l2(o1[1], o2[1]))
    print("split vs unsplit: mean update %.2e, covariance %.2e, realisations %.2e %.2e" % e, flush=True)
    assert e[0] < 2e-5 and e[1] < 2e-6 and e[2] < 1e-5 and e[3] < 1e-5


def test_split_z_fp64_resolves_ill_conditioned_tiles(ctx):
    """S/N 1e8: most inverted wavenumbers need the FP64 re-solve.  Their tiles are left to the
    second (FP64) pass, which must reproduce what the unsplit mixed path gives: the same
    re-solved wavenumbers, the same covariance (it depends on the tables only, so bit for bit) and
    the update to the same accuracy against the FP64 oracle."""
    kw = dict(CASES[0], sn=1e8, seed=44)
    p = P.make_problem(**kw)
    (mean, cov, _), (ninv, nfail, n64), zl = _run(ctx, p, 0, 1)
    assert zl == 0 and nfail == 0 and n64 > ninv // 2
    (mu, cu, _), (ninv2, nfail2, n64u), _ = _run(ctx, p, 0, 1, engine=C.FFT_GENERIC)
    ms, ds = P.forward_spectra(p, np.float64)
    om, _, oc, _, _ = O.posterior(p, ms, ds, None, "f64")
    want = np.stack([O.grid_ifft(m, p["nxp"], O.PARAMETER) for m in om])
    e_mean, e_mean_u = (rel_l2(m - p["mean_real"], want - p["mean_real"]) for m in (mean, mu))
    print("S/N 1e8 split-z: %d of %d re-solved in FP64 (unsplit %d); mean update %.2e (unsplit %.2e), "
          "covariance %.2e (unsplit %.2e)" % (n64, ninv, n64u, e_mean, e_mean_u, rel_l2(cov, oc), rel_l2(cu, oc)),
          flush=True)
    assert ninv2 == ninv and nfail2 == 0 and n64u == n64
    assert np.array_equal(cov, cu)
    assert e_mean < max(5e-3, 2 * e_mean_u)


def test_split_z_reference_precision_keeps_prior_on_failure(ctx):
    # a negative noise variance fails every inverted wavenumber's factorisation: the prior is kept
    # (avoinversion.cpp:1040-1042); the means come back through the transform round trip
    kw = dict(CASES[1], seed=45)
    p = P.make_problem(**kw)
    p = dict(p, sigma0e=-10.0 * np.abs(p["sigma0e"]) - 1.0, prior_var0=p["prior_var0"] * 1e-12)
    (mean, cov, _), (ninv, nfail, n64), zl = _run(ctx, p, 0, 1, precision=C.PREC_REFERENCE)
    assert zl == 0 and ninv == nfail > 0 and n64 == 0
    assert rel_l2(mean, p["mean_real"]) < 1e-6
    ms, ds = P.forward_spectra(p, np.float64)
    _, _, oc, _, onf = O.posterior(p, ms.astype(np.complex64), ds.astype(np.complex64), None, "f32")
    assert onf == nfail and rel_l2(cov, oc) < 1e-6


def test_split_z_colouring_pair_equals_singles_and_unsplit(ctx):
    """Realisations on a split geometry: the colouring carries the inverse z transform's second
    stage, the x pass its first.  A pair equals two single realisations bit for bit, and both
    match the unsplit colouring to FP32 rounding and the FP64 oracle's colouring."""
    p = P.make_problem(**dict(CASES[0], seed=46))
    it = _item(ctx, p, 0, 1)
    C.invert(ctx, it)
    mean, cov = it["mean"], it["cov"]
    mk = lambda: [C.FFTGrid(ctx, p["nx"], p["ny"], p["nz"], p["nxp"], p["nyp"], p["nzp"], C.PARAMETER) for _ in range(3)]
    a, b, s0, s1, u0 = mk(), mk(), mk(), mk(), mk()
    ctx.set_profiling(True)
    ctx.reset_kernel_time()
    C.simulate(ctx, cov, mean, a, seed=9, realisation=4, next_out=b)
    C.simulate(ctx, cov, mean, s0, seed=9, realisation=4)
    C.simulate(ctx, cov, mean, s1, seed=9, realisation=5)
    ctx.sync()
    assert ctx.kernel_time(C.K_FFT_Z)[0] == 0
    ctx.set_profiling(False)
    ctx.set_fused_colour(0)
    try:
        C.simulate(ctx, cov, mean, u0, seed=9, realisation=4)
    finally:
        ctx.set_fused_colour(-1)
    ga, gb, g0, g1, gu = (np.stack([g.download_real() for g in x]) for x in (a, b, s0, s1, u0))
    assert np.array_equal(ga, g0) and np.array_equal(gb, g1)
    m = np.stack([g.download_real() for g in mean])
    e_u = rel_l2(g0 - m, gu - m)
    oc = np.stack([g.download_complex() for g in cov]).astype(np.complex128)
    noise = np.stack([O.fill_complex_noise_cb(p["nxp"], p["nyp"], p["nzp"], 9, 3 * 4 + g) for g in range(3)])
    z, _ = O.colour(oc, noise, "f64")
    want = np.stack([O.grid_ifft(z[g], p["nxp"], O.PARAMETER) for g in range(3)])
    e_o = rel_l2(g0 - m, want)
    print("split-z colouring: pair == singles; vs unsplit %.2e, vs FP64 oracle colouring %.2e" % (e_u, e_o), flush=True)
    assert e_u < 1e-6 and e_o < 1e-5
    for x in (a, b, s0, s1, u0):
        for g in x:
            g.close()
    _close(dict(it, out=[]))
