"""GPU per-wavenumber posterior (cg_posterior) vs the oracle on identical spectra."""
import numpy as np
import pytest

import crava_amd as C
import oracle as O
import problem as P
from conftest import rel_l2

pytestmark = pytest.mark.gpu


def _grids(ctx, p, n, ct):
    return [C.FFTGrid(ctx, p["nx"], p["ny"], p["nz"], p["nxp"], p["nyp"], p["nzp"], ct) for _ in range(n)]


def _run_gpu(ctx, p, mean, data, cov=None, lam_grid=None, prior_mode=C.PRIOR_SEPARABLE,
             precision=C.PREC_REFERENCE, want_residual=True):
    nt = p["ntheta"]
    gm, gd, gc = _grids(ctx, p, 3, C.PARAMETER), _grids(ctx, p, nt, C.DATA), _grids(ctx, p, 6, C.COVARIANCE)
    for g, s in zip(gm, mean):
        g.upload_complex(s)
    for g, s in zip(gd, data):
        g.upload_complex(s)
    for g in range(6):
        gc[g].upload_complex(cov[g] if cov is not None else np.zeros_like(mean[0]))
    ge = None
    if lam_grid is not None:
        ge = _grids(ctx, p, 1, C.COVARIANCE)[0]
        ge.upload_complex(lam_grid)
    stats = C.posterior(ctx, p, gm, gd, gc, prior_mode=prior_mode, rxy=p["rxy"], rt=p["rt"], err_corr=ge,
                        lam2d=None if ge is not None else p["lam2d"], want_residual=want_residual,
                        precision=precision)
    out = (np.stack([g.download_complex() for g in gm]), np.stack([g.download_complex() for g in gd]),
           np.stack([g.download_complex() for g in gc]), stats)
    for g in gm + gd + gc + ([ge] if ge is not None else []):
        g.close()
    return out


CASES = [
    dict(nx=16, ny=12, nz=20, seed=1),
    dict(nx=20, ny=10, nz=30, seed=2, angles_deg=(0.0, 10.0, 20.0, 30.0, 40.0), peak_hz=30.0,
         wavelet_scales=[0.9, 1.1, 1.0, 0.85, 1.15], sn_per_angle=[3.5, 7.0, 5.0, 4.0, 6.5]),
    dict(nx=12, ny=12, nz=24, seed=3, angles_deg=(0.0, 8.0, 16.0, 24.0, 32.0, 40.0, 48.0, 56.0)),
    dict(nx=9, ny=7, nz=25, seed=4, angles_deg=(0.0, 30.0)),
    dict(nx=10, ny=6, nz=18, seed=12, angles_deg=(20.0,)),  # one angle: 1 x 1 systems
]


@pytest.mark.parametrize("case", range(len(CASES)))
def test_separable_prior_matches_oracle_f32(ctx, case):
    kw = dict(CASES[case])
    print("posterior case", kw, flush=True)
    p = P.make_problem(**kw)
    mean, data = P.forward_spectra(p, np.float32)
    om, od, oc, oflags, onf = O.posterior(p, mean, data, None, "f32")
    gm, gd, gc, (ninv, nfail, n64) = _run_gpu(ctx, p, mean, data)
    assert ninv == int(p["invert"].sum()) * p["nyp"] * p["cnxp"]
    assert nfail == onf
    e_upd = rel_l2(gm - mean, om - mean)
    print("update vs the reference-order FP32 oracle %.2e" % e_upd, flush=True)
    assert rel_l2(gm, om) < 1e-5
    assert e_upd < 1e-5  # the update itself, ten times inside a 1e-4 agreement with the reference
    assert rel_l2(gc, oc) < 1e-5
    assert rel_l2(gd, od) < 1e-5
    # pass-through planes are untouched / the prior
    out = np.where(p["invert"] == 0)[0]
    assert np.array_equal(gm[:, out], mean[:, out])
    assert np.array_equal(gd[:, out], data[:, out])


def test_covariance_grid_prior_and_errcorr_grid_match_oracle(ctx):
    p = P.make_problem(16, 12, 20, seed=5)
    mean, data = P.forward_spectra(p, np.float32)
    covs = np.stack([O.grid_fft(g, O.COVARIANCE, np.float32) for g in P.prior_cov_grids(p)])
    lam = O.grid_fft(P.err_corr_grid(p), O.COVARIANCE, np.float32)
    q = dict(p, prior_mode=0, lam_grid=lam)
    om, od, oc, _, onf = O.posterior(q, mean, data, covs, "f32")
    gm, gd, gc, (_, nfail, _) = _run_gpu(ctx, q, mean, data, cov=covs, lam_grid=lam, prior_mode=C.PRIOR_GRIDS)
    assert nfail == onf
    assert rel_l2(gm, om) < 1e-5 and rel_l2(gc, oc) < 1e-5 and rel_l2(gd, od) < 1e-5


def test_mixed_precision_matches_fp64_oracle(ctx):
    p = P.make_problem(20, 10, 30, seed=6, angles_deg=(0.0, 10.0, 20.0, 30.0, 40.0), peak_hz=30.0)
    mean, data = P.forward_spectra(p, np.float32)
    om, od, oc, _, _ = O.posterior(p, mean.astype(np.complex128), data.astype(np.complex128), None, "f64")
    gm, gd, gc, (ninv, nfail, n64) = _run_gpu(ctx, p, mean, data, precision=C.PREC_MIXED)
    assert nfail == 0
    print("mixed update vs FP64 %.2e" % rel_l2(gm - mean, om - mean), flush=True)
    assert rel_l2(gm - mean, om - mean) < 1e-5
    assert rel_l2(gc, oc) < 1e-5
    assert rel_l2(gd, od) < 1e-4
    assert n64 <= ninv


def test_ill_conditioned_wavenumbers_resolved_in_fp64(ctx):
    """Seismic with S/N 1e8 on five angles: M = K Sigma K^H + noise is a rank-3 signal over a
    tiny noise floor.  FP32 throughout (the reference's arithmetic) fails the Cholesky
    factorisation at thousands of wavenumbers (keeping the prior there) and loses the update and
    the covariance; the mixed mode re-solves every wavenumber whose pivot ratio exceeds 1e6 (or
    whose FP32 factorisation fails) in FP64.  The covariance then matches the FP64 oracle; the
    update keeps an error of about eps |d| / |d - K m| from the residual of the wavenumbers left
    in FP32 (cancellation in the data, not the conditioning of M)."""
    sn = 1e8
    p = P.make_problem(20, 10, 30, seed=6, angles_deg=(0.0, 10.0, 20.0, 30.0, 40.0), peak_hz=30.0, sn=sn)
    mean, data = P.forward_spectra(p, np.float32)
    om, od, oc, _, _ = O.posterior(p, mean.astype(np.complex128), data.astype(np.complex128), None, "f64")
    rm, rd, rc, (ninv, rfail, _) = _run_gpu(ctx, p, mean, data, precision=C.PREC_REFERENCE)
    gm, gd, gc, (ninv2, nfail, n64) = _run_gpu(ctx, p, mean, data, precision=C.PREC_MIXED)
    e_ref = rel_l2(rm - mean, om - mean), rel_l2(rc, oc)
    e_mix = rel_l2(gm - mean, om - mean), rel_l2(gc, oc), rel_l2(gd, od)
    print("S/N %.0e: FP32 update %.2e cov %.2e (%d failures); mixed update %.2e cov %.2e residual %.2e, "
          "%d of %d wavenumbers re-solved in FP64" % (sn, e_ref[0], e_ref[1], rfail, e_mix[0], e_mix[1], e_mix[2],
                                                       n64, ninv2), flush=True)
    assert ninv == ninv2 and rfail > 1000 and nfail == 0 and n64 > ninv // 2
    assert e_ref[1] > 1.0 and e_mix[1] < 1e-5
    assert e_ref[0] > 0.5 and e_mix[0] < 5e-3 and e_mix[2] < 2e-3


def test_all_wavenumbers_through_fp64_pass(ctx):
    p = P.make_problem(16, 12, 20, seed=7)
    mean, data = P.forward_spectra(p, np.float32)
    om, od, oc, _, _ = O.posterior(p, mean.astype(np.complex128), data.astype(np.complex128), None, "f64")
    nt = p["ntheta"]
    gm, gd, gc = _grids(ctx, p, 3, C.PARAMETER), _grids(ctx, p, nt, C.DATA), _grids(ctx, p, 6, C.COVARIANCE)
    for g, s in zip(gm, mean):
        g.upload_complex(s)
    for g, s in zip(gd, data):
        g.upload_complex(s)
    for g in gc:
        g.upload_complex(np.zeros_like(mean[0]))
    ninv, nfail, n64 = C.posterior(ctx, p, gm, gd, gc, rxy=p["rxy"], rt=p["rt"], lam2d=p["lam2d"],
                                   precision=C.PREC_MIXED, cond_limit=1e-30)
    assert n64 == ninv and nfail == 0
    rm = np.stack([g.download_complex() for g in gm])
    rc = np.stack([g.download_complex() for g in gc])
    assert rel_l2(rm - mean, om - mean) < 1e-6
    assert rel_l2(rc, oc) < 1e-6


def test_cholesky_failure_keeps_prior(ctx):
    # a negative noise variance makes M[0][0] <= 0 at every inverted wavenumber: the reference
    # keeps the prior there (avoinversion.cpp:1040-1042, lib_matr.c:100-102)
    p = P.make_problem(12, 8, 16, seed=8)
    p = dict(p, sigma0e=-10.0 * np.abs(p["sigma0e"]) - 1.0, prior_var0=p["prior_var0"] * 1e-12)
    mean, data = P.forward_spectra(p, np.float32)
    om, od, oc, oflags, onf = O.posterior(p, mean, data, None, "f32")
    gm, gd, gc, (ninv, nfail, _) = _run_gpu(ctx, p, mean, data)
    assert onf == ninv == nfail > 0
    assert np.array_equal(gm, mean) and np.array_equal(gd, data)
    assert rel_l2(gc, oc) < 1e-6


def test_no_residual_leaves_data(ctx):
    p = P.make_problem(16, 12, 20, seed=9)
    mean, data = P.forward_spectra(p, np.float32)
    gm, gd, gc, _ = _run_gpu(ctx, p, mean, data, want_residual=False)
    assert np.array_equal(gd, data)
