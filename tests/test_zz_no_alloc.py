"""The allocation-free posterior (cg_posterior_noalloc) against cg_posterior on the same inputs:
bit-identical outputs and statistics on two calls in a row, the second on a smaller grid, in both
precisions (with every inverted wavenumber forced through the FP64 re-solve, which finds them
through the buffer's bit mask instead of a list), and a refusal beyond the buffer's limits."""
import numpy as np
import pytest

import crava_amd as C
import problem as P

pytestmark = pytest.mark.gpu


def _run(ctx, p, mean, data, precision, cond_limit, noalloc):
    nt = p["ntheta"]
    mk = lambda ct: C.FFTGrid(ctx, p["nx"], p["ny"], p["nz"], p["nxp"], p["nyp"], p["nzp"], ct)
    gm, gd, gc = [mk(C.PARAMETER) for _ in range(3)], [mk(C.DATA) for _ in range(nt)], [mk(C.COVARIANCE) for _ in range(6)]
    for g, x in zip(gm, mean):
        g.upload_complex(x)
    for g, x in zip(gd, data):
        g.upload_complex(x)
    for g in gc:
        g.upload_complex(np.zeros_like(mean[0]))
    st = C.posterior(ctx, p, gm, gd, gc, rxy=p["rxy"], rt=p["rt"], lam2d=p["lam2d"], want_residual=True,
                     precision=precision, cond_limit=cond_limit, noalloc=noalloc)
    out = [np.stack([g.download_complex() for g in gs]) for gs in (gm, gd, gc)]
    for g in gm + gd + gc:
        g.close()
    return out, st


@pytest.mark.parametrize("precision,cond_limit", [(C.PREC_REFERENCE, 1e6), (C.PREC_MIXED, 1e6),
                                                  (C.PREC_MIXED, 1e-30)])
def test_noalloc_posterior_equals_posterior(ctx, precision, cond_limit):
    for kw in (dict(nx=40, ny=36, nz=60, seed=41, angles_deg=(0.0, 10.0, 20.0, 30.0, 40.0)),
               dict(nx=16, ny=12, nz=20, seed=42)):  # the second call is smaller
        print("noalloc posterior", precision, cond_limit, kw, flush=True)
        p = P.make_problem(**kw)
        mean, data = P.forward_spectra(p, np.float32)
        want, st_w = _run(ctx, p, mean, data, precision, cond_limit, False)
        got, st_g = _run(ctx, p, mean, data, precision, cond_limit, True)
        assert st_g == st_w
        if cond_limit < 1e-20:
            assert st_g[2] == st_g[0] > 0  # every inverted wavenumber re-solved in FP64
        for a, b in zip(got, want):
            assert np.array_equal(a, b)


def test_noalloc_refuses_grids_beyond_its_buffer(ctx):
    p = P.make_problem(8, 8, 8, nxp=8, nyp=8, nzp=2048, seed=3)
    mk = lambda ct: C.FFTGrid(ctx, 8, 8, 8, 8, 8, 2048, ct)
    gm, gd, gc = [mk(C.PARAMETER) for _ in range(3)], [mk(C.DATA) for _ in range(3)], [mk(C.COVARIANCE) for _ in range(6)]
    for g in gm + gd:
        g.upload_complex(np.zeros((2048, 8, 5), np.complex64))
    with pytest.raises(C.CravaError, match="fixed buffer"):
        C.posterior(ctx, p, gm, gd, gc, rxy=p["rxy"], rt=p["rt"], lam2d=p["lam2d"], noalloc=True)
    for g in gm + gd + gc:
        g.close()
