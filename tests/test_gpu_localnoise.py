"""Local noise on the device (cg_local_noise_*, AVOInversion::correctVpVsRho,
avoinversion.cpp:2146-2302) against the oracle's restatement: the per-trace matrices in FP64,
the posterior update and the realisation scaling on real-domain grids."""
import numpy as np
import pytest

import crava_amd as C
import oracle as O
from crava_amd import model as M

pytestmark = pytest.mark.gpu


def _case(nt, seed, nx, ny):
    rng = np.random.default_rng(seed)
    a = rng.standard_normal((3, 3))
    sm = 0.01 * (a @ a.T / 3 + 0.5 * np.eye(3))
    r = np.real(np.linalg.cholesky(sm))
    u = np.linalg.qr(rng.standard_normal((3, 3)))[0]
    delta = r @ (u @ np.diag(rng.uniform(0.2, 0.9, 3)) @ u.T) @ r.T
    post = sm - delta
    post = 0.5 * (post + post.T)
    b = rng.standard_normal((nt, nt))
    e = 1e-3 * (b @ b.T / nt + 0.5 * np.eye(nt))
    scales = rng.uniform(1.0, 8.0, (nt, ny, nx))
    scales[:, ny // 2, nx // 3] = 0.75  # each angle's minimum at one trace
    return sm, post, e, scales, rng


@pytest.mark.parametrize("nt,seed", [(1, 6), (2, 7), (3, 1), (5, 2), (8, 3)])
def test_trace_matrices_match_oracle(ctx, nt, seed):
    nx, ny = 13, 9
    sm, post, e, scales, _ = _case(nt, seed, nx, ny)
    G = M.compute_g(sm, post, e)
    print("local noise matrices", nt, seed, flush=True)
    ln = C.LocalNoise(ctx, G, sm, post, e, list(scales))
    sn, kept = ln.matrices()
    ln.close()
    wn, wk = O.local_noise_matrices(G, scales, sm, post, e)
    # with fewer than three angles Sigma_m G^T M^+ G Sigma_m has rank nt < 3: its square root
    # takes sqrt of eigenvalues at rounding level (~1e-17), which any two eigen solvers place
    # differently (~1e-8 in sigmanew)
    tol = 1e-11 if nt >= 3 else 1e-7
    assert np.abs(sn - wn).max() < tol * np.abs(wn).max()
    assert np.abs(kept - wk).max() < 1e-11 * np.abs(wk).max()
    if nt >= 3:  # G reproduces the lag-0 posterior only with three or more angles
        assert np.abs(sn[ny // 2, nx // 3] - np.eye(3)).max() < 1e-9  # the trace at every minimum


def test_posterior_and_realisation_corrections_match_oracle(ctx):
    nx, ny, nz, nxp, nyp, nzp = 21, 14, 30, 24, 16, 40
    nt = 5
    sm, post, e, scales, rng = _case(nt, 4, nx, ny)
    G = M.compute_g(sm, post, e)
    ln = C.LocalNoise(ctx, G, sm, post, e, list(scales))
    sn, kept = ln.matrices()
    rnxp = 2 * (nxp // 2 + 1)
    bg = np.array([8.0, 7.4, 7.8], dtype=np.float32)
    mean = (bg[:, None, None, None] + 0.05 * rng.standard_normal((3, nzp, nyp, rnxp))).astype(np.float32)
    pst = (mean + 0.1 * rng.standard_normal(mean.shape)).astype(np.float32)
    noise = (0.1 * rng.standard_normal(mean.shape)).astype(np.float32)
    mk = lambda: C.FFTGrid(ctx, nx, ny, nz, nxp, nyp, nzp, C.PARAMETER)
    gp, gm, gr = [mk() for _ in range(3)], [mk() for _ in range(3)], [mk() for _ in range(3)]
    for g in range(3):
        gp[g].upload_real(pst[g])
        gm[g].upload_real(mean[g])
        gr[g].upload_real(noise[g])
    ln.correct(gp, gm)
    got = np.stack([g.download_real(padded=True) for g in gp])
    # the oracle's sample arithmetic with the device's matrices: bit for bit; with its own
    # matrices (LAPACK eigen decompositions): within one rounding of the float result
    assert np.array_equal(got, O.correct_vp_vs_rho(pst, mean, nx, ny, nz, sn))
    wn, wk = O.local_noise_matrices(G, scales, sm, post, e)
    want = O.correct_vp_vs_rho(pst, mean, nx, ny, nz, wn)
    diff = np.abs(got.astype(np.float64) - want)
    assert diff.max() <= np.spacing(np.float32(8.0)) and np.mean(got != want) < 1e-3
    ln.correct_realisation(gr, gp)
    real = np.stack([g.download_real(padded=True) for g in gr])
    assert np.array_equal(real, O.correct_realisation(noise, got, nx, ny, nz, kept))
    want_r = O.correct_realisation(noise, got, nx, ny, nz, wk)
    assert np.abs(real.astype(np.float64) - want_r).max() <= np.spacing(np.float32(8.0))
    ln.close()
    for g in gp + gm + gr:
        g.close()


def test_rejects_grids_of_another_box(ctx):
    sm, post, e, scales, _ = _case(3, 5, 6, 5)
    ln = C.LocalNoise(ctx, M.compute_g(sm, post, e), sm, post, e, list(scales))
    gs = [C.FFTGrid(ctx, 7, 5, 8, 8, 6, 10, C.PARAMETER) for _ in range(6)]
    with pytest.raises(C.CravaError):
        ln.correct(gs[:3], gs[3:])
    ln.close()
    for g in gs:
        g.close()
