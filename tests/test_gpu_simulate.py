"""GPU realisations (cg_simulate, counter-based noise) vs the oracle."""
import numpy as np
import pytest

import crava_amd as C
import oracle as O
import problem as P
from conftest import rel_l2

pytestmark = pytest.mark.gpu


def _posterior_cov(p):
    mean, data = P.forward_spectra(p, np.float32)
    om, od, oc, _, _ = O.posterior(p, mean, data, None, "f32")
    return om, oc


def _grids(ctx, p, n, ct):
    return [C.FFTGrid(ctx, p["nx"], p["ny"], p["nz"], p["nxp"], p["nyp"], p["nzp"], ct) for _ in range(n)]


@pytest.mark.parametrize("shape", [(12, 10, 16), (9, 7, 10), (16, 16, 8), (10, 1, 6)])
def test_counter_noise_matches_oracle(ctx, shape):
    nxp, nyp, nzp = shape
    g = C.FFTGrid(ctx, nxp, nyp, nzp, nxp, nyp, nzp, C.PARAMETER)
    g.fillInComplexNoise(1234567, 5)
    got = g.download_complex()
    want = O.fill_complex_noise_cb(nxp, nyp, nzp, 1234567, 5)
    assert np.max(np.abs(got - want)) < 5e-6 * max(1.0, np.max(np.abs(want)))
    g.close()


@pytest.mark.parametrize("source", ["grids", "counter"])
def test_realisation_matches_oracle(ctx, source):
    p = P.make_problem(16, 12, 20, seed=11)
    om, oc = _posterior_cov(p)
    nxp, nyp, nzp = p["nxp"], p["nyp"], p["nzp"]
    mean_real = np.stack([O.grid_ifft(m, nxp, O.PARAMETER, np.float32) for m in om])
    if source == "grids":
        st = np.array([23665], dtype=np.uint32)  # the reference stream (RandomGen default seed)
        noise = np.stack([O.fill_complex_noise(nxp, nyp, nzp, st) for _ in range(3)])
    else:
        noise = np.stack([O.fill_complex_noise_cb(nxp, nyp, nzp, 99, 3 * 4 + g) for g in range(3)])
    z, nf = O.colour(oc, noise, "f32")
    want = np.stack([O.grid_ifft(z[g], nxp, O.PARAMETER) + mean_real[g] for g in range(3)])
    gc, gmean, gout = _grids(ctx, p, 6, C.COVARIANCE), _grids(ctx, p, 3, C.PARAMETER), _grids(ctx, p, 3, C.PARAMETER)
    for g in range(6):
        gc[g].upload_complex(oc[g])
    for g in range(3):
        gmean[g].upload_real(mean_real[g])
        if source == "grids":
            gout[g].upload_complex(noise[g])
    fails = C.simulate(ctx, gc, gmean, gout, seed=99, realisation=4,
                       noise_source=C.NOISE_GRIDS if source == "grids" else C.NOISE_COUNTER, stats=True)
    got = np.stack([g.download_real() for g in gout])
    assert fails == nf
    assert rel_l2(got - mean_real, want - mean_real) < 1e-5
    for g in gc + gmean + gout:
        g.close()
