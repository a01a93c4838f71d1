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


# odd and even nxp (the parity rule of the x = cnxp-1 column), nyp/nzp of 1, the plain case's
# reference padding (84 x 84 x 180, 1.3 M normals)
@pytest.mark.parametrize("shape", [(12, 10, 16), (9, 7, 10), (16, 1, 1), (15, 6, 1), (84, 84, 180)])
def test_reference_noise_stream_is_bit_identical(ctx, shape):
    """The device's parallel generation of the reference's RandomGen + fillInComplexNoise stream
    equals the oracle (itself bit-identical to the reference-compiled RandomGen) bit for bit,
    including the advanced seed, over three consecutive fills (one realisation's seed grids)."""
    nxp, nyp, nzp = shape
    print("reference noise", shape, flush=True)
    g = C.FFTGrid(ctx, nxp, nyp, nzp, nxp, nyp, nzp, C.PARAMETER)
    seed = 23665  # RandomGen's default seed (random.cpp:136)
    state = np.array([seed], dtype=np.uint32)
    for _ in range(3):
        seed = g.fillInComplexNoiseReference(seed)
        want = O.fill_complex_noise(nxp, nyp, nzp, state)
        got = g.download_complex()
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
        assert seed == int(state[0])
    g.close()


@pytest.mark.parametrize("source,fused,spectral", [("grids", -1, False), ("device_reference", -1, False),
                                                   ("counter", 0, False), ("counter", 1, False),
                                                   ("counter", 0, True), ("counter", 1, True),
                                                   ("grids", -1, True)])
def test_realisation_matches_oracle(ctx, source, fused, spectral):
    """grids: the reference stream drawn on the host and uploaded; device_reference: the same
    stream drawn on the device (cg_grid_fill_noise_reference); counter: the counter generator,
    with the colouring as its own kernel (fused 0) or inside the inverse z pass (fused 1).
    spectral: the posterior means handed over as spectra (added to the coloured spectra) instead
    of in the real domain (added after the inverse transform)."""
    ctx.set_fused_colour(fused)
    p = P.make_problem(16, 12, 20, seed=11)
    om, oc = _posterior_cov(p)
    nxp, nyp, nzp = p["nxp"], p["nyp"], p["nzp"]
    mean_real = np.stack([O.grid_ifft(m, nxp, O.PARAMETER, np.float32) for m in om])
    if source in ("grids", "device_reference"):
        st = np.array([23665], dtype=np.uint32)  # the reference stream (RandomGen default seed)
        noise = np.stack([O.fill_complex_noise(nxp, nyp, nzp, st) for _ in range(3)])
    else:
        noise = np.stack([O.fill_complex_noise_cb(nxp, nyp, nzp, 99, 3 * 4 + g) for g in range(3)])
    z, nf = O.colour(oc, noise, "f32")
    want = np.stack([O.grid_ifft(z[g], nxp, O.PARAMETER) + mean_real[g] for g in range(3)])
    gc, gmean, gout = _grids(ctx, p, 6, C.COVARIANCE), _grids(ctx, p, 3, C.PARAMETER), _grids(ctx, p, 3, C.PARAMETER)
    for g in range(6):
        gc[g].upload_complex(oc[g])
    seed = 23665
    for g in range(3):
        if spectral:
            gmean[g].upload_complex(om[g])
        else:
            gmean[g].upload_real(mean_real[g])
        if source == "grids":
            gout[g].upload_complex(noise[g])
        elif source == "device_reference":
            seed = gout[g].fillInComplexNoiseReference(seed)
    fails = C.simulate(ctx, gc, gmean, gout, seed=99, realisation=4,
                       noise_source=C.NOISE_COUNTER if source == "counter" else C.NOISE_GRIDS, stats=True)
    got = np.stack([g.download_real() for g in gout])
    ctx.set_fused_colour(-1)
    assert fails == nf
    # with the mean inside the transformed spectrum, FP32 rounding is relative to mean + noise
    ratio = max(1.0, np.linalg.norm(mean_real) / np.linalg.norm(want - mean_real)) if spectral else 1.0
    assert rel_l2(got - mean_real, want - mean_real) < 1e-5 * ratio
    assert rel_l2(got, want) < 3e-6
    for g in gc + gmean + gout:
        g.close()


@pytest.mark.parametrize("nz,nzp", [(256, 320), (320, 384)])
def test_fused_colour_at_production_depths(ctx, nz, nzp):
    """The colouring fused into the inverse z pass at the z lengths of the single-cube and
    large-cube workloads (workgroups of 8 columns, the last one partly outside the plane) and the
    unfused schedule both match the oracle."""
    p = P.make_problem(11, 7, nz, nzp=nzp, seed=5)
    om, oc = _posterior_cov(p)
    nxp, nyp = p["nxp"], p["nyp"]
    assert ((nxp // 2 + 1) * nyp) % 8 != 0
    mean_real = np.stack([O.grid_ifft(m, nxp, O.PARAMETER, np.float32) for m in om])
    noise = np.stack([O.fill_complex_noise_cb(nxp, nyp, nzp, 7, 3 * 2 + g) for g in range(3)])
    z, nf = O.colour(oc, noise, "f32")
    want = np.stack([O.grid_ifft(z[g], nxp, O.PARAMETER) + mean_real[g] for g in range(3)])
    gc, gmean, gout = _grids(ctx, p, 6, C.COVARIANCE), _grids(ctx, p, 3, C.PARAMETER), _grids(ctx, p, 3, C.PARAMETER)
    for g in range(6):
        gc[g].upload_complex(oc[g])
    for g in range(3):
        gmean[g].upload_real(mean_real[g])
    for fused in (0, 1):
        print("fused colour", nz, nzp, fused, flush=True)
        ctx.set_fused_colour(fused)
        fails = C.simulate(ctx, gc, gmean, gout, seed=7, realisation=2, noise_source=C.NOISE_COUNTER, stats=True)
        got = np.stack([g.download_real() for g in gout])
        assert fails == nf
        assert rel_l2(got - mean_real, want - mean_real) < 1e-5
    ctx.set_fused_colour(-1)
    for g in gc + gmean + gout:
        g.close()


@pytest.mark.parametrize("fused,nz,nzp,r0", [(1, 256, 320, 2), (1, 256, 320, 3), (-1, 256, 320, 2), (0, 256, 320, 2),
                                           (0, 256, 320, 3), (1, 20, None, 2)])
def test_realisation_pair_equals_two_singles(ctx, fused, nz, nzp, r0):
    """cg_simulate_pair: realisations r and r + 1 coloured in one launch (fused) or one after the
    other (unfused), bit for bit the two single calls, failures summed.  From an even r the pair's
    six streams are three whole Philox blocks; from an odd r they straddle four."""
    ctx.set_fused_colour(fused)
    p = P.make_problem(11, 7, nz, nzp=nzp, seed=9)
    print("realisation pair", fused, nz, p["nzp"], r0, flush=True)
    om, oc = _posterior_cov(p)
    nxp = p["nxp"]
    mean_real = np.stack([O.grid_ifft(m, nxp, O.PARAMETER, np.float32) for m in om])
    gc, gmean = _grids(ctx, p, 6, C.COVARIANCE), _grids(ctx, p, 3, C.PARAMETER)
    outs = [_grids(ctx, p, 3, C.PARAMETER) for _ in range(4)]
    for g in range(6):
        gc[g].upload_complex(oc[g])
    for g in range(3):
        gmean[g].upload_real(mean_real[g])
    fp = C.simulate(ctx, gc, gmean, outs[0], seed=5, realisation=r0, stats=True, next_out=outs[1])
    f1 = C.simulate(ctx, gc, gmean, outs[2], seed=5, realisation=r0, stats=True)
    f2 = C.simulate(ctx, gc, gmean, outs[3], seed=5, realisation=r0 + 1, stats=True)
    ctx.set_fused_colour(-1)
    assert fp == f1 + f2
    for a, b in ((outs[0], outs[2]), (outs[1], outs[3])):
        for ga, gb in zip(a, b):
            assert np.array_equal(ga.download_real().view(np.uint32), gb.download_real().view(np.uint32))
    for g in gc + gmean + [g for o in outs for g in o]:
        g.close()


@pytest.mark.parametrize("fused", [1, 0])
def test_realisation_pair_spectral_means_and_noise_grids(ctx, fused):
    """cg_simulate_pair with the means handed over as spectra (added to both coloured spectra)
    equals two single calls bit for bit; with noise grids (the second in next_out) it makes the
    two single calls itself and matches the oracle."""
    ctx.set_fused_colour(fused)
    p = P.make_problem(11, 7, 256, nzp=320, seed=13)
    print("realisation pair, spectral means / noise grids", fused, flush=True)
    om, oc = _posterior_cov(p)
    nxp, nyp, nzp = p["nxp"], p["nyp"], p["nzp"]
    gc, gmean = _grids(ctx, p, 6, C.COVARIANCE), _grids(ctx, p, 3, C.PARAMETER)
    outs = [_grids(ctx, p, 3, C.PARAMETER) for _ in range(4)]
    for g in range(6):
        gc[g].upload_complex(oc[g])
    for g in range(3):
        gmean[g].upload_complex(om[g])
    C.simulate(ctx, gc, gmean, outs[0], seed=8, realisation=6, next_out=outs[1])
    C.simulate(ctx, gc, gmean, outs[2], seed=8, realisation=6)
    C.simulate(ctx, gc, gmean, outs[3], seed=8, realisation=7)
    for a, b in ((outs[0], outs[2]), (outs[1], outs[3])):
        for ga, gb in zip(a, b):
            assert np.array_equal(ga.download_real().view(np.uint32), gb.download_real().view(np.uint32))
    # noise grids: the reference stream for realisation 0 in outs[0], realisation 1 in outs[1]
    st = np.array([23665], dtype=np.uint32)
    noise = [np.stack([O.fill_complex_noise(nxp, nyp, nzp, st) for _ in range(3)]) for _ in range(2)]
    mean_real = np.stack([O.grid_ifft(m, nxp, O.PARAMETER, np.float32) for m in om])
    for g in range(3):
        gmean[g].upload_real(mean_real[g])
        outs[0][g].upload_complex(noise[0][g])
        outs[1][g].upload_complex(noise[1][g])
    C.simulate(ctx, gc, gmean, outs[0], noise_source=C.NOISE_GRIDS, next_out=outs[1])
    ctx.set_fused_colour(-1)
    for s in range(2):
        z, _ = O.colour(oc, noise[s], "f32")
        want = np.stack([O.grid_ifft(z[g], nxp, O.PARAMETER) + mean_real[g] for g in range(3)])
        got = np.stack([g.download_real() for g in outs[s]])
        assert rel_l2(got - mean_real, want - mean_real) < 1e-5
    for g in gc + gmean + [g for o in outs for g in o]:
        g.close()
