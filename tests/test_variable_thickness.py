"""The non-constant-thickness branch of the inversion (CPU): its per-k_z tables
(avoinversion.cpp:976-1014, fillInverseAbskWRobust :1296-1322) from the library against the oracle,
and the oracle pinned by closed forms."""
import math

import numpy as np
import pytest

import crava_amd as C
import oracle as O


def _wavelets(nzp, nt, seed):
    rng = np.random.default_rng(seed)
    k = np.arange(nzp)
    t = np.where(k < nzp // 2 + 1, k, k - nzp).astype(np.float64)
    out = []
    for _ in range(nt):
        sh, wd = rng.uniform(-1.0, 1.0), rng.uniform(2.0, 4.0)
        out.append((1 - 2 * ((t - sh) / wd) ** 2) * np.exp(-((t - sh) / wd) ** 2) + 0.05 * rng.standard_normal(nzp))
    return np.array(out, dtype=np.float32)


@pytest.mark.parametrize("nzp", [30, 45, 64, 180])
def test_oracle_tables_closed_forms(nzp):
    nt = 3
    wav = _wavelets(nzp, nt, nzp)
    A = np.array([[0.5, -0.1, 0.4], [0.55, -0.2, 0.35], [0.6, -0.3, 0.3]])
    K, e1, e2, e3, inv = O.kz_tables_variable_thickness(nzp - 6, nzp, 4.0, wav, A, 5.0, 55.0, 0.8)
    w = 2 * np.pi * np.arange(nzp) / nzp
    D = np.exp(1j * w) - 1
    D3 = 1j * np.sin(w) * (2 * np.cos(w) - 2)
    for l in range(nt):
        wl = wav[l].astype(np.float64)
        assert np.allclose(K[:, l, :], D[:, None] * A[l][None, :], atol=1e-12)
        assert np.allclose(e1[:, l], 1.0 / np.linalg.norm(wl), rtol=1e-12)
        S = np.fft.fft(wl) * D3
        s_norm = math.sqrt(np.sum(np.abs(S) ** 2) / nzp)  # Parseval
        assert np.allclose(e2[:, l], D3 / s_norm, atol=1e-10 * np.max(np.abs(D3)) / s_norm)
        # the robust inverse is at most 1 / |W(k)|
        assert np.all(e3[:, l].real <= 1.0 / np.abs(np.fft.fft(wl)) * (1 + 1e-12))
    # band: f = 1000 / (dt nzp) min(k, nzp - k) with the low cut scaled by the minimum thickness
    f = 1000.0 / (4.0 * nzp) * np.minimum(np.arange(nzp), nzp - np.arange(nzp))
    assert np.array_equal(inv.astype(bool), (f > 5.0 * 0.8) & (f < 55.0))


@pytest.mark.parametrize("nzp", [24, 25])
def test_oracle_robust_inverse_of_a_spike(nzp):
    """A spike wavelet has a flat spectrum (1): stretched by 0.999 it reads 1 / 0.999 where k / 0.999
    stays within the half spectrum and 0 beyond, so e3 = 0.999 there and 1 elsewhere."""
    w = np.zeros((1, nzp), dtype=np.float32)
    w[0, 0] = 1.0
    _, _, _, e3, _ = O.kz_tables_variable_thickness(nzp - 4, nzp, 4.0, w, np.ones((1, 3)))
    cnzp = nzp // 2 + 1
    for k in range(nzp):
        kk = k if k < cnzp else nzp - k
        want = 0.999 if kk / 0.999 < cnzp else 1.0
        assert abs(e3[k, 0].real - want) < 1e-12, k


@pytest.mark.parametrize("nzp,nt", [(64, 3), (180, 5), (320, 6)])
def test_library_tables_match_oracle(nzp, nt):
    wav = _wavelets(nzp, nt, 100 + nzp)
    rng = np.random.default_rng(nzp)
    A = rng.uniform(-0.5, 0.8, (nt, 3)).astype(np.float32)
    got = C.build_kz_tables(nzp - 20, nzp, 4.0, wav, A, 5.0, 55.0, 0.7, constant_thickness=False)
    want = O.kz_tables_variable_thickness(nzp - 20, nzp, 4.0, wav, A, 5.0, 55.0, 0.7)
    # e3 reads the spectrum at k / 0.999 with the position and weights in float, as the reference
    # does (the oracle in double): up to 1e-5 of the largest value
    for g, w, name, tol in zip(got[:4], want[:4], ("K", "e1", "e2", "e3"), (2e-6, 2e-6, 2e-6, 1e-5)):
        err = np.max(np.abs(g - w)) / np.max(np.abs(w))
        assert err < tol, (name, err)
    assert np.array_equal(got[4], want[4])


# ------------------------------------------------------------------------------ data rescaling
def _trace_grid(nx, ny, nz, nxp, nyp, nzp, seed):
    rng = np.random.default_rng(seed)
    rv = np.zeros((nzp, nyp, 2 * (nxp // 2 + 1)))
    rv[:, :, :nxp] = rng.standard_normal((nzp, nyp, nxp))
    return rv


def _bandpass(x, nz, nzp, dt, lo, hi, rel=1.0):
    """The closed form of a flat-spectrum (spike) wavelet: the taper beyond nz, then the band."""
    X = np.fft.rfft(x, axis=0)
    k = np.arange(nzp // 2 + 1)
    dF = nz * 1000.0 / (rel * nz * dt * nzp)
    band = (dF * k < hi) & (dF * k > lo)
    return np.fft.irfft(X * band[:, None], n=nzp, axis=0)


def _taper(nz, nzp):
    t = np.ones(nzp)
    for k in range(nz + 1, nzp):
        d = O.dist_to_boundary(k, nz, nzp)
        t[k] = max(1.0 - d * d, 0.0)
    return t


def test_oracle_divide_with_a_spike_wavelet_is_a_band_pass():
    """A spike wavelet has a flat spectrum: with no prior signal term the adjustment factor is 1, so
    dividing by it tapers the z padding and keeps the band; an integer time shift of n samples in the
    local wavelet delays every trace by n samples."""
    nx, ny, nz, nxp, nyp, nzp, dt = 5, 4, 24, 8, 6, 32, 4.0
    rv = _trace_grid(nx, ny, nz, nxp, nyp, nzp, 1)
    spike = np.zeros(nzp)
    spike[0] = 1.0
    rel = np.ones((nyp, nxp))
    got = O.scale_wavelet_filter(rv, nx, ny, nz, nxp, dt, spike, "divide", rel, 5.0, 55.0)
    x = rv[:, :, :nxp] * _taper(nz, nzp)[:, None, None]
    want = _bandpass(x.reshape(nzp, -1), nz, nzp, dt, 5.0, 55.0).reshape(nzp, nyp, nxp)
    assert np.allclose(got[:, :, :nxp], want, atol=1e-12)
    assert np.array_equal(got[:, :, nxp:], rv[:, :, nxp:])  # the row padding is not a trace
    n = 3
    shifted = O.scale_wavelet_filter(rv, nx, ny, nz, nxp, dt, spike, "divide", rel, 5.0, 55.0,
                                     shift=np.full((ny, nx), n * dt))
    assert np.allclose(shifted[:, :, :nxp], np.roll(want, n, axis=0), atol=1e-12)


def test_oracle_divide_undoes_multiply_when_the_signal_dominates():
    """With a prior signal term far above the noise the adjustment factor is 1 / conj(W_loc): dividing
    undoes the multiplication by conj(W_loc) inside the band (no z padding beyond nz to taper; bins
    where |W|^2 falls below 5% of its scale are adjusted and left out)."""
    nx, ny, nzp, dt = 4, 3, 48, 4.0
    nz = nzp - 1
    rng = np.random.default_rng(2)
    k = np.arange(nzp)
    t = np.where(k < nzp // 2 + 1, k, k - nzp).astype(np.float64)
    w = (1 - 2 * (t / 2.5) ** 2) * np.exp(-(t / 2.5) ** 2)
    rv = _trace_grid(nx, ny, nz, nx, ny, nzp, 3)
    rel = np.ones((ny, nx))
    gain = rng.uniform(0.8, 1.25, (ny, nx))
    shift = rng.uniform(-2.0, 2.0, (ny, nx))
    mult = O.scale_wavelet_filter(rv, nx, ny, nz, nx, dt, w, "multiply", rel, gain=gain, shift=shift)
    tau = np.full(nzp // 2 + 1, 1e12, dtype=complex)
    back = O.scale_wavelet_filter(mult, nx, ny, nz, nx, dt, w, "divide", rel, 5.0, 55.0, tau_spec=tau,
                                  gain=gain, shift=shift)
    W = np.fft.rfft(w)
    modW = np.sum(w ** 2) * ((nzp // 2) * 1000.0 / (dt * nzp)) / 55.0
    kk = np.arange(nzp // 2 + 1)
    dF = 1000.0 / (dt * nzp)
    keep = (dF * kk > 5.0) & (dF * kk < 55.0) & (np.abs(W) ** 2 * gain.min() ** 2 >= 0.05 * modW)
    Xb = np.fft.rfft(back[:, :, :nx], axis=0)
    X0 = np.fft.rfft(rv[:, :, :nx], axis=0)
    assert np.allclose(Xb[keep], X0[keep], rtol=1e-6, atol=1e-6 * np.abs(X0).max())
    assert np.allclose(Xb[~((dF * kk > 5.0) & (dF * kk < 55.0))], 0.0, atol=1e-9)


@pytest.mark.gpu
@pytest.mark.parametrize("shape", [(20, 16, 30, 24, 20, 40), (15, 13, 25, 21, 16, 36)])
@pytest.mark.parametrize("mode", ["divide", "multiply"])
def test_device_scale_wavelet_matches_oracle(ctx, shape, mode):
    nx, ny, nz, nxp, nyp, nzp = shape
    print("scale wavelet", mode, shape, flush=True)
    dt = 4.0
    rng = np.random.default_rng(nx + nzp)
    rv = _trace_grid(nx, ny, nz, nxp, nyp, nzp, 7).astype(np.float32)
    k = np.arange(nzp)
    t = np.where(k < nzp // 2 + 1, k, k - nzp).astype(np.float64)
    w = ((1 - 2 * ((t - 0.4) / 3.0) ** 2) * np.exp(-((t - 0.4) / 3.0) ** 2)).astype(np.float32)
    rel = rng.uniform(0.6, 1.0, (nyp, nxp))
    shift = rng.uniform(-3.0, 3.0, (ny, nx)).astype(np.float32)
    gain = rng.uniform(0.7, 1.3, (ny, nx)).astype(np.float32)
    stretch = rng.uniform(0.9, 1.1, (ny, nx)).astype(np.float32)
    A = np.array([0.55, -0.12, 0.42], dtype=np.float32)
    S0 = np.array([[0.01, 0.0105, 0.002], [0.0105, 0.0225, 0.00225], [0.002, 0.00225, 0.0025]])
    ct = np.exp(-3.0 * dt * np.minimum(k, nzp - k) / 40.0).astype(np.float32)
    ct[np.minimum(k, nzp - k) >= nz] = 0.0  # getPriorCorrTFiltered keeps lags < nz
    g = C.FFTGrid(ctx, nx, ny, nz, nxp, nyp, nzp, C.DATA)
    g.upload_real(rv)
    mcode = C.SCALE_WAVELET_DIVIDE if mode == "divide" else C.SCALE_WAVELET_MULTIPLY
    g.scale_wavelet(mcode, dt, w, rel, 5.0, 55.0, 0.1, 0.8, corr_t=ct, A=A, prior_var0=S0, shift_ms=shift,
                    gain=gain, stretch=stretch)
    got = g.download_real(padded=True)
    g.close()
    tau = O.reflection_time_covariance_spectrum(ct, A, S0)
    want = O.scale_wavelet_filter(rv, nx, ny, nz, nxp, dt, w, mode, rel, 5.0, 55.0, 0.1, 0.8, tau_spec=tau,
                                  shift=shift, gain=gain, stretch=stretch)
    err = np.linalg.norm(got[:, :, :nxp] - want[:, :, :nxp]) / np.linalg.norm(want[:, :, :nxp])
    print("scale wavelet %s %s: relative L2 %.2e" % (mode, shape, err), flush=True)
    assert err < 1e-5
