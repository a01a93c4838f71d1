"""Host-side setup of the product (crava_amd.model, cg_build_kz_tables, cg_factorable) against
the oracle's restatement; host-only entry points, no GPU needed."""
import math

import numpy as np
import pytest

import crava_amd as C
import oracle as O
from crava_amd import model as M


def test_kz_tables_match_oracle():
    for nz, nzp, dt, ang, f in [(100, 120, 4.0, (0.0, 15.0, 30.0), 25.0), (256, 320, 4.0, (0, 10, 20, 30, 40), 30.0),
                                (50, 64, 2.0, (5.0,), 40.0)]:
        A = M.aki_richards(ang, 0.5)
        w = np.stack([M.wavelet_fft_order(nzp, dt, f, s) for s in np.linspace(0.8, 1.2, len(ang))])
        got = C.build_kz_tables(nz, nzp, dt, w, A)
        want = O.kz_tables(nz, nzp, dt, w, A)
        for g, o in zip(got[:4], want[:4]):
            assert np.max(np.abs(g - o)) < 1e-5 * max(1.0, np.max(np.abs(o)))
        assert np.array_equal(got[4], want[4])


def test_model_setup_matches_oracle():
    assert M.padding_size(512, 0.125) == O.lib().orc_padding_size(512, 0.125) == 576
    for a in (0.0, 12.5, 40.0):
        row = np.zeros(3)
        O.lib().orc_aki_richards(math.radians(a), 0.5, O._ptr(row, O.ctypes.c_double))
        assert np.allclose(M.aki_richards([a], 0.5)[0], row, atol=1e-7)
    cxy = M.spherical_corr_xy(40, 30, 25.0, 25.0, 600.0)
    assert np.allclose(cxy, O.corr_xy(40, 30, 25.0, 25.0, 600.0), atol=1e-6)
    ct = M.circ_corr_t(4.0, 40.0, 180, 5.0)
    n_corr = 180 // 2 + 1
    ref = O.circ_corr_t(np.exp(-3.0 * np.float32(4.0) * np.arange(n_corr + 1) / np.float32(40.0)),
                        int(math.floor(5.0 * 180 * 0.001 * 4.0)), 180)
    assert np.allclose(ct, ref, atol=1e-6)
    w = M.wavelet_fft_order(64, 4.0, 25.0)
    assert w[0] == 1.0 and abs(w[1] - O.lib().orc_ricker(4.0, 25.0)) < 1e-7 and w[-1] == w[1]


def test_error_covariance_is_positive_definite():
    S = M.angular_error_covariance([1.0, 2.0, 0.5, 1.5], [0, 10, 20, 30])
    assert np.allclose(S, S.T) and np.linalg.eigvalsh(S).min() > 0
    assert S[0, 1] == pytest.approx(2.0 * math.exp(-3.0))


def _cross_cov(nzp, seed, indefinite=False):
    """Nine circular covariance series of a 3-variate stationary process: c_jk(t) = sum_l
    B_jl B_kl r_l(t) with r_l symmetric autocorrelations of positive spectra (so every 3 x 3
    spectrum is positive definite and the pairs are Hermitian, c_jk(t) = c_kj(-t)).  With
    `indefinite`, one component's spectrum is made negative at some frequencies."""
    rng = np.random.default_rng(seed)
    f = np.fft.rfftfreq(nzp)
    B = rng.normal(size=(3, 3))
    specs = []
    for l in range(3):
        s = 1.0 / (1.0 + (f / (0.05 + 0.1 * l)) ** 2)
        if indefinite and l == 2:
            s = s - 0.6 * s.max() * (f > 0.2)
        specs.append(s)
    r = np.stack([np.fft.irfft(s, n=nzp) for s in specs])
    return np.einsum("jl,kl,lt->jkt", B, B, r).astype(np.float32)


def test_circ_auto_cov_pos_def_oracle_pins():
    """The oracle: a positive-definite spectrum comes back unchanged; an indefinite one becomes
    D |M| D, |M| from numpy's Hermitian eigendecomposition (eigh, not the general eig the oracle
    restates), and every output spectrum is positive semi-definite."""
    nzp = 64
    c = _cross_cov(nzp, 1)
    assert np.max(np.abs(O.make_circ_auto_cov_pos_def(c) - c)) < 1e-5 * np.max(np.abs(c))
    c = _cross_cov(nzp, 2, indefinite=True)
    got = O.make_circ_auto_cov_pos_def(c)
    d = np.sqrt(np.array([c[i, i, 0] for i in range(3)], dtype=np.float64))
    spec = np.fft.rfft(c.astype(np.float64), axis=2)
    want = np.empty_like(spec)
    neg = 0
    for m in range(spec.shape[2]):
        M = spec[:, :, m] / d[:, None] / d[None, :]
        w, V = np.linalg.eigh(M)
        neg += int((w < 0).sum())
        want[:, :, m] = d[:, None] * (V @ np.diag(np.abs(w)) @ V.conj().T) * d[None, :]
    assert neg > 0  # the case really is indefinite
    assert np.max(np.abs(got - np.fft.irfft(want, n=nzp, axis=2))) < 1e-9 * np.max(np.abs(got))
    gs = np.fft.rfft(got, axis=2)
    for m in range(gs.shape[2]):
        assert np.linalg.eigvalsh(gs[:, :, m]).min() > -1e-9 * np.abs(gs[:, :, m]).max()


@pytest.mark.parametrize("nzp", [64, 45, 320])
def test_circ_auto_cov_pos_def_matches_oracle(nzp):
    """cg_make_circ_auto_cov_pos_def (host code of the library) against the oracle, even and odd
    nzp, definite and indefinite spectra (float series, float spectra as FFTW returns them)."""
    for indefinite in (False, True):
        c = _cross_cov(nzp, 3 + nzp, indefinite)
        got = C.make_circ_auto_cov_pos_def(c)
        want = O.make_circ_auto_cov_pos_def(c)
        assert np.max(np.abs(got - want)) < 2e-6 * np.max(np.abs(want)), (nzp, indefinite)
    bad = _cross_cov(nzp, 4)
    bad[1, 1, 0] = 0.0
    with pytest.raises(C.CravaError):
        C.make_circ_auto_cov_pos_def(bad)
