"""Host-side model setup for the inversion: the small tables the GPU path consumes.

Each function restates the reference's setup step it replaces (paths relative to the CRAVA
source tree).  These run once per inversion on the host, in numpy; the per-wavenumber and
per-cell work is on the GPU.
"""
import math

import numpy as np

from . import build_kz_tables, lib


def padding_size(n, pad_fraction):
    """MultiIntervalGrid::FindPaddingSize (multiintervalgrid.cpp:922-928): smallest 2^a 3^b 5^c 7^d
    >= ceil(n (1 + pad))."""
    return int(lib().cg_factorable(int(math.ceil(n * (np.float32(1.0) + pad_fraction)))))


def aki_richards(angles_deg, vsvp):
    """PP rows of CommonData::SetupDefaultReflectionMatrix (commondata.cpp:2915-2942)."""
    A = np.zeros((len(angles_deg), 3), dtype=np.float32)
    g2 = vsvp * vsvp
    for l, a in enumerate(angles_deg):
        t = math.radians(a)
        s2 = math.sin(t) ** 2
        A[l] = [(1.0 + math.tan(t) ** 2) / 2.0, -4 * g2 * s2, (1.0 - 4.0 * g2 * s2) / 2.0]
    return A


def ricker(t_ms, peak_hz):
    """Wavelet::Ricker (wavelet.cpp:1484-1488)."""
    c = math.pi * math.pi * peak_hz * peak_hz * np.asarray(t_ms, dtype=np.float64) ** 2 * 1e-6
    return (1 - 2 * c) * np.exp(-c)


def ormsby(t_ms, f1, f2, f3, f4):
    """Zero-phase Ormsby band-pass wavelet with corner frequencies f1 < f2 < f3 < f4 (Hz), peak 1
    (the configs' band-limited wavelet; the reference reads such wavelets from file)."""
    t = np.asarray(t_ms, dtype=np.float64) * 1e-3

    def term(f):
        return np.pi * f * f * np.sinc(f * t) ** 2  # numpy sinc(x) = sin(pi x) / (pi x)

    w = (term(f4) - term(f3)) / (f4 - f3) - (term(f2) - term(f1)) / (f2 - f1)
    w0 = (np.pi * (f4 * f4 - f3 * f3) / (f4 - f3)) - (np.pi * (f2 * f2 - f1 * f1) / (f2 - f1))
    return w / w0


def wavelet_fft_order(nzp, dt_ms, peak_hz, scale=1.0, ormsby_hz=None):
    """A zero-phase wavelet sampled at z = dt k (k <= nzp/2) and dt (k - nzp) otherwise: the FFT-order
    layout Wavelet::resample produces (wavelet.cpp:424-475), sampled analytically.  Ricker of the
    peak frequency, or Ormsby with the corner frequencies `ormsby_hz`."""
    k = np.arange(nzp)
    t = np.where(k < nzp // 2 + 1, k, k - nzp) * dt_ms
    w = ricker(t, peak_hz) if ormsby_hz is None else ormsby(t, *ormsby_hz)
    return (scale * w).astype(np.float32)


def sigma0(std, corr):
    """3x3 prior covariance of (ln Vp, ln Vs, ln rho) from stds and correlations (01, 02, 12)."""
    s = np.asarray(std, dtype=np.float64)
    c01, c02, c12 = corr
    return np.array([[1, c01, c02], [c01, 1, c12], [c02, c12, 1]]) * np.outer(s, s)


def spherical_corr_xy(nxp, nyp, dx, dy, range1, range2=None, angle=0.0):
    """CommonData::FindCorrXYGrid (commondata.cpp:9180-9218) with SphericalVario::corr
    (vario.cpp:35-74): correlation at wrap-around lags on the padded lattice, [nyp][nxp]."""
    r2 = range1 if range2 is None else range2
    i = np.arange(nxp)
    j = np.arange(nyp)
    ri = np.where(i < nxp // 2 + 1, i, i - nxp).astype(np.float32) * np.float32(dx)
    rj = np.where(j < nyp // 2 + 1, j, j - nyp).astype(np.float32) * np.float32(dy)
    txx = np.float32(math.cos(angle) ** 2 / range1 ** 2 + math.sin(angle) ** 2 / r2 ** 2)
    tyy = np.float32(math.sin(angle) ** 2 / range1 ** 2 + math.cos(angle) ** 2 / r2 ** 2)
    txy = np.float32(2 * math.cos(angle) * math.sin(angle) * (1.0 / range1 ** 2 - 1.0 / r2 ** 2))
    X, Y = np.meshgrid(ri, rj)
    d = np.sqrt(txx * X * X + tyy * Y * Y + txy * X * Y).astype(np.float32)
    return np.where(d >= 1.0, 0.0, 1.0 + d ** 3 / 2.0 - 1.5 * d).astype(np.float32)


def circ_corr_t(dt_ms, temporal_range_ms, nzp, low_cut_hz):
    """Exponential temporal correlation (commondata.cpp:8597-8604) mirrored onto the circle and made
    positive definite with frequencies <= low_int_cut removed (seismicparametersholder.cpp:624-656,
    805-837; main.cpp:258)."""
    n_corr = nzp // 2 + 1 if nzp % 2 == 0 else nzp // 2
    corr = np.exp(-3.0 * np.float32(dt_ms) * np.arange(n_corr + 1) / np.float32(temporal_range_ms))
    k = np.arange(nzp)
    refk = np.where(k < nzp // 2 + 1, k, nzp - k)
    c = np.where(refk < len(corr), corr[np.minimum(refk, len(corr) - 1)].astype(np.float32), 0.0)
    low_int_cut = int(math.floor(low_cut_hz * (nzp * 0.001 * dt_ms)))
    f = np.abs(np.fft.fft(c.astype(np.float64)))
    f[refk <= low_int_cut] = 0.0
    back = np.fft.ifft(f).real * nzp
    if not back[0] > 1e-5:
        raise ValueError("circular temporal correlation undefined (too few layers)")
    return (back * np.float32(1.0 / back[0])).astype(np.float32)


def lateral_spectrum(corr_xy):
    """Unnormalised 2D FFT of the lateral correlation on the half-spectrum x axis, [nyp][cnxp]:
    both the separable prior's lateral factor and the error-correlation spectrum lambda
    (fillInErrCorr puts the lateral correlation in the k = 0 plane only, fftgrid.cpp:777-828)."""
    nyp, nxp = corr_xy.shape
    return np.fft.fft2(corr_xy.astype(np.float64)).real[:, : nxp // 2 + 1].astype(np.float32)


def temporal_spectrum(circ_t):
    return np.fft.fft(circ_t.astype(np.float64)).real.astype(np.float32)


def angular_error_covariance(sigma_e, angles_deg, range_deg=10.0):
    """Sigma0e_lm = s_l s_m exp(-3 |theta_l - theta_m| / range) (ModelAVODynamic::SetupErrorCorrelation,
    modelavodynamic.cpp:623-651, with the default generalised-exponential angular correlation,
    commondata.cpp:9862-9881)."""
    s = np.asarray(sigma_e, dtype=np.float64)
    a = np.asarray(angles_deg, dtype=np.float64)
    return np.outer(s, s) * np.exp(-3.0 * np.abs(a[:, None] - a[None, :]) / range_deg)


def inversion_tables(nz, nzp, dt_ms, angles_deg, peak_hz, wavelet_scales=None, vsvp=0.5, low_cut=5.0,
                     high_cut=55.0, ormsby_hz=None):
    """Reflection matrix, wavelets and the per-k_z tables (K, e1, e2, e3, invert)."""
    nt = len(angles_deg)
    scales = wavelet_scales if wavelet_scales is not None else [1.0] * nt
    A = aki_richards(angles_deg, vsvp)
    wav = np.stack([wavelet_fft_order(nzp, dt_ms, peak_hz, s, ormsby_hz) for s in scales])
    K, e1, e2, e3, inv = build_kz_tables(nz, nzp, dt_ms, wav, A, low_cut, high_cut, 1.0)
    return dict(A=A, wavelets=wav, K=K, e1=e1, e2=e2, e3=e3, invert=inv)


def synthetic_operator(wavelets, A):
    """Per-k_z operator of CravaResult::ComputeSyntSeismic (cravaresult.cpp:2312-2377) for
    cg_forward_model: K[k, l, j] = A[l, j] (1 - e^{-2 pi i k / nzp}) conj(W_l(k)), the backward
    difference of the impedance (Wavelet1D FIRSTORDERBACKWARDDIFF, wavelet1D.cpp:462-467) and the
    product with the conjugate wavelet spectrum.  wavelets (ntheta, nzp) real, in FFT order;
    A (ntheta, 3) the reflection coefficients.  Returns (nzp, ntheta, 3) complex64."""
    w = np.asarray(wavelets, dtype=np.float64)
    a = np.asarray(A, dtype=np.float64)
    nzp = w.shape[1]
    k = np.arange(nzp)
    diff = 1.0 - np.exp(-2j * np.pi * k / nzp)
    W = np.fft.fft(w, axis=1)  # sign -1, unnormalised, every k
    K = diff[:, None, None] * np.conj(W).T[:, :, None] * a[None, :, :]
    return K.astype(np.complex64)


def synthetic_seismic(ctx, vp, vs, rho, wavelets, A, nxp, nyp, nzp):
    """CravaResult::ComputeSyntSeismic on the device: vp, vs, rho cubes (nz, ny, nx) -> one
    synthetic seismic cube (nz, ny, nx) per angle.  Each parameter gets the reference's trace
    taper over the z padding, is transformed (3D, the x / y transforms cancel for an operator
    acting along z only), combined by the operator of synthetic_operator and transformed back."""
    from . import FFTGrid, PARAMETER, DATA, forward_model
    nz, ny, nx = np.shape(vp)
    nt = len(wavelets)
    m = [FFTGrid(ctx, nx, ny, nz, nxp, nyp, nzp, PARAMETER) for _ in range(3)]
    d = [FFTGrid(ctx, nx, ny, nz, nxp, nyp, nzp, DATA) for _ in range(nt)]
    try:
        for g, cube in zip(m, (vp, vs, rho)):
            g.fill_trace_taper(cube)
            g.fftInPlace()
        forward_model(ctx, synthetic_operator(wavelets, A), m, d)
        out = []
        for g in d:
            g.invFFTInPlace()
            out.append(g.download_real()[:nz, :ny, :nx])
        return out
    finally:
        for g in m + d:
            g.close()


def _eig_general(m):
    """Eigenvalues (real parts) and right eigenvectors (columns) by LAPACK dgeev, the routine
    NRLib::ComputeEigenVectors reaches through flens::ev (libs/nrlib/flens/nrlib_flens.cpp:114-121)."""
    w, v = np.linalg.eig(np.asarray(m, dtype=np.float64))
    return np.real(w).copy(), np.real(v).copy()


def compute_g(prior_var0, post_var0, err_theta_cov):
    """AVOInversion::computeG (avoinversion.cpp:1943-2026), the host side of the local-noise
    correction (crava_amd.LocalNoise): the n_theta x 3 operator
    G = Sigma_e^1/2 Lg V^T Sigma_m^-1/2, where V, lambda are the eigen pairs of
    A = Sigma_m^-1/2 (Sigma_m - postVar0) Sigma_m^-1/2 in decreasing order (NRLib::Sort3x3) and
    Lg = diag(sqrt(lambda / (1 - lambda))), so that the posterior of G with noise Sigma_e is
    postVar0.  V's signs are LAPACK's (they matter once the scales vary between angles)."""
    sm = np.asarray(prior_var0, dtype=np.float64)
    se = np.asarray(err_theta_cov, dtype=np.float64)
    nt = se.shape[0]

    def vfv(v, f):
        return (v * f[None, :]) @ v.T

    w, v = _eig_general(sm - np.asarray(post_var0, dtype=np.float64))
    sdelta = vfv(v, np.maximum(w, 0.0))
    w, v = _eig_general(sm)
    sinv = vfv(v, np.array([1.0 / math.sqrt(x) if x > 1e-7 else 0.0 for x in w]))
    w, v = _eig_general((sinv @ sdelta) @ sinv)
    order = []
    for _ in range(3):  # Sort3x3: decreasing, the first of equal values first
        best, bi = -100000.0, -1
        for i in range(3):
            if i not in order and w[i] > best:
                best, bi = w[i], i
        order.append(bi)
    w, v = w[order], v[:, order]
    lg = np.zeros((nt, 3))
    for i in range(min(nt, 3)):
        if w[i] > 0:
            lg[i, i] = math.sqrt(w[i] / (1.0 - w[i]))
    we, ve = _eig_general(se)
    esq = vfv(ve, np.array([math.sqrt(x) if x > 0.0 else 0.0 for x in we]))
    return ((esq @ lg) @ v.T) @ sinv
