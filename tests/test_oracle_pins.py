"""Independent pins of the oracle's model conventions (CPU).

* e2, the errorSmooth3 multiplier (avoinversion.cpp:865-872, wavelet1D.cpp:441-500): the central,
  backward and forward differences of the wavelet are circulant operators, so their product has
  the closed-form spectrum  S(k) = W(k) (2 cos w - 2) i sin w,  w = 2 pi k / nzp, and the
  time-domain norm follows from Parseval.
* The wavelet spectra against the reference's own FFTW 2.1.2 (1D rfftw, the transform
  Wavelet::fft1DInPlace uses, wavelet.cpp:255-270).
* The whole per-wavenumber algebra (K, Sigma_m, Sigma_e assembly, the update, FFT scaling and the
  Hermitian half spectrum) against Gaussian conditioning written as dense real-space matrices on a
  tiny grid (1_theory.tex:293-300): the forward model is built from its definition -- Aki-Richards
  mix, forward difference, correlation with the wavelet along z -- and the error from its
  generating model, not from the per-wavenumber formulas.
"""
import math

import numpy as np
import pytest

import oracle as O

needs_ref = pytest.mark.skipif(not O.ref_available(), reason="reference-built oracle/_ref absent")


def _wavelets(nzp, nt, seed):
    """Asymmetric wavelets in FFT order (so that a missing or misplaced conjugate shows)."""
    rng = np.random.default_rng(seed)
    k = np.arange(nzp)
    t = np.where(k < nzp // 2 + 1, k, k - nzp).astype(np.float64)
    out = []
    for l in range(nt):
        shift = rng.uniform(-1.5, 1.5)
        width = rng.uniform(2.0, 4.0)
        w = (1 - 2 * ((t - shift) / width) ** 2) * np.exp(-((t - shift) / width) ** 2) + 0.1 * rng.standard_normal(nzp)
        out.append(w)
    return np.array(out, dtype=np.float32)


@pytest.mark.parametrize("nzp", [24, 45, 64, 180])
def test_e2_closed_form(nzp):
    nt = 3
    wav = _wavelets(nzp, nt, nzp)
    A = np.array([[0.5, -0.1, 0.4], [0.55, -0.2, 0.35], [0.6, -0.3, 0.3]], dtype=np.float32)
    _, e1, e2, _, _ = O.kz_tables(nzp - 4, nzp, 4.0, wav, A, -1.0, 1e9)
    w = 2 * np.pi * np.arange(nzp) / nzp
    for l in range(nt):
        W = np.fft.fft(wav[l].astype(np.float64))  # sign -1, unnormalised (FFTW's convention)
        S = W * (2 * np.cos(w) - 2) * 1j * np.sin(w)
        s_norm = math.sqrt(np.sum(np.abs(S) ** 2) / nzp)  # Parseval: time-domain L2 norm
        want = np.conj(S) / s_norm
        assert np.max(np.abs(e2[:, l] - want)) < 2e-5 * np.max(np.abs(want))
        # e1 = conj(W) / |w|
        want1 = np.conj(W) / np.linalg.norm(wav[l].astype(np.float64))
        assert np.max(np.abs(e1[:, l] - want1)) < 2e-6 * np.max(np.abs(want1))


@needs_ref
@pytest.mark.parametrize("nzp", [32, 45, 180, 320])
def test_wavelet_spectra_match_reference_fftw(nzp):
    """conj(e1) |w| is the wavelet's spectrum; FFTW 2.1.2 (the reference's own, from _ref)
    computes it from the same samples."""
    wav = _wavelets(nzp, 2, 7 + nzp)
    A = np.ones((2, 3), dtype=np.float32)
    K, e1, _, _, _ = O.kz_tables(nzp - 4, nzp, 4.0, wav, A, -1.0, 1e9)
    for l in range(2):
        buf = np.zeros(2 * (nzp // 2 + 1), dtype=np.float32)
        buf[:nzp] = wav[l]
        O.ref().ref_fft1d_r2c(nzp, O._ptr(buf, O.ctypes.c_float))
        half = buf.view(np.complex64)
        spec = np.concatenate([half, np.conj(half[1:(nzp + 1) // 2][::-1])])  # Hermitian completion
        norm = float(np.sqrt(np.sum(wav[l].astype(np.float64) ** 2)))
        got = np.conj(e1[:, l]).astype(np.complex128) * norm
        assert np.max(np.abs(got - spec)) < 3e-6 * np.max(np.abs(spec))
        # K = D conj(W) A with D the forward difference e^{+i w} - 1
        D = np.exp(2j * np.pi * np.arange(nzp) / nzp) - 1
        assert np.max(np.abs(K[:, l, 0] - D * np.conj(spec))) < 3e-6 * np.max(np.abs(D * spec))


# ------------------------------------------------------------------------------ dense brute force
def _circ_z(kernel, nzp):
    """Matrix of x -> y, y[t] = sum_s kernel[s] x[(t + s) mod nzp] (cyclic correlation along z)."""
    M = np.zeros((nzp, nzp))
    for t in range(nzp):
        for s in range(nzp):
            M[t, (t + s) % nzp] += kernel[s]
    return M


def _forward_diff(nzp):
    """y[t] = x[t + 1] - x[t], cyclic (Wavelet1D FIRSTORDERFORWARDDIFF as an operator)."""
    M = -np.eye(nzp)
    for t in range(nzp):
        M[t, (t + 1) % nzp] += 1.0
    return M


def _error_smooth3(w):
    """central(backward(forward(w))), the sequence definitions of wavelet1D.cpp:441-500."""
    f = np.roll(w, -1) - w
    b = f - np.roll(f, 1)
    return 0.5 * (np.roll(b, -1) - np.roll(b, 1))


def _positive_corr(shape, length, rng):
    """A real circular correlation whose spectrum is non-negative (the autocorrelation of a smooth
    random field, normalised to 1 at lag 0): the reference's |Re(.)| of the spectrum is then the
    spectrum itself."""
    x = rng.standard_normal(shape)
    grids = np.meshgrid(*[np.minimum(np.arange(n), n - np.arange(n)) for n in shape], indexing="ij")
    r2 = sum(g.astype(np.float64) ** 2 for g in grids)
    x = np.real(np.fft.ifftn(np.fft.fftn(x) * np.fft.fftn(np.exp(-r2 / length ** 2))))
    c = np.real(np.fft.ifftn(np.abs(np.fft.fftn(x)) ** 2))
    return c / c.flat[0]


@pytest.mark.parametrize("nt,seed", [(2, 1), (3, 2)])
def test_posterior_equals_dense_real_space_conditioning(nt, seed):
    rng = np.random.default_rng(seed)
    nxp, nyp, nzp = 4, 4, 8
    cnxp = nxp // 2 + 1
    N = nxp * nyp * nzp
    wav = _wavelets(nzp, nt, 100 + seed)
    angles = np.linspace(0.0, 30.0, nt)
    A = np.array([[0.5 + 0.2 * math.sin(math.radians(a)), -0.8 * math.sin(math.radians(a)) ** 2,
                   0.5 - 0.3 * math.sin(math.radians(a))] for a in angles], dtype=np.float32)
    K, e1, e2, e3, inv = O.kz_tables(nzp, nzp, 4.0, wav, A, -1.0, 1e9)  # every k_z plane inverted
    assert inv.all()
    S0 = np.array([[0.010, 0.0105, 0.002], [0.0105, 0.0225, 0.00225], [0.002, 0.00225, 0.0025]])
    cz = _positive_corr((nzp,), 1.5, rng)                     # prior: temporal correlation
    cxy = _positive_corr((nyp, nxp), 1.2, rng)                # prior: lateral correlation
    ce = _positive_corr((nyp, nxp), 1.0, rng)                 # lateral error correlation (k = 0 plane)
    sig = rng.uniform(0.5, 1.5, nt)
    s0e = np.array([[sig[l] * sig[m] * math.exp(-3 * abs(angles[l] - angles[m]) / 10.0) for m in range(nt)]
                    for l in range(nt)])
    wnc = 0.1
    # ---- the oracle, per wavenumber: separable prior r(k) = |rt(kz) rxy(kx, ky)| and the 2D error
    # spectrum, both unnormalised (COVARIANCE) transforms of the real correlations
    rt = np.fft.fft(cz).real
    rxy = np.fft.fft2(cxy).real[:, :cnxp]
    lam2d = np.fft.fft2(ce).real[:, :cnxp]
    p = dict(nxp=nxp, nyp=nyp, nzp=nzp, ntheta=nt, K=K, e1=e1, e2=e2, e3=e3, invert=inv, sigma0e=s0e, wnc=wnc,
             prior_var0=S0, prior_mode=1, rxy=rxy.astype(np.float32), rt=rt.astype(np.float32),
             lam2d=lam2d.astype(np.float32))
    mu = np.stack([np.full((nzp, nyp, nxp), v) for v in (8.0, 7.3, 0.8)]) + 0.05 * rng.standard_normal((3, nzp, nyp, nxp))
    d = rng.standard_normal((nt, nzp, nyp, nxp))
    ms = np.stack([O.grid_fft(x, O.PARAMETER) for x in mu])
    ds = np.stack([O.grid_fft(x, O.DATA) for x in d])
    om, _, oc, _, nf = O.posterior(p, ms, ds, None, "f64")
    assert nf == 0
    got_mean = np.stack([O.grid_ifft(x, nxp, O.PARAMETER) for x in om])
    # FFTGrid's unitary scale factor is a float (fftgrid.cpp:1596-1598), so the transform pair is the
    # identity only to ~1e-7: the oracle's update is its posterior mean minus its round trip of mu
    mu_rt = np.stack([O.grid_ifft(x, nxp, O.PARAMETER) for x in ms])
    got_cov = np.stack([O.grid_ifft(x, nxp, O.COVARIANCE) for x in oc])  # covariance functions
    # ---- dense real space, cells ordered (z, y, x).  The prior and error correlations are
    # stationary: Cov(x, x') = c(x' - x) on the torus (rt, rxy, lam2d above are rounded to float, so
    # the real-space correlations are taken back from those rounded spectra)
    cz = np.real(np.fft.ifft(p["rt"].astype(np.float64)))
    cxy = np.fft.irfft2(p["rxy"].astype(np.float64), s=(nyp, nxp))
    ce = np.fft.irfft2(p["lam2d"].astype(np.float64), s=(nyp, nxp))
    Z, Y, X = np.unravel_index(np.arange(N), (nzp, nyp, nxp))
    dz, dy, dx = ((Z[None, :] - Z[:, None]) % nzp, (Y[None, :] - Y[:, None]) % nyp, (X[None, :] - X[:, None]) % nxp)
    Cm = cz[dz] * cxy[dy, dx]
    Ce = (dz == 0) * ce[dy, dx]                               # white in z: only the k = 0 plane is filled
    Sm = np.kron(S0, Cm)
    Iy = np.eye(nyp * nxp)

    def along_z(Mz):  # an operator on z applied to every (y, x) column
        return np.kron(Mz, Iy)

    Dz = along_z(_forward_diff(nzp))
    G = np.zeros((nt * N, 3 * N))
    for l in range(nt):
        Wl = along_z(_circ_z(wav[l].astype(np.float64), nzp))  # correlation with the wavelet
        for j in range(3):
            G[l * N:(l + 1) * N, j * N:(j + 1) * N] = float(A[l, j]) * (Wl @ Dz)
    # error model: e_l = (w_l / |w_l|) * eps_l + (s_l / |s_l|) * eps'_l (correlations along z), with
    # Cov(eps_l, eps_m) = Cov(eps'_l, eps'_m) = 0.5 (1 - wnc) s0e_lm Ce, eps and eps' independent,
    # plus white noise of variance wnc s0e_ll
    Se = np.zeros((nt * N, nt * N))
    U = [along_z(_circ_z(w / np.linalg.norm(w), nzp)) for w in wav.astype(np.float64)]
    V = [along_z(_circ_z(s / np.linalg.norm(s), nzp)) for s in (_error_smooth3(w) for w in wav.astype(np.float64))]
    for l in range(nt):
        for m in range(nt):
            blk = 0.5 * (1 - wnc) * s0e[l, m] * (U[l] @ Ce @ U[m].T + V[l] @ Ce @ V[m].T)
            if l == m:
                blk = blk + wnc * s0e[l, l] * np.eye(N)
            Se[l * N:(l + 1) * N, m * N:(m + 1) * N] = blk
    Msys = G @ Sm @ G.T + Se
    X = np.linalg.solve(Msys, G @ Sm)
    m_post = mu.reshape(-1) + X.T @ (d.reshape(-1) - G @ mu.reshape(-1))
    S_post = Sm - (G @ Sm).T @ X
    # the oracle's per-plane tables (K, e1, e2) are float32, as the reference's are: agreement to
    # their rounding, relative L2 of the update and of the covariance functions
    upd = m_post - mu.reshape(-1)
    err_upd = np.linalg.norm(got_mean.reshape(-1) - mu_rt.reshape(-1) - upd) / np.linalg.norm(upd)
    print("dense brute force: update rel L2 %.2e" % err_upd)
    assert err_upd < 1e-6
    # the six covariance functions.  Spectral entry ab is E[m_a(k) conj(m_b(k))], whose inverse
    # transform at lag u is Cov(m_a(x + u), m_b(x)): column x = 0 of block (a, b)
    for g, (a, b) in enumerate([(0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2)]):
        col = S_post[a * N:(a + 1) * N, b * N].reshape(nzp, nyp, nxp)
        err = np.linalg.norm(got_cov[g] - col) / np.linalg.norm(col)
        print("dense brute force: covariance %d rel L2 %.2e" % (g, err))
        assert err < 1e-6, g


# ------------------------------------------------------------------------------ data variance
def test_data_variance_definition():
    """ComputeDataVariance: the mean of float(x * x) over the non-zero box samples (the padding and
    the zeros left out), against an exactly rounded sum; all-zero data gives 0."""
    rng = np.random.default_rng(11)
    nx, ny, nz, nxp, nyp, nzp = 13, 9, 7, 16, 12, 10
    rv = np.zeros((nzp, nyp, 2 * (nxp // 2 + 1)), dtype=np.float32)
    rv[:] = rng.standard_normal(rv.shape).astype(np.float32) * 1e3  # padding must not count
    box = rv[:nz, :ny, :nx]
    box[rng.random(box.shape) < 0.2] = 0.0
    v, n = O.data_variance(rv, nx, ny, nz, nxp)
    sq = (box[box != 0] * box[box != 0]).astype(np.float32)  # float products, as the reference
    assert n == np.count_nonzero(box)
    assert abs(v - math.fsum(sq.astype(np.float64)) / n) <= 2e-7 * v
    assert O.data_variance(np.zeros_like(rv), nx, ny, nz, nxp) == (0.0, 0)
