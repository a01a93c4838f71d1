"""Pins of the oracle that do not retype its own formulas: naive DFTs, closed forms, textbook
Gaussian conditioning with numpy's solver, invariants, known-answer vectors, and (where the
CRAVA tree was available at build time) the reference's own compiled code."""
import math

import numpy as np
import pytest

import oracle as O
import problem as P
from conftest import rel_l2

needs_ref = pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built (no reference tree)")


# ------------------------------------------------------------------------------ padding
def _smooth7_brute(n):
    m = max(n, 1)
    while True:
        k = m
        for p in (2, 3, 5, 7):
            while k % p == 0:
                k //= p
        if k == 1:
            return m
        m += 1


def test_factorable_is_smallest_7_smooth():
    for n in list(range(1, 600)) + [1000, 1153, 2049, 4097]:
        assert O.lib().orc_factorable(n) == _smooth7_brute(n), n


def test_padding_of_the_plain_case():
    # 64 x 64 x 128, dx 25 m, dt 4 ms, 1000 m range: half a range laterally, a 200 ms wavelet in z
    assert O.lib().orc_padding_size(64, 500.0 / (64 * 25.0)) == 84
    assert O.lib().orc_padding_size(128, 200.0 / (128 * 4.0)) == 180
    assert O.lib().orc_padding_size(512, 0.125) == 576
    assert O.lib().orc_padding_size(256, 0.25) == 320
    assert O.lib().orc_padding_size(1024, 0.125) == 1152
    assert O.lib().orc_padding_size(320, 0.2) == 384


# ------------------------------------------------------------------------------ FFT
@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 6, 7, 8, 9, 12, 14, 16, 25, 27, 32, 35, 49, 60, 84, 96, 105])
def test_cfft_matches_naive_dft(n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(2 * n)
    for sign in (-1, 1):
        want = np.zeros(2 * n)
        O.lib().orc_cdft_naive(n, sign, O._ptr(x, O.ctypes.c_double), O._ptr(want, O.ctypes.c_double))
        got = x.copy()
        O.lib().orc_cfft(n, sign, O._ptr(got, O.ctypes.c_double), 1, 1, n)
        assert np.max(np.abs(got - want)) < 1e-11 * max(1, n)


@pytest.mark.parametrize("shape", [(6, 4, 10), (5, 7, 9), (12, 9, 14), (1, 1, 8), (3, 1, 1), (7, 5, 1)])
def test_grid_fft_conventions(shape):
    rng = np.random.default_rng(sum(shape))
    x = rng.standard_normal(shape)
    nzp, nyp, nxp = shape
    N = nxp * nyp * nzp
    s = O.grid_fft(x, O.COVARIANCE)
    # unnormalised, sign -1, x the half axis: numpy's fftn restricted to kx <= nxp/2
    assert np.max(np.abs(s - np.fft.fftn(x)[:, :, : nxp // 2 + 1])) < 1e-10 * N
    # DATA/PARAMETER: unitary (1/sqrt N computed in float)
    u = O.grid_fft(x, O.DATA)
    assert rel_l2(u, s / math.sqrt(N)) < 1e-7
    # round trips: COVARIANCE x 1/N on the inverse, unitary pair for DATA
    assert np.max(np.abs(O.grid_ifft(s, nxp, O.COVARIANCE) - x)) < 1e-6
    assert np.max(np.abs(O.grid_ifft(u, nxp, O.DATA) - x)) < 1e-6


def test_c2r_ignores_imaginary_part_of_dc_and_nyquist():
    rng = np.random.default_rng(3)
    nzp, nyp, nxp = 4, 3, 8
    z = rng.standard_normal((nzp, nyp, 5)) + 1j * rng.standard_normal((nzp, nyp, 5))
    a = O.grid_ifft(z, nxp, O.COVARIANCE)
    z2 = z.copy()
    # the x pass is last and reads only Re of its DC / Nyquist bins; an impulse in Im of the
    # (kx=0, ky=0, kz=0) and (kx=4, ky=0, kz=0) bins changes nothing after the y/z passes only if
    # it is in those bins of the *x rows*; test through a single row grid instead
    x1 = O.grid_ifft(z[:1, :1, :], nxp, O.COVARIANCE)
    z3 = z[:1, :1, :].copy()
    z3[0, 0, 0] += 5j
    z3[0, 0, 4] -= 3j
    assert np.max(np.abs(O.grid_ifft(z3, nxp, O.COVARIANCE) - x1)) < 1e-12
    assert a.shape == (nzp, nyp, nxp) and np.array_equal(z, z2)


def test_hermitian_symmetry_of_real_input():
    rng = np.random.default_rng(11)
    x = rng.standard_normal((6, 5, 8))
    s = O.grid_fft(x, O.COVARIANCE)
    # X[0, ky, kz] = conj X[0, -ky, -kz]
    for ky in range(5):
        for kz in range(6):
            assert abs(s[kz, ky, 0] - np.conj(s[(-kz) % 6, (-ky) % 5, 0])) < 1e-10


@needs_ref
@pytest.mark.parametrize("shape", [(10, 6, 8), (5, 7, 9), (30, 42, 49), (36, 84, 84), (16, 15, 21)])
def test_fft_matches_reference_fftw(shape):
    rng = np.random.default_rng(len(shape) + shape[0])
    x = rng.standard_normal(shape).astype(np.float32)
    nxp = shape[2]
    for ct in (O.DATA, O.COVARIANCE):
        assert rel_l2(O.grid_fft(x, ct), O.ref_grid_fft(x, ct)) < 1e-6
    z = (rng.standard_normal(O.grid_fft(x, O.DATA).shape) * (1 + 1j)).astype(np.complex64)
    assert rel_l2(O.grid_ifft(z, nxp, O.PARAMETER), O.ref_grid_ifft(z, nxp, O.PARAMETER)) < 1e-6


# ------------------------------------------------------------------------------ posterior
def _sigma_e(p, k, lam):
    """The error covariance of one wavenumber (avoinversion.cpp:1236-1272)."""
    nt = p["ntheta"]
    e1, e2, e3 = p["e1"][k].astype(complex), p["e2"][k].astype(complex), p["e3"][k].astype(complex)
    S = np.zeros((nt, nt), complex)
    for l in range(nt):
        for m in range(nt):
            S[l, m] = 0.5 * (1 - p["wnc"]) * p["sigma0e"][l, m] * lam * (e1[l] * np.conj(e1[m]) + e2[l] * np.conj(e2[m]))
        S[l, l] = S[l, l].real + p["wnc"] * p["sigma0e"][l, l] * e3[l].real ** 2
    return S


def test_fp64_posterior_is_gaussian_conditioning():
    # Sigma' = P - P K^H (K P K^H + Se)^-1 K P, m' = m + P K^H (K P K^H + Se)^-1 (d - K m), by
    # numpy's dense solver, on a sample of inverted wavenumbers
    p = P.make_problem(12, 10, 16, seed=21, angles_deg=(0.0, 15.0, 30.0, 45.0))
    mean, data = P.forward_spectra(p, np.float64)
    om, od, oc, flags, nf = O.posterior(p, mean, data, None, "f64")
    assert nf == 0
    S0 = p["prior_var0"]
    rng = np.random.default_rng(0)
    ks = np.where(p["invert"])[0]
    checked = 0
    for _ in range(60):
        k = int(rng.choice(ks))
        j = int(rng.integers(p["nyp"]))
        i = int(rng.integers(p["cnxp"]))
        r = abs(float(p["rxy"][j, i]) * float(p["rt"][k]))
        Pm = r * S0
        K = p["K"][k].astype(complex)
        Se = _sigma_e(p, k, abs(float(p["lam2d"][j, i])))
        M = K @ Pm @ K.conj().T + Se
        gain = Pm @ K.conj().T @ np.linalg.inv(M)
        post = Pm - gain @ K @ Pm
        mvec = mean[:, k, j, i]
        dvec = data[:, k, j, i]
        mpost = mvec + gain @ (dvec - K @ mvec)
        got_cov = oc[[0, 1, 2, 3, 4, 5], k, j, i]
        want_cov = post[[0, 1, 2, 0, 0, 1], [0, 1, 2, 1, 2, 2]]
        # double rounding amplified by cond(M) (up to ~1e6 here); an algorithmic error is O(1)
        assert np.max(np.abs(got_cov - want_cov)) < 1e-7 * np.max(np.abs(Pm))
        assert np.max(np.abs(om[:, k, j, i] - mpost)) < 1e-7 * (1 + np.max(np.abs(mpost)))
        assert np.max(np.abs(od[:, k, j, i] - (dvec - K @ mpost))) < 1e-7 * (1 + np.max(np.abs(dvec)))
        checked += 1
    assert checked == 60


def test_posterior_invariants():
    p = P.make_problem(12, 10, 16, seed=22)
    mean, data = P.forward_spectra(p, np.float64)
    om, od, oc, flags, nf = O.posterior(p, mean, data, None, "f64")
    # pass-through planes keep the prior mean, the data, and the prior covariance
    out = np.where(p["invert"] == 0)[0]
    assert np.array_equal(om[:, out], mean[:, out]) and np.array_equal(od[:, out], data[:, out])
    r = np.abs(p["rxy"][None].astype(np.float64) * p["rt"][:, None, None].astype(np.float64))
    assert np.allclose(oc[0, out].real, (r * p["prior_var0"][0, 0])[out], rtol=1e-12, atol=0)
    # Sigma' is Hermitian PSD and Sigma_m - Sigma' is PSD
    S0 = p["prior_var0"]
    inv = np.where(p["invert"])[0][:4]
    for k in inv:
        for j in range(0, p["nyp"], 3):
            for i in range(0, p["cnxp"], 2):
                c = oc[:, k, j, i]
                Sp = np.array([[c[0], c[3], c[4]], [np.conj(c[3]), c[1], c[5]], [np.conj(c[4]), np.conj(c[5]), c[2]]])
                ev = np.linalg.eigvalsh((Sp + Sp.conj().T) / 2)
                rr = abs(float(p["rxy"][j, i]) * float(p["rt"][k]))
                ev2 = np.linalg.eigvalsh(rr * S0 - (Sp + Sp.conj().T) / 2)
                assert ev.min() > -1e-12 * rr and ev2.min() > -1e-12 * rr


def test_huge_noise_returns_the_prior():
    p = P.make_problem(10, 8, 12, seed=23)
    p = dict(p, sigma0e=p["sigma0e"] * 1e12)
    mean, data = P.forward_spectra(p, np.float64)
    om, od, oc, _, _ = O.posterior(p, mean, data, None, "f64")
    assert rel_l2(om, mean) < 1e-8


@needs_ref
def test_fp32_posterior_bit_identical_to_reference_lib_matr_in_both_prior_forms():
    p = P.make_problem(16, 12, 20, seed=24, angles_deg=(0.0, 10.0, 20.0, 30.0, 40.0, 50.0))
    mean, data = P.forward_spectra(p, np.float32)
    a = O.posterior(p, mean, data, None, "f32")
    b = O.ref_posterior(p, mean, data, None)
    for x, y in zip(a[:4], b[:4]):
        assert np.array_equal(x, y)
    covs = np.stack([O.grid_fft(g, O.COVARIANCE, np.float32) for g in P.prior_cov_grids(p)])
    q = dict(p, prior_mode=0, lam_grid=O.grid_fft(P.err_corr_grid(p), O.COVARIANCE, np.float32))
    a = O.posterior(q, mean, data, covs, "f32")
    b = O.ref_posterior(q, mean, data, covs)
    for x, y in zip(a[:4], b[:4]):
        assert np.array_equal(x, y)


@needs_ref
def test_cholesky_failure_paths_match_reference():
    p = P.make_problem(10, 8, 12, seed=25)
    p = dict(p, sigma0e=-np.abs(p["sigma0e"]) - 1.0)
    mean, data = P.forward_spectra(p, np.float32)
    a = O.posterior(p, mean, data, None, "f32")
    b = O.ref_posterior(p, mean, data, None)
    assert a[4] == b[4] > 0
    for x, y in zip(a[:4], b[:4]):
        assert np.array_equal(x, y)


# ------------------------------------------------------------------------------ RNG / noise
def test_lcg_closed_form():
    st = np.array([23665], dtype=np.uint32)
    s = 23665
    for _ in range(50):
        s = (69069 * s + 1) % 2 ** 32
        assert O.lib().orc_unif01(O._ptr(st, O.ctypes.c_uint32)) == s / 2.0 ** 32


def test_marsaglia_bray_moments():
    st = np.array([12345], dtype=np.uint32)
    x = np.zeros(400000)
    O.lib().orc_rnorm01_n(O._ptr(st, O.ctypes.c_uint32), O._ptr(x, O.ctypes.c_double), len(x))
    assert abs(x.mean()) < 0.01 and abs(x.std() - 1) < 0.01
    assert abs(np.mean(x ** 4) - 3) < 0.05  # kurtosis of N(0,1)
    assert np.mean(np.abs(x) > 3) == pytest.approx(0.0027, abs=0.0005)


@needs_ref
def test_rng_and_noise_fill_bit_identical_to_reference():
    st = np.array([987654], dtype=np.uint32)
    O.ref().ref_rng_seed(987654)
    for shape in [(12, 10, 16), (9, 7, 10), (10, 6, 8)]:
        a = O.fill_complex_noise(*shape, st)
        b = np.zeros_like(a)
        O.ref().ref_fill_complex_noise(*shape, O._ptr(b.view(np.float32), O.ctypes.c_float))
        assert np.array_equal(a, b)


@pytest.mark.parametrize("shape", [(12, 10, 16), (10, 6, 8), (9, 7, 6)])
def test_noise_layout(shape):
    nxp, nyp, nzp = shape
    cnxp = nxp // 2 + 1
    for z in (O.fill_complex_noise(nxp, nyp, nzp, np.array([5], dtype=np.uint32)),
              O.fill_complex_noise_cb(nxp, nyp, nzp, 17, 3)):
        flat = z.ravel()
        for idx in range(flat.size):
            xs = idx % cnxp
            if xs == 0 or (xs == cnxp - 1 and idx % 2 == 1):
                jk = idx // cnxp
                j, k = jk % nyp, jk // nyp
                jc, kc = (-j) % nyp, (-k) % nzp
                pjk = jc + kc * nyp
                partner = pjk * cnxp + xs
                if pjk == jk:  # self-conjugate: real
                    assert flat[idx].imag == 0
                elif pjk < jk:  # copied from the earlier partner (fftgrid.cpp:1032-1037)
                    assert flat[idx] == np.conj(flat[partner])
                if xs == 0:  # on the x = 0 column both cells are special: the pair is conjugate
                    assert flat[idx] == np.conj(flat[partner]) or pjk == jk


def test_philox_known_answers():
    # Random123 kat_vectors, philox4x32 10 rounds
    kat = [([0, 0, 0, 0], [0, 0], [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]),
           ([0xffffffff] * 4, [0xffffffff] * 2, [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]),
           ([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0],
            [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1])]
    for ctr, key, want in kat:
        c = np.array(ctr, dtype=np.uint32)
        k = np.array(key, dtype=np.uint32)
        o = np.zeros(4, dtype=np.uint32)
        O.lib().orc_philox4x32_10(O._ptr(c, O.ctypes.c_uint32), O._ptr(k, O.ctypes.c_uint32),
                                  O._ptr(o, O.ctypes.c_uint32))
        assert list(o) == want


def test_counter_noise_stream_pairs():
    """Stream s at cell idx: Box-Muller on words 0-1 (even s) or 2-3 (odd s) of the Philox block
    of counter (idx, s >> 1, 0) -- recomputed here from the raw block (itself pinned by the known
    answers above); the two streams of a pair are uncorrelated."""
    seed = 0x123456789AB
    for s in (6, 7, 12):
        for idx in (0, 5, (1 << 33) + 3):
            c = np.array([idx & 0xFFFFFFFF, idx >> 32, s >> 1, 0], dtype=np.uint32)
            k = np.array([seed & 0xFFFFFFFF, seed >> 32], dtype=np.uint32)
            o = np.zeros(4, dtype=np.uint32)
            O.lib().orc_philox4x32_10(O._ptr(c, O.ctypes.c_uint32), O._ptr(k, O.ctypes.c_uint32),
                                      O._ptr(o, O.ctypes.c_uint32))
            w0, w1 = (int(o[2]), int(o[3])) if s & 1 else (int(o[0]), int(o[1]))
            u1 = ((w0 >> 8) + 1) / 16777216.0
            u2 = (w1 >> 8) / 16777216.0
            r = math.sqrt(-2.0 * math.log(u1))
            z0, z1 = O.ctypes.c_float(), O.ctypes.c_float()
            O.lib().orc_cb_normal2(seed, s, idx, O.ctypes.byref(z0), O.ctypes.byref(z1))
            assert abs(z0.value - r * math.cos(2 * math.pi * u2)) < 1e-6 * max(1.0, r)
            assert abs(z1.value - r * math.sin(2 * math.pi * u2)) < 1e-6 * max(1.0, r)
    a = O.fill_complex_noise_cb(32, 32, 32, 77, 8).ravel()
    b = O.fill_complex_noise_cb(32, 32, 32, 77, 9).ravel()
    assert abs(np.mean(a * np.conj(b))) < 0.02
    assert not np.array_equal(a, b)


def test_counter_noise_moments():
    z = O.fill_complex_noise_cb(64, 64, 64, 2024, 0)
    v = z[:, 1:-1, 1:-1].ravel()
    assert abs(np.mean(np.abs(v) ** 2) - 1) < 0.01
    assert abs(np.mean(v.real * v.imag)) < 0.01


# ------------------------------------------------------------------------------ colouring
def test_sample_covariance_of_coloured_noise_is_the_posterior_covariance():
    p = P.make_problem(8, 6, 10, seed=26)
    mean, data = P.forward_spectra(p, np.float64)
    _, _, oc, _, _ = O.posterior(p, mean, data, None, "f64")
    n = 4000
    acc = np.zeros((3, 3) + oc.shape[1:], complex)
    for r in range(n):
        noise = np.stack([O.fill_complex_noise_cb(p["nxp"], p["nyp"], p["nzp"], 99, 3 * r + g) for g in range(3)])
        z, _ = O.colour(oc, noise, "f64")
        acc += np.einsum("a...,b...->ab...", z, np.conj(z))
    est = acc / n
    k, j, i = int(np.where(p["invert"])[0][1]), 2, 2
    want = oc[:, k, j, i]
    got = est[:, :, k, j, i]
    scale = abs(want[0])
    assert abs(got[0, 0] - want[0]) < 0.1 * scale
    assert abs(got[1, 1] - want[1]) < 0.1 * abs(want[1])
    assert abs(got[0, 1] - want[3]) < 0.1 * np.sqrt(abs(want[0] * want[1]))


@needs_ref
def test_colouring_bit_identical_to_reference():
    p = P.make_problem(10, 8, 12, seed=27)
    mean, data = P.forward_spectra(p, np.float32)
    _, _, oc, _, _ = O.posterior(p, mean, data, None, "f32")
    noise = np.stack([O.fill_complex_noise_cb(p["nxp"], p["nyp"], p["nzp"], 5, g) for g in range(3)])
    a = O.colour(oc, noise, "f32")
    b = O.ref_colour(oc, noise)
    assert a[1] == b[1] and np.array_equal(a[0], b[0])


# ------------------------------------------------------------------------------ setup
def test_aki_richards_hand_values():
    want = {0.0: [0.5, 0.0, 0.5], 15.0: [0.5359, -0.0670, 0.4665], 30.0: [0.6667, -0.25, 0.375]}
    for a, w in want.items():
        row = np.zeros(3)
        O.lib().orc_aki_richards(math.radians(a), 0.5, O._ptr(row, O.ctypes.c_double))
        assert np.allclose(row, w, atol=5e-5)


def test_ricker_and_spherical_closed_forms():
    assert O.lib().orc_ricker(0.0, 25.0) == 1.0
    # zero crossing at t = 1 / (pi f sqrt 2) s
    t0 = 1000.0 / (math.pi * 25.0 * math.sqrt(2.0))
    assert abs(O.lib().orc_ricker(t0, 25.0)) < 1e-12
    assert O.lib().orc_spherical_corr(0.0, 0.0, 1000.0, 1000.0, 0.0) == 1.0
    assert O.lib().orc_spherical_corr(1000.0, 0.0, 1000.0, 1000.0, 0.0) == 0.0
    assert O.lib().orc_spherical_corr(500.0, 0.0, 1000.0, 1000.0, 0.0) == pytest.approx(1 + 0.125 / 2 - 0.75)


def test_circular_temporal_correlation():
    nzp, dt = 180, 4.0
    ct = np.exp(-3.0 * dt * np.arange(nzp // 2 + 2) / 40.0)
    c = O.circ_corr_t(ct, int(math.floor(5.0 * nzp * 0.001 * dt)), nzp)
    assert c[0] == pytest.approx(1.0, abs=1e-6)
    assert np.allclose(c[1:], c[1:][::-1], atol=1e-6)  # symmetric on the circle
    f = np.fft.fft(c.astype(np.float64))
    assert np.all(f.real > -1e-4) and np.max(np.abs(f[:4])) < 1e-4  # PSD, low frequencies removed


def test_kz_tables_closed_forms():
    nz, nzp, dt = 100, 120, 4.0
    k = np.arange(nzp)
    t = np.where(k < nzp // 2 + 1, k, k - nzp) * dt
    w = np.array([[O.lib().orc_ricker(float(x), 25.0) for x in t]], dtype=np.float32)
    A = np.array([[0.5, -0.1, 0.4]], dtype=np.float32)
    K, e1, e2, e3, inv = O.kz_tables(nz, nzp, dt, w, A)
    W = np.fft.fft(w[0].astype(np.float64))
    D = np.exp(2j * np.pi * k / nzp) - 1.0  # forward difference operator
    assert np.allclose(K[:, 0, 0], D * np.conj(W) * 0.5, atol=1e-5)
    assert np.allclose(e1[:, 0], np.conj(W) / np.linalg.norm(w[0]), atol=1e-5)
    f = 1000.0 / (dt * nzp) * np.minimum(k, nzp - k)
    assert np.array_equal(inv.astype(bool), (f > 5.0) & (f < 55.0))
    assert np.all(e3 == 1)
