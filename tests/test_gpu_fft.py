"""GPU FFT passes vs the oracle (FFTGrid::fftInPlace / invFFTInPlace semantics)."""
import numpy as np
import pytest

import crava_amd as C
import oracle as O
from conftest import rel_l2

pytestmark = pytest.mark.gpu

# even/odd x, every radix (2,3,4,5,7,8), y/z lengths of 1 and primes
SIZES = [(8, 6, 10), (12, 10, 14), (9, 7, 5), (16, 1, 8), (84, 84, 30), (30, 42, 49), (64, 64, 128),
         (15, 21, 35), (2, 2, 2), (40, 1, 1)]


def _grid(ctx, shape, ct):
    nzp, nyp, nxp = shape
    return C.FFTGrid(ctx, nxp, nyp, nzp, nxp, nyp, nzp, ct)


@pytest.mark.parametrize("nxp,nyp,nzp", SIZES)
@pytest.mark.parametrize("ct", [C.DATA, C.COVARIANCE])
def test_forward_matches_oracle(ctx, nxp, nyp, nzp, ct):
    print("forward", nxp, nyp, nzp, ct, flush=True)
    rng = np.random.default_rng(nxp * 1000 + nyp * 10 + nzp)
    x = rng.standard_normal((nzp, nyp, nxp)).astype(np.float32)
    g = _grid(ctx, x.shape, ct)
    g.upload_real(x)
    g.fftInPlace()
    assert g.getIsTransformed()
    got = g.download_complex()
    want = O.grid_fft(x, ct)
    assert rel_l2(got, want) < 2e-6
    g.close()


@pytest.mark.parametrize("nxp,nyp,nzp", SIZES)
def test_inverse_matches_oracle_including_non_hermitian_input(ctx, nxp, nyp, nzp):
    print("inverse", nxp, nyp, nzp, flush=True)
    rng = np.random.default_rng(7 + nxp)
    cnxp = nxp // 2 + 1
    z = (rng.standard_normal((nzp, nyp, cnxp)) + 1j * rng.standard_normal((nzp, nyp, cnxp))).astype(np.complex64)
    for ct in (C.PARAMETER, C.COVARIANCE):
        g = _grid(ctx, (nzp, nyp, nxp), ct)
        g.upload_complex(z)
        g.invFFTInPlace()
        got = g.download_real()
        want = O.grid_ifft(z, nxp, ct)
        assert rel_l2(got, want) < 2e-6
        g.close()


def test_round_trip_and_parseval(ctx):
    rng = np.random.default_rng(1)
    x = rng.standard_normal((40, 36, 50)).astype(np.float32)
    g = _grid(ctx, x.shape, C.DATA)
    g.upload_real(x)
    g.fftInPlace()
    s = g.download_complex()
    # unitary: sum |x|^2 = |X_0|^2 + 2 sum_{0<kx<n/2} |X|^2 + |X_{n/2}|^2
    e = np.sum(np.abs(s[:, :, 0]) ** 2) + 2 * np.sum(np.abs(s[:, :, 1:-1]) ** 2) + np.sum(np.abs(s[:, :, -1]) ** 2)
    assert abs(e / np.sum(x.astype(np.float64) ** 2) - 1) < 1e-5
    g.invFFTInPlace()
    assert rel_l2(g.download_real(), x) < 1e-6
    g.close()


def test_accessors_and_elementwise(ctx):
    rng = np.random.default_rng(3)
    g = C.FFTGrid(ctx, 10, 6, 8, 12, 8, 10, C.PARAMETER)
    x = rng.standard_normal((10, 8, 12)).astype(np.float32)
    g.upload_real(x)
    g.multiplyByScalar(2.0)
    h = C.FFTGrid(ctx, 10, 6, 8, 12, 8, 10, C.PARAMETER)
    h.upload_real(x)
    g.add(h)
    assert rel_l2(g.download_real(), 3 * x) < 1e-7
    g.fftInPlace()
    s = g.download_complex()
    assert g.getComplexValue(3, 2, 5) == pytest.approx(complex(s[5, 2, 3]), rel=1e-6)
    # outside the unpadded box -> RMISSING; inside the padded box with extSimbox
    assert g.getComplexValue(3, 7, 5) == complex(C.RMISSING, C.RMISSING)
    assert g.getComplexValue(3, 7, 9, extSimbox=True) == pytest.approx(complex(s[9, 7, 3]), rel=1e-6)
    assert g.setComplexValue(1, 1, 1, 2 + 3j) == 0
    assert g.getComplexValue(1, 1, 1) == 2 + 3j
    assert g.setComplexValue(11, 1, 1, 1j) == 1
    with pytest.raises(C.CravaError):
        g.fftInPlace()  # already transformed
    g.close()
    h.close()


# Every compile-time plan length (160 192 256 288 320 384 512 576 640 768 1152) on every axis:
# the x pass transforms nxp/2 points, y and z their full lengths.
FIXED = [(320, 192, 8), (576, 288, 4), (8, 4, 320), (1152, 4, 4), (8, 576, 4), (8, 4, 384), (1024, 2, 4),
         (8, 512, 2), (8, 2, 640), (8, 768, 2), (8, 2, 1152), (1536, 2, 2), (2304, 2, 2), (512, 4, 160),
         (640, 256, 2), (768, 2, 192), (8, 160, 256)]


@pytest.mark.parametrize("nxp,nyp,nzp", FIXED)
def test_fixed_plan_lengths_match_oracle_and_generic_engine(ctx, nxp, nyp, nzp):
    print("fixed plans", nxp, nyp, nzp, flush=True)
    rng = np.random.default_rng(nxp + 7 * nyp + 13 * nzp)
    x = rng.standard_normal((nzp, nyp, nxp)).astype(np.float32)
    cnxp = nxp // 2 + 1
    z = (rng.standard_normal((nzp, nyp, cnxp)) + 1j * rng.standard_normal((nzp, nyp, cnxp))).astype(np.complex64)
    res = {}
    for eng in (C.FFT_AUTO, C.FFT_GENERIC):
        ctx.set_fft_engine(eng)
        g = _grid(ctx, x.shape, C.DATA)
        g.upload_real(x)
        g.fftInPlace()
        fwd = g.download_complex()
        g.upload_complex(z)
        g.invFFTInPlace()
        res[eng] = (fwd, g.download_real())
        g.close()
    ctx.set_fft_engine(C.FFT_AUTO)
    assert rel_l2(res[C.FFT_AUTO][0], O.grid_fft(x, C.DATA)) < 2e-6
    assert rel_l2(res[C.FFT_AUTO][1], O.grid_ifft(z, nxp, C.DATA)) < 2e-6
    assert rel_l2(res[C.FFT_AUTO][0], res[C.FFT_GENERIC][0]) < 2e-6
    assert rel_l2(res[C.FFT_AUTO][1], res[C.FFT_GENERIC][1]) < 2e-6


def test_out_of_place_forward_keeps_source(ctx):
    rng = np.random.default_rng(5)
    x = rng.standard_normal((16, 12, 576)).astype(np.float32)
    src = _grid(ctx, x.shape, C.DATA)
    dst = _grid(ctx, x.shape, C.PARAMETER)
    src.upload_real(x)
    dst.fft_from(src)
    assert not src.getIsTransformed() and dst.getIsTransformed()
    assert np.array_equal(src.download_real(), x)
    assert rel_l2(dst.download_complex(), O.grid_fft(x, C.DATA)) < 2e-6
    src.close()
    dst.close()


def _factorable(n):
    for p in (2, 3, 5, 7):
        while n % p == 0:
            n //= p
    return n == 1


# every length the compile-time engine instantiates (2^a 3^b 5^c 7^d up to 1152)
PLAN_LENGTHS = [n for n in range(2, 1153) if _factorable(n)]


@pytest.mark.parametrize("variant", [0, 1])
def test_every_compiled_length_matches_oracle(ctx, variant):
    """y and z passes of every length, x passes of every even nxp = 2n (n <= 576) and every odd
    nxp = n, in both tile variants: forward and inverse against the FP64 oracle."""
    ctx.set_fft_variant(variant)
    rng = np.random.default_rng(11 + variant)
    shapes = []
    for n in PLAN_LENGTHS:
        if n >= 4:
            shapes += [(2, n, 4), (n, 2, 4)]  # (nzp, nyp, nxp): y length n; z length n
        if n <= 576:
            shapes.append((2, 3, 2 * n))
        if n % 2:
            shapes.append((3, 2, n))
    worst = 0.0
    for shape in shapes:
        nzp, nyp, nxp = shape
        x = rng.standard_normal(shape).astype(np.float32)
        cnxp = nxp // 2 + 1
        z = (rng.standard_normal((nzp, nyp, cnxp)) + 1j * rng.standard_normal((nzp, nyp, cnxp))).astype(np.complex64)
        g = _grid(ctx, shape, C.DATA)
        g.upload_real(x)
        g.fftInPlace()
        e1 = rel_l2(g.download_complex(), O.grid_fft(x, C.DATA))
        g.upload_complex(z)
        g.invFFTInPlace()
        e2 = rel_l2(g.download_real(), O.grid_ifft(z, nxp, C.DATA))
        g.close()
        worst = max(worst, e1, e2)
        assert e1 < 2e-6 and e2 < 2e-6, (shape, variant, e1, e2)
    ctx.set_fft_variant(-1)
    print("compiled lengths, variant %d: %d grids, worst rel L2 %.2e" % (variant, len(shapes), worst), flush=True)


@pytest.mark.parametrize("variant", [0, 1])
def test_x_pass_tails_and_multi_tile_walks(ctx, variant):
    """Even-nxp x passes on row counts that leave a partial last tile with an odd number of rows
    (the double-buffered kernel copies an odd float2 count, its last element by a plain load),
    and on grids with more tiles than the persistent launch has workgroups (each walks several
    tiles, alternating its two buffers), forward and inverse."""
    ctx.set_fft_variant(variant)
    rng = np.random.default_rng(31 + variant)
    for nzp, nyp, nxp in ((7, 9, 576), (3, 5, 50), (1, 1, 64), (96, 96, 576), (100, 200, 288)):
        print("x pass", (nzp, nyp, nxp), "variant", variant, flush=True)
        x = rng.standard_normal((nzp, nyp, nxp)).astype(np.float32)
        cnxp = nxp // 2 + 1
        z = (rng.standard_normal((nzp, nyp, cnxp)) + 1j * rng.standard_normal((nzp, nyp, cnxp))).astype(np.complex64)
        g = _grid(ctx, x.shape, C.DATA)
        g.upload_real(x)
        g.fftInPlace()
        assert rel_l2(g.download_complex(), O.grid_fft(x, C.DATA)) < 2e-6
        g.upload_complex(z)
        g.invFFTInPlace()
        want = O.grid_ifft(z, nxp, C.DATA)
        assert rel_l2(g.download_real(), want) < 2e-6
        g.close()
    ctx.set_fft_variant(-1)


def test_autotuned_transforms_match_oracle(ctx):
    """After the autotuner has picked tile variants for a geometry, transforms of that geometry
    (plain and with the inverse's added grid) still match the oracle."""
    rng = np.random.default_rng(12)
    for nxp, nyp, nzp in ((576, 96, 40), (63, 70, 48), (160, 576, 12)):
        ctx.autotune(nxp, nyp, nzp)
        x = rng.standard_normal((nzp, nyp, nxp)).astype(np.float32)
        g = _grid(ctx, x.shape, C.DATA)
        g.upload_real(x)
        g.fftInPlace()
        assert rel_l2(g.download_complex(), O.grid_fft(x, C.DATA)) < 2e-6
        g.invFFTInPlace()
        assert rel_l2(g.download_real(), x) < 2e-6
        g.close()


def test_grids_carved_from_an_arena(ctx):
    """Grids in one arena are independent (distinct memory, same results as stand-alone grids);
    the arena refuses to go while a grid of it is alive."""
    rng = np.random.default_rng(9)
    shape = (12, 10, 18)
    xs = [rng.standard_normal(shape).astype(np.float32) for _ in range(3)]
    arena = C.Arena(ctx, 3 * C.Arena.grid_bytes(18, 10, 12))
    gs = [C.FFTGrid(ctx, 18, 10, 12, 18, 10, 12, C.DATA, arena=arena) for _ in range(3)]
    with pytest.raises(C.CravaError):
        C.FFTGrid(ctx, 18, 10, 12, 18, 10, 12, C.DATA, arena=arena)  # full
    for g, x in zip(gs, xs):
        g.upload_real(x)
    for g in gs:
        g.fftInPlace()
    for g, x in zip(gs, xs):
        assert rel_l2(g.download_complex(), O.grid_fft(x, C.DATA)) < 2e-6
    with pytest.raises(C.CravaError):
        arena.close()  # grids still alive
    for g in gs:
        g.close()
    arena.close()
