"""ctypes front end of the oracle (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this module.
It wraps two shared libraries built by oracle/Makefile:

* ``liboracle.so`` -- the CPU restatement of CRAVA's path (see oracle.h for what each
  function restates and what pins it);
* ``_ref/libcrava_ref.so`` -- the reference's own FFTW 2.1.2 / lib_matr.c / random.cpp,
  compiled from the CRAVA tree, driven by ref_harness.cpp (optional: absent where the
  reference tree was not available at build time).
"""
import ctypes
import math
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libcrava_ref.so")

DATA, PARAMETER, COVARIANCE = 1, 2, 3

_c_float_p = ctypes.POINTER(ctypes.c_float)
_c_double_p = ctypes.POINTER(ctypes.c_double)
_c_int_p = ctypes.POINTER(ctypes.c_int)
_c_u8_p = ctypes.POINTER(ctypes.c_uint8)
_c_u32_p = ctypes.POINTER(ctypes.c_uint32)


def _ptr(a, ctype):
    if a is None:
        return None
    return a.ctypes.data_as(ctypes.POINTER(ctype))


class PostSpec(ctypes.Structure):
    _fields_ = [
        ("nxp", ctypes.c_int), ("nyp", ctypes.c_int), ("nzp", ctypes.c_int), ("ntheta", ctypes.c_int),
        ("K", _c_float_p), ("e1", _c_float_p), ("e2", _c_float_p), ("e3", _c_float_p),
        ("invert", _c_int_p), ("sigma0e", _c_double_p), ("wnc", ctypes.c_float),
        ("prior_var0", _c_double_p), ("prior_mode", ctypes.c_int),
        ("rxy", _c_float_p), ("rt", _c_float_p), ("lam_grid", _c_float_p), ("lam2d", _c_float_p),
    ]


_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError("oracle library not built: run `make -C oracle` (or __graft_entry__.build())")
        L = ctypes.CDLL(ORACLE_SO)
        L.orc_factorable.restype = ctypes.c_int
        L.orc_factorable.argtypes = [ctypes.c_int]
        L.orc_padding_size.restype = ctypes.c_int
        L.orc_padding_size.argtypes = [ctypes.c_int, ctypes.c_double]
        L.orc_cdft_naive.argtypes = [ctypes.c_int, ctypes.c_int, _c_double_p, _c_double_p]
        L.orc_cfft.argtypes = [ctypes.c_int, ctypes.c_int, _c_double_p, ctypes.c_long, ctypes.c_long, ctypes.c_long]
        for f in ("orc_rfft3d_r2c", "orc_rfft3d_c2r"):
            getattr(L, f).argtypes = [ctypes.c_int] * 3 + [_c_double_p]
        for f in ("orc_grid_fft", "orc_grid_ifft"):
            getattr(L, f).argtypes = [ctypes.c_int] * 4 + [_c_double_p]
        for f in ("orc_grid_fft_f32", "orc_grid_ifft_f32"):
            getattr(L, f).argtypes = [ctypes.c_int] * 4 + [_c_float_p]
        L.orc_posterior_f32.restype = ctypes.c_long
        L.orc_posterior_f32.argtypes = [ctypes.POINTER(PostSpec), _c_float_p, _c_float_p, _c_float_p, _c_u8_p]
        L.orc_posterior_f64.restype = ctypes.c_long
        L.orc_posterior_f64.argtypes = [ctypes.POINTER(PostSpec), _c_double_p, _c_double_p, _c_double_p, _c_u8_p]
        L.orc_unif01.restype = ctypes.c_double
        L.orc_unif01.argtypes = [_c_u32_p]
        L.orc_rnorm01.restype = ctypes.c_double
        L.orc_rnorm01.argtypes = [_c_u32_p]
        L.orc_rnorm01_n.argtypes = [_c_u32_p, _c_double_p, ctypes.c_long]
        L.orc_fill_complex_noise.argtypes = [ctypes.c_int] * 3 + [_c_u32_p, _c_float_p]
        L.orc_cb_normal2.argtypes = [ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint64, _c_float_p, _c_float_p]
        L.orc_fill_complex_noise_cb.argtypes = [ctypes.c_int] * 3 + [ctypes.c_uint64, ctypes.c_uint32, _c_float_p]
        L.orc_fill_complex_noise_cb_planes.argtypes = ([ctypes.c_int] * 3 + [ctypes.c_uint64, ctypes.c_uint32, _c_int_p,
                                                       ctypes.c_int, _c_float_p])
        L.orc_philox4x32_10.argtypes = [_c_u32_p, _c_u32_p, _c_u32_p]
        L.orc_colour_f32.restype = ctypes.c_long
        L.orc_colour_f32.argtypes = [ctypes.c_long, _c_float_p, _c_float_p]
        L.orc_colour_f64.restype = ctypes.c_long
        L.orc_colour_f64.argtypes = [ctypes.c_long, _c_double_p, _c_double_p]
        L.orc_ricker.restype = ctypes.c_double
        L.orc_ricker.argtypes = [ctypes.c_double, ctypes.c_double]
        L.orc_aki_richards.argtypes = [ctypes.c_double, ctypes.c_double, _c_double_p]
        L.orc_spherical_corr.restype = ctypes.c_float
        L.orc_spherical_corr.argtypes = [ctypes.c_float] * 5
        L.orc_corr_xy.argtypes = [ctypes.c_int, ctypes.c_int] + [ctypes.c_float] * 5 + [_c_float_p]
        L.orc_circ_corr_t.restype = ctypes.c_int
        L.orc_circ_corr_t.argtypes = [_c_double_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, _c_float_p]
        L.orc_fill_padded.argtypes = [ctypes.c_int] * 6 + [_c_float_p, _c_float_p]
        L.orc_fill_number.restype = ctypes.c_int
        L.orc_fill_number.argtypes = [ctypes.c_int] * 3
        L.orc_dist_to_boundary.restype = ctypes.c_float
        L.orc_dist_to_boundary.argtypes = [ctypes.c_int] * 3
        L.orc_extrapolate_seismic.argtypes = [ctypes.c_int] * 6 + [_c_float_p] + [ctypes.c_int] * 4
        L.orc_interpolate_seismic.restype = ctypes.c_int
        L.orc_interpolate_seismic.argtypes = [ctypes.c_int] * 6 + [_c_float_p, ctypes.c_float, _c_int_p]
        L.orc_fill_param_corr.argtypes = [ctypes.c_int] * 3 + [_c_float_p, _c_float_p] + [ctypes.c_float] * 3 + [_c_float_p]
        L.orc_fill_err_corr.argtypes = [ctypes.c_int] * 3 + [_c_float_p] + [ctypes.c_float] * 2 + [_c_float_p]
        L.orc_data_variance.restype = ctypes.c_float
        L.orc_data_variance.argtypes = [ctypes.c_int] * 6 + [_c_float_p, ctypes.POINTER(ctypes.c_long)]
        L.orc_kz_tables.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int, _c_float_p,
                                    _c_float_p, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                                    _c_float_p, _c_float_p, _c_float_p, _c_float_p, _c_int_p]
        _lib = L
    return _lib


def ref_available():
    return os.path.exists(REF_SO)


def ref():
    """The reference's own compiled code (FFTW 2.1.2, lib_matr, RandomGen)."""
    global _ref
    if _ref is None:
        if not ref_available():
            raise RuntimeError("oracle/_ref not built (needs the reference tree at build time)")
        R = ctypes.CDLL(REF_SO)
        for f in ("ref_grid_fft", "ref_grid_ifft"):
            getattr(R, f).argtypes = [ctypes.c_int] * 4 + [_c_float_p]
        R.ref_fft1d_r2c.argtypes = [ctypes.c_int, _c_float_p]
        R.ref_posterior.restype = ctypes.c_long
        R.ref_posterior.argtypes = ([ctypes.c_int] * 4 + [_c_float_p] * 4 + [_c_int_p, _c_double_p, ctypes.c_float,
                                    _c_double_p, ctypes.c_int] + [_c_float_p] * 4 + [_c_float_p] * 3 + [_c_u8_p])
        R.ref_colour.restype = ctypes.c_long
        R.ref_colour.argtypes = [ctypes.c_long, _c_float_p, _c_float_p]
        R.ref_rng_seed.argtypes = [ctypes.c_uint]
        R.ref_unif01.restype = ctypes.c_double
        R.ref_rnorm01.restype = ctypes.c_double
        R.ref_rnorm01_n.argtypes = [_c_double_p, ctypes.c_long]
        R.ref_fill_complex_noise.argtypes = [ctypes.c_int] * 3 + [_c_float_p]
        R.ref_matr_eigen.restype = ctypes.c_int
        R.ref_matr_eigen.argtypes = [ctypes.c_int, _c_double_p, _c_double_p, _c_double_p]
        _ref = R
    return _ref


# ----------------------------------------------------------------------------- helpers
def rsize(nxp, nyp, nzp):
    return 2 * (nxp // 2 + 1) * nyp * nzp


def csize(nxp, nyp, nzp):
    return (nxp // 2 + 1) * nyp * nzp


def to_padded(real, nxp):
    """(nzp, nyp, nxp) real array -> FFTW in-place layout (nzp, nyp, rnxp)."""
    nzp, nyp, _ = real.shape
    out = np.zeros((nzp, nyp, 2 * (nxp // 2 + 1)), dtype=real.dtype)
    out[:, :, :nxp] = real
    return out


def grid_fft(real, cubetype, dtype=np.float64):
    """FFTGrid::fftInPlace on a (nzp, nyp, nxp) real array; returns (nzp, nyp, cnxp) complex."""
    nzp, nyp, nxp = real.shape
    buf = np.ascontiguousarray(to_padded(np.asarray(real, dtype=dtype), nxp))
    if dtype == np.float64:
        lib().orc_grid_fft(nxp, nyp, nzp, cubetype, _ptr(buf, ctypes.c_double))
        return buf.view(np.complex128)
    lib().orc_grid_fft_f32(nxp, nyp, nzp, cubetype, _ptr(buf, ctypes.c_float))
    return buf.view(np.complex64)


def grid_ifft(spec, nxp, cubetype, dtype=np.float64):
    """FFTGrid::invFFTInPlace on a (nzp, nyp, cnxp) complex array; returns (nzp, nyp, nxp) real."""
    nzp, nyp, cnxp = spec.shape
    cdt = np.complex128 if dtype == np.float64 else np.complex64
    buf = np.ascontiguousarray(spec, dtype=cdt).view(dtype).copy()
    if dtype == np.float64:
        lib().orc_grid_ifft(nxp, nyp, nzp, cubetype, _ptr(buf, ctypes.c_double))
    else:
        lib().orc_grid_ifft_f32(nxp, nyp, nzp, cubetype, _ptr(buf, ctypes.c_float))
    return buf.reshape(nzp, nyp, 2 * cnxp)[:, :, :nxp].copy()


def ref_grid_fft(real, cubetype):
    nzp, nyp, nxp = real.shape
    buf = np.ascontiguousarray(to_padded(np.asarray(real, dtype=np.float32), nxp))
    ref().ref_grid_fft(nxp, nyp, nzp, cubetype, _ptr(buf, ctypes.c_float))
    return buf.view(np.complex64)


def ref_grid_ifft(spec, nxp, cubetype):
    nzp, nyp, cnxp = spec.shape
    buf = np.ascontiguousarray(spec, dtype=np.complex64).view(np.float32).copy()
    ref().ref_grid_ifft(nxp, nyp, nzp, cubetype, _ptr(buf, ctypes.c_float))
    return buf.reshape(nzp, nyp, 2 * cnxp)[:, :, :nxp].copy()


def make_spec(p, keep):
    """Build a PostSpec from a problem dict (see tests/problem.py); `keep` holds references."""
    s = PostSpec()
    s.nxp, s.nyp, s.nzp, s.ntheta = p["nxp"], p["nyp"], p["nzp"], p["ntheta"]
    arrs = {
        "K": np.ascontiguousarray(p["K"], dtype=np.complex64).view(np.float32),
        "e1": np.ascontiguousarray(p["e1"], dtype=np.complex64).view(np.float32),
        "e2": np.ascontiguousarray(p["e2"], dtype=np.complex64).view(np.float32),
        "e3": np.ascontiguousarray(p["e3"], dtype=np.complex64).view(np.float32),
        "invert": np.ascontiguousarray(p["invert"], dtype=np.int32),
        "sigma0e": np.ascontiguousarray(p["sigma0e"], dtype=np.float64),
        "prior_var0": np.ascontiguousarray(p["prior_var0"], dtype=np.float64),
    }
    keep.update(arrs)
    s.K = _ptr(arrs["K"], ctypes.c_float)
    s.e1 = _ptr(arrs["e1"], ctypes.c_float)
    s.e2 = _ptr(arrs["e2"], ctypes.c_float)
    s.e3 = _ptr(arrs["e3"], ctypes.c_float)
    s.invert = _ptr(arrs["invert"], ctypes.c_int)
    s.sigma0e = _ptr(arrs["sigma0e"], ctypes.c_double)
    s.wnc = p["wnc"]
    s.prior_var0 = _ptr(arrs["prior_var0"], ctypes.c_double)
    s.prior_mode = p.get("prior_mode", 1)
    for name in ("rxy", "rt", "lam2d"):
        if p.get(name) is not None:
            keep[name] = np.ascontiguousarray(p[name], dtype=np.float32)
            setattr(s, name, _ptr(keep[name], ctypes.c_float))
    if p.get("lam_grid") is not None:
        keep["lam_grid"] = np.ascontiguousarray(p["lam_grid"], dtype=np.complex64).view(np.float32)
        s.lam_grid = _ptr(keep["lam_grid"], ctypes.c_float)
    return s


def posterior(p, mean, data, cov=None, precision="f32"):
    """Per-wavenumber update.  mean (3, nzp, nyp, cnxp), data (nt, ...), cov (6, ...) complex.
    Returns (mean_post, residual, cov_post, flags, nfail)."""
    keep = {}
    s = make_spec(p, keep)
    dt = np.complex64 if precision == "f32" else np.complex128
    rt = np.float32 if precision == "f32" else np.float64
    m = np.ascontiguousarray(mean, dtype=dt).copy()
    d = np.ascontiguousarray(data, dtype=dt).copy()
    shape = m.shape[1:]
    c = np.zeros((6,) + shape, dtype=dt) if cov is None else np.ascontiguousarray(cov, dtype=dt).copy()
    flags = np.zeros(shape, dtype=np.uint8)
    if precision == "f32":
        nf = lib().orc_posterior_f32(ctypes.byref(s), _ptr(m.view(rt), ctypes.c_float),
                                     _ptr(d.view(rt), ctypes.c_float), _ptr(c.view(rt), ctypes.c_float),
                                     _ptr(flags, ctypes.c_uint8))
    else:
        nf = lib().orc_posterior_f64(ctypes.byref(s), _ptr(m.view(rt), ctypes.c_double),
                                     _ptr(d.view(rt), ctypes.c_double), _ptr(c.view(rt), ctypes.c_double),
                                     _ptr(flags, ctypes.c_uint8))
    return m, d, c, flags, nf


def ref_posterior(p, mean, data, cov=None):
    keep = {}
    s = make_spec(p, keep)
    m = np.ascontiguousarray(mean, dtype=np.complex64).copy()
    d = np.ascontiguousarray(data, dtype=np.complex64).copy()
    shape = m.shape[1:]
    c = np.zeros((6,) + shape, dtype=np.complex64) if cov is None else np.ascontiguousarray(cov, dtype=np.complex64).copy()
    flags = np.zeros(shape, dtype=np.uint8)
    f = ctypes.c_float
    nf = ref().ref_posterior(s.nxp, s.nyp, s.nzp, s.ntheta, s.K, s.e1, s.e2, s.e3, s.invert, s.sigma0e, s.wnc,
                             s.prior_var0, s.prior_mode, s.rxy, s.rt, s.lam_grid, s.lam2d,
                             _ptr(m.view(np.float32), f), _ptr(d.view(np.float32), f),
                             _ptr(c.view(np.float32), f), _ptr(flags, ctypes.c_uint8))
    return m, d, c, flags, nf


def colour(cov, seed, precision="f32"):
    dt = np.complex64 if precision == "f32" else np.complex128
    c = np.ascontiguousarray(cov, dtype=dt)
    z = np.ascontiguousarray(seed, dtype=dt).copy()
    n = c[0].size
    if precision == "f32":
        nf = lib().orc_colour_f32(n, _ptr(c.view(np.float32), ctypes.c_float), _ptr(z.view(np.float32), ctypes.c_float))
    else:
        nf = lib().orc_colour_f64(n, _ptr(c.view(np.float64), ctypes.c_double), _ptr(z.view(np.float64), ctypes.c_double))
    return z, nf


def ref_colour(cov, seed):
    c = np.ascontiguousarray(cov, dtype=np.complex64)
    z = np.ascontiguousarray(seed, dtype=np.complex64).copy()
    nf = ref().ref_colour(c[0].size, _ptr(c.view(np.float32), ctypes.c_float), _ptr(z.view(np.float32), ctypes.c_float))
    return z, nf


def fill_complex_noise(nxp, nyp, nzp, state):
    """Reference-stream noise; `state` is a 1-element uint32 array advanced in place."""
    out = np.zeros((nzp, nyp, nxp // 2 + 1), dtype=np.complex64)
    lib().orc_fill_complex_noise(nxp, nyp, nzp, _ptr(state, ctypes.c_uint32), _ptr(out.view(np.float32), ctypes.c_float))
    return out


def fill_complex_noise_cb(nxp, nyp, nzp, seed, stream):
    out = np.zeros((nzp, nyp, nxp // 2 + 1), dtype=np.complex64)
    lib().orc_fill_complex_noise_cb(nxp, nyp, nzp, seed, stream, _ptr(out.view(np.float32), ctypes.c_float))
    return out


def data_variance(rv, nx, ny, nz, nxp):
    """ModelAVODynamic::ComputeDataVariance of a padded real grid rv (nzp, nyp, rnxp): (float variance,
    non-zero box samples)."""
    a = np.ascontiguousarray(rv, dtype=np.float32)
    nzp, nyp, _ = a.shape
    n = ctypes.c_long()
    v = lib().orc_data_variance(nx, ny, nz, nxp, nyp, nzp, _ptr(a, ctypes.c_float), ctypes.byref(n))
    return float(np.float32(v)), n.value


def fill_complex_noise_cb_planes(nxp, nyp, nzp, seed, stream, ks):
    """Counter noise of fill_complex_noise_cb on the k planes `ks` only: (len(ks), nyp, cnxp)."""
    kk = np.ascontiguousarray(ks, dtype=np.int32)
    out = np.zeros((len(kk), nyp, nxp // 2 + 1), dtype=np.complex64)
    lib().orc_fill_complex_noise_cb_planes(nxp, nyp, nzp, seed, stream, _ptr(kk, ctypes.c_int), len(kk),
                                           _ptr(out.view(np.float32), ctypes.c_float))
    return out


def spectrum_planes(real, ks, cubetype, nxp=None):
    """k planes `ks` of FFTGrid::fftInPlace of a real grid, in FP64 (numpy), without the whole 3D
    transform: the z transform at those k is a matrix product over the planes, then a 2D FFT.
    real: (nzp, nyp, nxp) or padded (nzp, nyp, rnxp) with nxp given.  Returns (len(ks), nyp, cnxp)."""
    nzp, nyp = real.shape[:2]
    nxp = real.shape[2] if nxp is None else nxp
    cnxp = nxp // 2 + 1
    z = np.arange(nzp)
    E = np.exp(-2j * np.pi * np.outer(np.asarray(ks), z) / nzp)
    out = np.empty((len(ks), nyp, cnxp), dtype=np.complex128)
    for p in range(len(ks)):
        acc = np.zeros((nyp, nxp), dtype=np.complex128)
        for z0 in range(0, nzp, 32):  # bounded memory: 32 planes at a time
            sl = np.asarray(real[z0:z0 + 32, :, :nxp], dtype=np.float64)
            acc += np.tensordot(E[p, z0:z0 + 32], sl, axes=(0, 0))
        out[p] = np.fft.fft2(acc)[:, :cnxp]
    if cubetype != COVARIANCE:  # FFTGrid's float scale factor (fftgrid.cpp:1601)
        out *= float(np.float32(1.0) / np.sqrt(np.float32(float(nxp) * nyp * nzp)))
    return out


def real_planes(spec, zs, nxp, cubetype):
    """z planes `zs` of FFTGrid::invFFTInPlace of a (nzp, nyp, cnxp) spectrum, in FP64 (numpy): the
    inverse z transform at those z as a matrix product, then inverse y and the half-complex x
    inverse (imaginary parts of the x DC and Nyquist bins ignored, as FFTW's hc2r).  Returns
    (len(zs), nyp, nxp)."""
    nzp, nyp, cnxp = spec.shape
    k = np.arange(nzp)
    E = np.exp(2j * np.pi * np.outer(np.asarray(zs), k) / nzp)
    out = np.empty((len(zs), nyp, nxp))
    for p in range(len(zs)):
        acc = np.zeros((nyp, cnxp), dtype=np.complex128)
        for k0 in range(0, nzp, 32):
            acc += np.tensordot(E[p, k0:k0 + 32], np.asarray(spec[k0:k0 + 32], dtype=np.complex128), axes=(0, 0))
        out[p] = np.fft.irfft(np.fft.ifft(acc, axis=0) * nyp, n=nxp, axis=1) * nxp  # unnormalised
    N = float(nxp) * nyp * nzp
    if cubetype == COVARIANCE:
        return out / N
    return out * float(np.float32(1.0 / np.sqrt(np.float32(N))))  # fftgrid.cpp:1634


def kz_tables(nz, nzp, dt_ms, wavelets, A, low_cut=5.0, high_cut=55.0, min_rel_thick=1.0):
    nt = wavelets.shape[0]
    w = np.ascontiguousarray(wavelets, dtype=np.float32)
    a = np.ascontiguousarray(A, dtype=np.float32)
    K = np.zeros((nzp, nt, 3), dtype=np.complex64)
    e1 = np.zeros((nzp, nt), dtype=np.complex64)
    e2 = np.zeros((nzp, nt), dtype=np.complex64)
    e3 = np.zeros((nzp, nt), dtype=np.complex64)
    inv = np.zeros(nzp, dtype=np.int32)
    f = ctypes.c_float
    lib().orc_kz_tables(nz, nzp, dt_ms, nt, _ptr(w, f), _ptr(a, f), low_cut, high_cut, min_rel_thick,
                        _ptr(K.view(np.float32), f), _ptr(e1.view(np.float32), f), _ptr(e2.view(np.float32), f),
                        _ptr(e3.view(np.float32), f), _ptr(inv, ctypes.c_int))
    return K, e1, e2, e3, inv


def stretched_spectrum(half, nzp, k, scale):
    """Wavelet::getCAmp(k, scale) (wavelet.cpp:337-394): the half spectrum `half` (bins 0..nzp/2)
    read at k / scale by linear interpolation between bins (zero at or beyond nzp/2 + 1), the
    Hermitian mirror for k > nzp/2, divided by scale."""
    cnzp = nzp // 2 + 1
    mirror = k >= cnzp
    om = (nzp - k if mirror else k) / scale
    if om >= cnzp:
        return 0j
    lo = int(np.floor(om))
    hi = lo + 1 if lo + 1 < cnzp else lo
    d = om - lo
    v = half[lo] * (1 - d) + half[hi] * d
    return (np.conj(v) if mirror else v) / scale


def kz_tables_variable_thickness(nz, nzp, dt_ms, wavelets, A, low_cut=5.0, high_cut=55.0, min_rel_thick=1.0):
    """Per-k_z tables of the non-constant-thickness branch (avoinversion.cpp:976-1014), in double:
    K = D A, e1 = 1 / |w|, e2 = D3 / |errorSmooth3|, e3 = fillInverseAbskWRobust (:1296-1322);
    D = FFT of the forward-difference sequence, D3 = FFT of central(backward(forward-difference))
    (wavelet1D.cpp:389-500)."""
    w = np.asarray(wavelets, dtype=np.float64)
    nt = w.shape[0]
    d1 = np.zeros(nzp)
    d1[0], d1[nzp - 1] = -1.0, 1.0
    d2 = d1 - np.roll(d1, 1)
    d3 = 0.5 * (np.roll(d2, -1) - np.roll(d2, 1))
    D, D3 = np.fft.fft(d1), np.fft.fft(d3)
    K = np.zeros((nzp, nt, 3), dtype=np.complex128)
    e1 = np.zeros((nzp, nt), dtype=np.complex128)
    e2 = np.zeros((nzp, nt), dtype=np.complex128)
    e3 = np.zeros((nzp, nt), dtype=np.complex128)
    for l in range(nt):
        f = np.roll(w[l], -1) - w[l]
        b = f - np.roll(f, 1)
        c = 0.5 * (np.roll(b, -1) - np.roll(b, 1))
        W = np.fft.fft(w[l])
        half = W[:nzp // 2 + 1]
        K[:, l, :] = D[:, None] * np.asarray(A[l], dtype=np.float64)[None, :]
        e1[:, l] = 1.0 / np.linalg.norm(w[l])
        e2[:, l] = D3 / np.linalg.norm(c)
        for k in range(nzp):
            m2 = max(abs(W[k]) ** 2, abs(stretched_spectrum(half, nzp, k, 0.999)) ** 2)
            e3[k, l] = 1.0 / np.sqrt(m2) if m2 > 0 else np.linalg.norm(w[l]) * nzp * nzp * 100.0
    k = np.arange(nzp)
    rf = nz * 1000.0 / (nz * dt_ms * nzp) * np.minimum(k, nzp - k)
    inv = ((rf > low_cut * min_rel_thick) & (rf < high_cut)).astype(np.int32)
    return K, e1, e2, e3, inv


def reflection_time_covariance_spectrum(corr_t_filtered, A, prior_var0):
    """AVOInversion::computeReflectionCoefficientTimeCovariance (avoinversion.cpp:368-401) and its
    operator FFT (:649-650): the elastic-impedance variance A S0 A^T times the filtered circular
    temporal correlation, the second difference 2 c[i] - c[i+1] - c[i-1] (cyclic), half spectrum."""
    ct = np.asarray(corr_t_filtered, dtype=np.float64)
    A = np.asarray(A, dtype=np.float64)
    ei = float(A @ np.asarray(prior_var0, dtype=np.float64) @ A) * ct
    rc = 2 * ei - np.roll(ei, -1) - np.roll(ei, 1)
    return np.fft.rfft(rc)


def scale_wavelet_filter(rv, nx, ny, nz, nxp, dt_ms, wavelet, mode, rel_thick, low_cut=5.0, high_cut=55.0,
                         wnc=0.1, error_var=1.0, tau_spec=None, shift=None, gain=None, stretch=None):
    """AVOInversion::divideDataByScaleWavelet (mode "divide", avoinversion.cpp:485-579, with
    computeAdjustmentFactor :582-719) or multiplyDataByScaleWaveletAndWriteToFile (mode "multiply",
    :721-790) on a copy of the padded real grid rv (nzp, nyp, rnxp), trace by trace in double.
    rel_thick (nyp, nxp): the simbox's relative thickness at every padded trace; shift (ms), gain,
    stretch (ny, nx): the local wavelet grids (None: 0, 1, 1); tau_spec: the half spectrum of the
    reflection-coefficient time covariance (reflection_time_covariance_spectrum).  The wavelet is
    (nzp,) in FFT order; its time-domain norm is |w| (shiftAndScale leaves the norm unchanged)."""
    out = np.array(rv, dtype=np.float64, copy=True)
    nzp, nyp, _ = out.shape
    cnzp = nzp // 2 + 1
    w = np.asarray(wavelet, dtype=np.float64)
    W = np.fft.rfft(w)
    wnorm = np.linalg.norm(w)
    om = 2 * np.pi * np.arange(cnzp) / nzp
    D3 = 1j * np.sin(om) * (2 * np.cos(om) - 2)  # central(backward(forward)) difference operator
    lz = nz * dt_ms
    tol_fac = 0.05
    maxfreq = (nzp // 2) * 1000.0 * nz / (lz * nzp)
    modW = wnorm ** 2 * maxfreq / high_cut
    d_omega = 1.0 / (nzp * dt_ms * 0.001)  # findNormWithinFrequencyBand's bins (wavelet.cpp:903-925)
    lo_i = max(0, int(np.floor(low_cut / d_omega)))
    hi_i = min(cnzp, int(np.ceil(high_cut / d_omega)))
    ks = np.arange(cnzp)

    def band_norm(c):
        n2 = 2.0 / nzp * np.sum(np.abs(c[lo_i:hi_i]) ** 2)
        if lo_i == 0:
            n2 -= np.abs(c[0]) ** 2 / nzp
        return np.sqrt(n2)

    def stretched(c, sf):
        return np.array([stretched_spectrum(c, nzp, k, sf) for k in range(cnzp)])

    def local(i, j):
        ts = 0.0 if shift is None or not (0 <= i < nx and 0 <= j < ny) else float(shift[j, i])
        g = 1.0 if gain is None or not (0 <= i < nx and 0 <= j < ny) else float(gain[j, i])
        st = 1.0 if stretch is None or not (0 <= i < nx and 0 <= j < ny) else float(stretch[j, i])
        # Wavelet1D::shiftAndScale (wavelet1D.cpp:580-602)
        amp = g * np.exp(-2j * np.pi * (ts / dt_ms) * ks / nzp)
        return W * amp, st

    taper = np.ones(nzp)
    for k in range(nz + 1, nzp):
        d = dist_to_boundary(k, nz, nzp)
        taper[k] = max(1.0 - d * d, 0.0)
    xs = range(nxp) if mode == "divide" else range(nx)
    ys = range(nyp) if mode == "divide" else range(ny)
    for j in ys:
        for i in xs:
            if mode == "divide":  # the local wavelet of the mirrored box trace
                ii = 0 if i > 3 * nx - 1 else i
                jj = 0 if j > 3 * ny - 1 else j
                if ii > (nxp + nx) // 2:
                    ii = nxp - ii
                if ii >= nx:
                    ii = 2 * nx - ii - 1
                if jj > (nyp + ny) // 2:
                    jj = nyp - jj
                if jj >= ny:
                    jj = 2 * ny - jj - 1
            else:
                ii, jj = i, j
            Wl, st = local(ii, jj)
            rel = float(rel_thick[j, i])
            sf = rel * st
            tr = out[:, j, i] / np.sqrt(nzp)
            if mode == "divide":
                tr = tr * taper
            X = np.fft.rfft(tr)
            if mode == "multiply":
                Y = X * np.conj(stretched(Wl, sf))
            else:
                Sl = D3 * Wl
                Sl = Sl / band_norm(Sl)
                wloc = stretched(Wl, sf)
                sloc = stretched(Sl, sf)
                size_loc = np.abs(wloc) ** 2 / wnorm ** 2 + np.abs(sloc) ** 2
                size_glob = np.abs(Wl) ** 2 / wnorm ** 2 + np.abs(Sl) ** 2
                s2loc = error_var * (wnc + (1 - wnc) * 0.5 * size_loc)
                s2glob = error_var * (wnc + (1 - wnc) * 0.5 * size_glob)
                tau2 = np.abs(np.real(tau_spec)) if tau_spec is not None else np.zeros(cnzp)
                wl2 = np.abs(wloc) ** 2
                wg2 = np.abs(W) ** 2
                wg2a = np.where(wg2 < tol_fac * modW, 0.5 * (wg2 + tol_fac * modW), wg2)
                wl2a = np.where(wl2 < tol_fac * modW, 0.5 * (wl2 + tol_fac * modW), wl2)
                low = wl2a < tol_fac * wg2a
                wl2a_new = np.where(low, 0.5 * (wl2 + tol_fac * wg2a), wl2a)
                with np.errstate(divide="ignore", invalid="ignore"):
                    s2loc = np.where(low, s2loc * wl2a_new / wl2, s2loc)
                wl2a = wl2a_new
                factor = (wloc / (tau2 * wl2a + s2loc)) / (wg2a / (tau2 * wg2 + s2glob))
                deltaF = nz * 1000.0 / (rel * lz * nzp)
                band = (deltaF * ks < high_cut) & (deltaF * ks > low_cut)
                Y = np.where(band, X * factor, 0.0)
            out[:, j, i] = np.fft.irfft(Y, n=nzp) * nzp / np.sqrt(nzp)
    return out


def corr_xy(nxp, nyp, dx, dy, range1, range2=None, angle=0.0):
    out = np.zeros((nyp, nxp), dtype=np.float32)
    lib().orc_corr_xy(nxp, nyp, dx, dy, range1, range1 if range2 is None else range2, angle,
                      _ptr(out, ctypes.c_float))
    return out


def circ_corr_t(corr_t, min_int_fq, nzp):
    c = np.ascontiguousarray(corr_t, dtype=np.float64)
    out = np.zeros(nzp, dtype=np.float32)
    err = lib().orc_circ_corr_t(_ptr(c, ctypes.c_double), len(c), min_int_fq, nzp, _ptr(out, ctypes.c_float))
    if err:
        raise ValueError("circular temporal correlation undefined (too few layers)")
    return out


def fill_padded(cube, nxp, nyp, nzp):
    """FFTGrid(const NRLib::Grid<float> *) padding fill: cube (nz, ny, nx) float32 -> padded real
    grid (nzp, nyp, rnxp)."""
    c = np.ascontiguousarray(cube, dtype=np.float32)
    nz, ny, nx = c.shape
    out = np.zeros((nzp, nyp, 2 * (nxp // 2 + 1)), dtype=np.float32)
    f = ctypes.c_float
    lib().orc_fill_padded(nx, ny, nz, nxp, nyp, nzp, _ptr(c, f), _ptr(out, f))
    return out


def fill_number(i, n, np_):
    return lib().orc_fill_number(i, n, np_)


def dist_to_boundary(i, n, np_):
    return lib().orc_dist_to_boundary(i, n, np_)


def extrapolate_seismic(rv, nx, ny, nz, nxp, rect):
    """FFTGrid::extrapolateSeismic on a copy of the padded real grid rv (nzp, nyp, rnxp)."""
    out = np.array(rv, dtype=np.float32, order="C", copy=True)
    nzp, nyp, _ = out.shape
    lib().orc_extrapolate_seismic(nx, ny, nz, nxp, nyp, nzp, _ptr(out, ctypes.c_float), *rect)
    return out


def interpolate_seismic(rv, nx, ny, nz, nxp, threshold):
    """FFTGrid::interpolateSeismic on a copy of rv; returns (grid, bad traces, rect)."""
    out = np.array(rv, dtype=np.float32, order="C", copy=True)
    nzp, nyp, _ = out.shape
    rect = np.zeros(4, dtype=np.int32)
    nbad = lib().orc_interpolate_seismic(nx, ny, nz, nxp, nyp, nzp, _ptr(out, ctypes.c_float), threshold,
                                         _ptr(rect, ctypes.c_int))
    return out, nbad, tuple(int(v) for v in rect)


CRAVA_TAG = b"crava_fftgrid_binary"
CRAVA_GEOMETRY = (("x0", "d"), ("y0", "d"), ("dx", "d"), ("dy", "d"), ("nx", "i"), ("ny", "i"), ("il0", "d"),
                  ("xl0", "d"), ("il_step_x", "d"), ("il_step_y", "d"), ("xl_step_x", "d"), ("xl_step_y", "d"),
                  ("angle", "d"))


def write_crava(path, rv, geometry):
    """FFTGrid::writeCravaFile (fftgrid.cpp:2183-2221): tag line, geometry, rnxp, nyp, nzp, then the
    padded real grid rv (nzp, nyp, rnxp), everything big-endian (NRLib's default)."""
    import struct
    nzp, nyp, rnxp = rv.shape
    with open(path, "wb") as f:
        f.write(CRAVA_TAG + b"\n")
        for name, code in CRAVA_GEOMETRY:
            f.write(struct.pack(">" + code, geometry[name]))
        f.write(struct.pack(">iii", rnxp, nyp, nzp))
        f.write(np.ascontiguousarray(rv, dtype=">f4").tobytes())


def read_crava(path):
    """FFTGrid::readCravaFile (fftgrid.cpp:2224-2271): returns (geometry, rv (nzp, nyp, rnxp))."""
    import struct
    with open(path, "rb") as f:
        f.readline()
        geometry = {}
        for name, code in CRAVA_GEOMETRY:
            geometry[name] = struct.unpack(">" + code, f.read(struct.calcsize(code)))[0]
        rnxp, nyp, nzp = struct.unpack(">iii", f.read(12))
        rv = np.frombuffer(f.read(4 * rnxp * nyp * nzp), dtype=">f4").astype(np.float32)
    return geometry, rv.reshape(nzp, nyp, rnxp)


def storm_header(cubetype, nx, ny, nz, geometry):
    """Simbox::getStormHeader (simbox.cpp:1012-1034), binary form: numbers with std::fixed and
    precision 6 (NRLib::ToString(v, 6)); a flat simbox (no surface names) writes top 0.0 and base lz."""
    f6 = lambda v: "%.6f" % v
    g = geometry
    h = "storm_petro_binary\n0 %d %s\nFFTGrid\n" % (cubetype, f6(-99999.0))
    if g.get("top_name") and g.get("bot_name"):
        h += " ".join([f6(g["x_min"]), f6(g["lx"]), f6(g["y_min"]), f6(g["ly"]), g["top_name"], g["bot_name"],
                       "0.0", "0.0"]) + "\n"
    else:
        h += " ".join([f6(g["x_min"]), f6(g["lx"]), f6(g["y_min"]), f6(g["ly"]), "0.0", f6(g["lz"]), "0.0", "0.0"]) + "\n"
    h += f6(g["lz"]) + " " + f6(g["angle"] * 180.0 / np.pi) + "\n\n"
    h += "%d %d %d\n" % (nx, ny, nz)
    return h


def write_storm(path, cube, cubetype, geometry):
    """FFTGrid::writeStormFile (fftgrid.cpp:1889-1957), binary: the header, then the cube (nz, ny, nx)
    as big-endian floats, i fastest, then "0\n"."""
    nz, ny, nx = cube.shape
    with open(path, "wb") as f:
        f.write(storm_header(cubetype, nx, ny, nz, geometry).encode())
        f.write(np.ascontiguousarray(cube, dtype=">f4").tobytes())
        f.write(b"0\n")


def read_storm(path):
    """StormContGrid::ReadFromFile (stormcontgrid.cpp:82-170) of a binary grid: (tokens of the header,
    cube (nz, ny, nx))."""
    with open(path, "rb") as f:
        data = f.read()
    toks, pos = [], 0
    while len(toks) < 18:  # format, 4 more, the volume (8 + 2), nx ny nz
        while data[pos:pos + 1].isspace():
            pos += 1
        end = pos
        while end < len(data) and not data[end:end + 1].isspace():
            end += 1
        toks.append(data[pos:end].decode())
        pos = end
    pos = data.index(b"\n", pos) + 1  # the rest of the nx ny nz line
    nx, ny, nz = int(toks[15]), int(toks[16]), int(toks[17])
    cube = np.frombuffer(data[pos:pos + 4 * nx * ny * nz], dtype=">f4").astype(np.float32).reshape(nz, ny, nx)
    return toks, cube


def ieee_to_ibm(x):
    """IBM single (hexadecimal) floats of float32 values by the definition, fraction truncated:
    (-1)^s 0.F 16^(E - 64), F a 24-bit fraction in [1/16, 1); both zeros give 0 (as the writer of
    libs/nrlib/iotools/fileio.hpp:1127-1139 does for normal floats)."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.int64)
    e = (u >> 23) & 0xFF
    mant = np.where(e > 0, (u & 0x7FFFFF) | 0x800000, u & 0x7FFFFF)
    p = np.where(e > 0, e, 1) - 126
    q = (p + 3) >> 2
    frac = mant >> (4 * q - p)
    out = (u & 0x80000000) | ((q + 64) << 24) | frac
    return np.where((u & 0x7FFFFFFF) == 0, 0, out).astype(np.uint32)


def ibm_to_float(b):
    """Values (float64) of IBM single floats given as uint32."""
    b = np.asarray(b, dtype=np.uint64)
    sign = np.where(b & 0x80000000, -1.0, 1.0)
    expo = ((b >> 24) & 0x7F).astype(np.int64) - 64
    frac = (b & 0xFFFFFF).astype(np.float64) / float(1 << 24)
    return sign * frac * np.power(16.0, expo)


def segy_write(real, nx, ny, nz, geo):
    """FFTGrid::writeSegyFile (fftgrid.cpp:1960-2050) with NRLib's SegY writer for a real padded grid
    (nzp, nyp, rnxp): the file as bytes.  geo: x0, y0, dx, dy, angle, il0, xl0, il_step_x,
    il_step_y, xl_step_x, xl_step_y, dz, z0 (float), zmax, top (ny, nx), rel_thick (ny, nx), text
    (lines from the second on).  Binary header as BinaryHeader::Write (commonheaders.cpp:234-270);
    trace headers at the SeisWorks positions with scalco 1 (traceheader.cpp:38-47, 527-586);
    traces sorted by inline then crossline (segy.cpp:1364-1430); samples as
    getRegularZInterpolatedRealValue (fftgrid.cpp:2275-2306), big-endian IBM floats."""
    import struct
    f32 = np.float32
    grid = np.asarray(real, dtype=np.float32)
    dzf = f32(math.floor(geo["dz"] + 0.5))
    if dzf == 0:
        dzf = f32(1.0)
    z0 = f32(geo["z0"])
    ns = int(math.ceil((geo["zmax"] - float(z0)) / float(dzf)))
    text = list(geo.get("text", []))
    # NRLib's TextualHeader (commonheaders.cpp:60-213): the caller's lines on the empty header
    # (last line 79 characters, then the string's terminating zero), or with none the standard
    # header with the writing program's credits on lines 12-28
    standard = {12: "THIS SEGY FILE WAS WRITTEN BY", 14: "THE CRAVA PROGRAM", 16: "DEVELOPED BY",
                18: "RAGNAR HAUGE", 20: "ANNE RANDI SYVERSVEEN           AND", 22: "ODD KOLBJORNSEN", 24: "AT",
                26: "THE NORWEGIAN COMPUTING CENTER", 28: "2003"}
    lines = []
    for l in range(39):
        if text:
            t = text[l - 1] if 1 <= l <= len(text) else ""
        else:
            t = " " * 7 + standard[l + 1] if l + 1 in standard else ""
        lines.append(("C%2d  %-75.75s" % (l + 1, t))[:80])
    lines.append("%-80s" % "C40 END EBCDIC")
    head = bytearray("".join(lines).encode("cp037"))
    if text:
        head[3199] = 0
    head = bytes(head)
    bh = bytearray(400)
    struct.pack_into(">iii", bh, 0, 9999, 1, 1)
    struct.pack_into(">h", bh, 12, nx)
    struct.pack_into(">h", bh, 16, int(1000 * dzf))
    struct.pack_into(">h", bh, 20, ns)
    struct.pack_into(">hhh", bh, 24, 1, 1, 4)
    struct.pack_into(">h", bh, 54, 1)
    struct.pack_into(">h", bh, 302, 1)
    cr, sr = math.cos(geo["angle"]), math.sin(geo["angle"])
    traces = []
    for j in range(ny):
        for i in range(nx):
            top = float(geo["top"][j, i])
            if top in (-99999.0, -88888.0):
                continue
            rx, ry = (i + 0.5) * geo["dx"], (j + 0.5) * geo["dy"]
            xf = float(f32(rx * cr - ry * sr + geo["x0"]))
            yf = float(f32(rx * sr + ry * cr + geo["y0"]))
            il = int(0.5 + geo["il0"] + (xf - geo["x0"]) * geo["il_step_x"] + (yf - geo["y0"]) * geo["il_step_y"])
            xl = int(0.5 + geo["xl0"] + (xf - geo["x0"]) * geo["xl_step_x"] + (yf - geo["y0"]) * geo["xl_step_y"])
            gdz = geo["dz"] * float(geo["rel_thick"][j, i])
            first = int(math.floor(0.5 + (top - float(z0)) / float(dzf)))
            end = min(ns, int(math.floor(0.5 + ((top - float(z0)) + nz * gdz) / float(dzf))))
            traces.append((il, xl, i, j, top, gdz, first, end, int(xf), int(yf)))
    traces.sort(key=lambda t: (t[0], t[1]))  # stable: equal pairs keep the scan order
    out = [head, bytes(bh)]
    k = np.arange(ns)
    for il, xl, i, j, top, gdz, first, end, ux, uy in traces:
        z = (np.float64(z0) + np.float64(dzf) * k).astype(f32)
        t = ((z.astype(np.float64) - top) / gdz).astype(f32)
        kg = np.trunc(t).astype(np.int64)
        t = (t - kg.astype(f32)).astype(f32)
        miss = (kg < -1) | (kg > nz - 1)
        lo = (kg < 0) & ~miss
        hi = (kg > nz - 2) & (kg >= 0) & ~miss
        kg = np.where(lo, 0, np.where(hi, nz - 2, kg))
        t = np.where(lo, f32(0), np.where(hi, f32(1), t)).astype(f32)
        col = grid[:, j, i]

        def at(kk):
            ok = (kk >= 0) & (kk < nz)
            return np.where(ok, col[np.clip(kk, 0, max(nz - 1, 0))], f32(-99999.0)).astype(f32)

        v = ((f32(1) - t) * at(kg) + t * at(kg + 1)).astype(f32)
        v = np.where(miss, f32(-99999.0), v).astype(f32)
        v = np.where((k >= first) & (k < end), v, f32(0)).astype(f32)
        th = bytearray(240)
        struct.pack_into(">i", th, 8, il)
        struct.pack_into(">i", th, 20, xl)
        struct.pack_into(">h", th, 70, 1)
        struct.pack_into(">ii", th, 72, ux, uy)
        struct.pack_into(">h", th, 108, int(z0))
        struct.pack_into(">hH", th, 114, ns, int(dzf * 1000))
        out.append(bytes(th))
        out.append(ieee_to_ibm(v).astype(">u4").tobytes())
    return b"".join(out)


def fill_param_corr(corr_xy, corr_t, grad_i=0.0, grad_j=0.0, scale=1.0):
    """fillInParamCorr / FillInLateralCorr then multiplyByScalar: corr_xy (nyp, nxp), corr_t (nzp,)
    -> padded real grid (nzp, nyp, rnxp)."""
    cxy = np.ascontiguousarray(corr_xy, dtype=np.float32)
    ct = np.ascontiguousarray(corr_t, dtype=np.float32)
    nyp, nxp = cxy.shape
    nzp = len(ct)
    out = np.zeros((nzp, nyp, 2 * (nxp // 2 + 1)), dtype=np.float32)
    f = ctypes.c_float
    lib().orc_fill_param_corr(nxp, nyp, nzp, _ptr(cxy, f), _ptr(ct, f), grad_i, grad_j, scale, _ptr(out, f))
    return out


def fill_err_corr(corr_xy, nzp, grad_i=0.0, grad_j=0.0):
    cxy = np.ascontiguousarray(corr_xy, dtype=np.float32)
    nyp, nxp = cxy.shape
    out = np.zeros((nzp, nyp, 2 * (nxp // 2 + 1)), dtype=np.float32)
    lib().orc_fill_err_corr(nxp, nyp, nzp, _ptr(cxy, ctypes.c_float), grad_i, grad_j, _ptr(out, ctypes.c_float))
    return out


def make_circ_auto_cov_pos_def(cac):
    """SeismicParametersHolder::MakeCircAutoCovPosDef (seismicparametersholder.cpp:661-767) in
    double: spectra of the nine series, per frequency D V |diag(w)| V^H D with (w, V) the general
    complex eigendecomposition of D^-1 C D^-1 (D = lag-0 standard deviations), back to lags / nzp."""
    a = np.asarray(cac, dtype=np.float64)
    nzp = a.shape[2]
    d = np.sqrt(np.array([a[i, i, 0] for i in range(3)]))
    spec = np.fft.rfft(a, axis=2)  # sign -1, unnormalised, bins 0..nzp/2
    out = np.empty_like(spec)
    for m in range(spec.shape[2]):
        M = spec[:, :, m] / d[:, None] / d[None, :]
        w, V = np.linalg.eig(M)
        out[:, :, m] = d[:, None] * (V @ np.diag(np.abs(w)) @ V.conj().T) * d[None, :]
    return np.fft.irfft(out, n=nzp, axis=2)


def synthetic_seismic(vp, vs, rho, wavelets, A, nzp):
    """CravaResult::ComputeSyntSeismic (cravaresult.cpp:2312-2406) trace by trace, in double:
    impedance A_l . (vp, vs, rho), the trace taper over the z padding
    (fac ((k - nz) r_0 + (nzp - k - 1) r_{nz-1}), fac = 1 / (nzp - nz - 1)), backward difference
    (cyclic), forward FFT, product with the conjugate wavelet spectrum, inverse FFT / nzp, the
    first nz samples."""
    p = np.stack([np.asarray(x, dtype=np.float64) for x in (vp, vs, rho)])
    nz = p.shape[1]
    out = []
    for l in range(len(wavelets)):
        imp = np.tensordot(np.asarray(A[l], dtype=np.float64), p, axes=1)  # (nz, ny, nx)
        tr = np.zeros((nzp,) + imp.shape[1:])
        tr[:nz] = imp
        k = np.arange(nz, nzp)[:, None, None]
        tr[nz:] = ((k - nz) * imp[0] + (nzp - k - 1) * imp[nz - 1]) / (nzp - nz - 1)
        r = tr - np.roll(tr, 1, axis=0)
        W = np.fft.rfft(np.asarray(wavelets[l], dtype=np.float64))
        s = np.fft.rfft(r, axis=0) * np.conj(W)[:, None, None]
        out.append(np.fft.irfft(s, n=nzp, axis=0)[:nz])
    return out


# ---------------------------------------------------------------- local noise (correctVpVsRho)
RMISSING = -99999.0


def _ev(m):
    """NRLib::ComputeEigenVectors (libs/nrlib/flens/nrlib_flens.cpp:114-121): flens::ev on a
    general matrix, i.e. LAPACK dgeev, keeping the real parts of the eigenvalues and the right
    eigenvectors as columns.  numpy.linalg.eig calls the same LAPACK driver."""
    w, v = np.linalg.eig(np.asarray(m, dtype=np.float64))
    return np.real(w).copy(), np.real(v).copy()


def _vfv(v, f):
    """(V diag(f)) V^T: the reference's `H = Evec * EvalMat; X = H * EvecT`."""
    return (v * np.asarray(f, dtype=np.float64)[None, :]) @ v.T


def sort3x3(w, v):
    """NRLib::Sort3x3 (nrlib_flens.cpp:247-280): eigenvalues in decreasing order (the first of
    equal ones first), eigenvector columns permuted alike."""
    idx = []
    for _ in range(3):
        best, bi = -100000.0, -1
        for i in range(3):
            if i not in idx and w[i] > best:
                best, bi = w[i], i
        idx.append(bi)
    return np.array([w[i] for i in idx]), v[:, idx].copy()


def compute_g(prior_var0, post_var0, err_theta_cov):
    """AVOInversion::computeG (avoinversion.cpp:1943-2026): the n_theta x 3 operator G with
    Sigma_m - Sigma_m G^T (G Sigma_m G^T + Sigma_e)^-1 G Sigma_m = postVar0 (the posterior the
    full inversion gave at lag 0), from the eigen decompositions of Sigma_m - postVar0, Sigma_m,
    A = Sigma_m^-1/2 (Sigma_m - postVar0) Sigma_m^-1/2 (sorted, Sort3x3) and Sigma_e."""
    Sm = np.asarray(prior_var0, dtype=np.float64)
    E = np.asarray(err_theta_cov, dtype=np.float64)
    nt = E.shape[0]
    w, V = _ev(Sm - np.asarray(post_var0, dtype=np.float64))
    Sdelta = _vfv(V, np.where(w > 0.0, w, 0.0))
    w, V = _ev(Sm)
    Sinv = _vfv(V, np.where(w > 1e-7, 1.0 / np.sqrt(np.where(w > 1e-7, w, 1.0)), 0.0))
    A = (Sinv @ Sdelta) @ Sinv
    w, V = sort3x3(*_ev(A))
    Lg = np.zeros((nt, 3))
    for i in range(min(nt, 3)):
        if w[i] > 0:
            Lg[i, i] = math.sqrt(w[i] / (1.0 - w[i]))
    we, Ve = _ev(E)
    Esq = _vfv(Ve, np.where(we > 0.0, np.sqrt(np.where(we > 0.0, we, 0.0)), 0.0))
    return ((Esq @ Lg) @ V.T) @ Sinv


def post_var0_inv_sqrt(post_var0, use_ref=False):
    """correctVpVsRho's postVar0^-1/2 (avoinversion.cpp:2182-2191): lib_matr_eigen, 1/sqrt of the
    positive eigenvalues, V D V^T.  use_ref: through the reference-compiled lib_matr_eigen."""
    a = np.ascontiguousarray(post_var0, dtype=np.float64)
    if use_ref:
        vec, val = np.zeros((3, 3)), np.zeros(3)
        ref().ref_matr_eigen(3, _ptr(a, ctypes.c_double), _ptr(vec, ctypes.c_double), _ptr(val, ctypes.c_double))
    else:
        val, vec = np.linalg.eigh(a)
    return _vfv(vec, np.where(val > 0.0, 1.0 / np.sqrt(np.where(val > 0.0, val, 1.0)), 0.0))


def new_posterior_cov_pointwise(G, scales, prior_var0, post_var0, err_theta_cov):
    """AVOInversion::newPosteriorCovPointwise (avoinversion.cpp:2028-2124) for one trace:
    sigmanew = sqrt(Sigma_m G^T M^-1 G Sigma_m) (Sigma_m - postVar0)^-1/2 with
    M = G Sigma_m G^T + D Sigma_e D, D = diag(sqrt(scales)) (eigenvalues <= 1e-8 of M dropped from
    its inverse, <= 1e-7 of Sigma_m - postVar0 from its inverse root), and
    sqrt(Sigma_m - Sigma_m G^T M^-1 G Sigma_m).  Returns (sigmanew, sigmamdnew)."""
    G = np.asarray(G, dtype=np.float64)
    Sm = np.asarray(prior_var0, dtype=np.float64)
    E = np.asarray(err_theta_cov, dtype=np.float64)
    nt = E.shape[0]
    D = np.zeros((nt, nt))
    for i in range(nt):
        D[i, i] = math.sqrt(scales[i])
    sigmaenew = (D @ E) @ D
    help_ = ((G @ Sm) @ G.T) + sigmaenew
    we, Ve = _ev(help_)
    Minv = _vfv(Ve, np.where(we > 1e-8, 1.0 / np.where(we > 1e-8, we, 1.0), 0.0))
    deltanew = (((Sm @ G.T) @ Minv) @ G) @ Sm
    w, V = _ev(Sm - deltanew)
    sigmamdnew = _vfv(V, np.where(w > 0.0, np.sqrt(np.where(w > 0.0, w, 0.0)), 0.0))
    w, V = _ev(deltanew)
    dsq = _vfv(V, np.where(w > 0.0, np.sqrt(np.where(w > 0.0, w, 0.0)), 0.0))
    w, V = _ev(Sm - np.asarray(post_var0, dtype=np.float64))
    sdi = _vfv(V, np.where(w > 1e-7, 1.0 / np.sqrt(np.where(w > 1e-7, w, 1.0)), 0.0))
    return dsq @ sdi, sigmamdnew


def find_min(a, missing=RMISSING):
    """NRLib::Grid2D::FindMin (libs/nrlib/grid/grid2d.hpp:182-191) over a grid in storage order."""
    flat = np.asarray(a, dtype=np.float64).ravel()
    m = flat[0]
    for v in flat:
        if (m == missing or v < m) and v != missing:
            m = v
    return m


def local_noise_matrices(G, scale_grids, prior_var0, post_var0, err_theta_cov):
    """correctVpVsRho's per-trace matrices (avoinversion.cpp:2206-2241).  scale_grids: (n_theta,
    ny, nx) local noise scales (Grid2D (i, j) at [j, i]).  Returns (sigmanew, stored), each
    (ny, nx, 3, 3): the posterior correction and sqrt(local posterior) postVar0^-1/2 (kept for
    the realisations, sigmamdnew_)."""
    sg = np.asarray(scale_grids, dtype=np.float64)
    nt, ny, nx = sg.shape
    mins = [find_min(sg[l]) for l in range(nt)]
    old = post_var0_inv_sqrt(post_var0)
    sn = np.zeros((ny, nx, 3, 3))
    st = np.zeros((ny, nx, 3, 3))
    for j in range(ny):
        for i in range(nx):
            sc = [sg[l, j, i] / mins[l] for l in range(nt)]
            a, b = new_posterior_cov_pointwise(G, sc, prior_var0, post_var0, err_theta_cov)
            sn[j, i] = a
            st[j, i] = b @ old
    return sn, st


def correct_vp_vs_rho(post, mean, nx, ny, nz, sigmanew):
    """The per-sample update of correctVpVsRho (avoinversion.cpp:2243-2263) on real-domain padded
    grids (3, nzp, nyp, rnxp) float32: diff = post - mean in float, then
    float(mean + s00 d0 + s01 d1 + s02 d2) evaluated in double, left to right."""
    out = np.array(post, dtype=np.float32, copy=True)
    p = np.asarray(post, dtype=np.float32)[:, :nz, :ny, :nx]
    m = np.asarray(mean, dtype=np.float32)[:, :nz, :ny, :nx]
    d = (p - m).astype(np.float32)  # float subtraction
    s = np.asarray(sigmanew, dtype=np.float64)[None]  # (1, ny, nx, 3, 3)
    for a in range(3):
        acc = m[a].astype(np.float64) + s[..., a, 0] * d[0].astype(np.float64)
        acc = acc + s[..., a, 1] * d[1].astype(np.float64)
        acc = acc + s[..., a, 2] * d[2].astype(np.float64)
        out[a, :nz, :ny, :nx] = acc.astype(np.float32)
    return out


def correct_realisation(seed, post, nx, ny, nz, stored):
    """simulate with local noise (avoinversion.cpp:1485-1504): inside the box the realisation's
    noise fields become float(s0 vp + s1 vs + s2 rho) (double, left to right), then the whole
    padded grid gets the posterior added in float (FFTGrid::add, fftgrid.cpp:1659-1678)."""
    out = np.array(seed, dtype=np.float32, copy=True)
    z = np.asarray(seed, dtype=np.float32)[:, :nz, :ny, :nx].astype(np.float64)
    s = np.asarray(stored, dtype=np.float64)[None]
    for a in range(3):
        acc = s[..., a, 0] * z[0] + s[..., a, 1] * z[1]
        acc = acc + s[..., a, 2] * z[2]
        out[a, :nz, :ny, :nx] = acc.astype(np.float32)
    return (out + np.asarray(post, dtype=np.float32)).astype(np.float32)


# ------------------------------------------------------------------ seismic traces onto the grid
# CommonData::FillInData (commondata.cpp:6451-6775), the is_segy, DATA (seismic) branch: every
# column of the padded grid takes its nearest SEG-Y trace, tapered in the guard zone, resampled
# ten times finer through its spectrum and interpolated at the grid's cell centres.

def smooth_trace_in_guard_zone(trace, dz_data, smooth_length):
    """CommonData::SmoothTraceInGuardZone (commondata.cpp:6847-6885): sin^2 taper over the first
    floor(smooth_length / dz) samples, cos^2 over the last as many (float products, the factors
    from double angles)."""
    t = np.array(trace, dtype=np.float32, copy=True)
    ns = int(math.floor(np.float32(smooth_length) / np.float32(dz_data)))
    n = len(t)
    for k in range(ns):
        s = np.float32(math.sin(0.5 * math.pi * (k / ns)))
        t[k] = np.float32(t[k] * np.float32(s * s))
    for k in range(ns):
        c = np.float32(math.cos(0.5 * math.pi * ((k + 1) / ns)))
        t[n - ns + k] = np.float32(t[n - ns + k] * np.float32(c * c))
    return t


def resample_trace(trace, nt):
    """CommonData::ResampleTrace (commondata.h:136-190) in double: the trace zero-padded to nt,
    its nt/2 + 1 bins as the lowest bins of a length-10 nt spectrum, inverse (Hermitian: a bin
    below the fine Nyquist counts twice, the imaginary part of DC is dropped), scaled by 1 / nt.
    Returns the rmt = 2 (5 nt + 1) floats the reference's in-place buffer holds: the fine trace
    and, past it, the zeros of the (always empty) fine Nyquist bin."""
    mt = 10 * nt
    x = np.zeros(nt)
    x[:len(trace)] = np.asarray(trace, dtype=np.float64)
    c = np.fft.rfft(x)
    cf = np.zeros(mt // 2 + 1, dtype=complex)
    cf[:len(c)] = c
    out = np.zeros(2 * (mt // 2 + 1))
    out[:mt] = np.fft.irfft(cf, mt) * (mt / nt)
    return out


def z_simbox_index(k, nz, nzp):
    """CommonData::GetZSimboxIndex (commondata.cpp:7052-7062)."""
    return k if k < (nz + nzp) // 2 else k - nzp


def interpolate_grid_values(fine, z0_grid, dz_grid, z0_data, dz_fine, nz, nzp):
    """CommonData::InterpolateGridValues (commondata.cpp:6947-6995): the value at cell k of a
    column (refk = GetZSimboxIndex(k)) from the fine trace, linear between the two fine samples
    around it, in the reference's float arithmetic; zero outside the fine trace."""
    n_fine = len(fine)
    f32 = np.float32
    z0_shift = f32(f32(z0_grid) - f32(z0_data))
    inv = f32(f32(1.0) / f32(dz_fine))
    out = np.zeros(nzp, dtype=np.float32)
    for k in range(nzp):
        refk = z_simbox_index(k, nz, nzp)
        dl = f32(f32(z0_shift + f32(f32(refk) * f32(dz_grid))) * inv)
        l1, l2 = int(math.floor(dl)), int(math.ceil(dl))
        if l2 < 0 or l1 > n_fine - 1:
            v = 0.0
        elif l1 < 0:
            v = fine[l2]
        elif l2 > n_fine - 1 or l1 == l2:
            v = fine[l1]
        else:
            w1, w2 = f32(math.ceil(dl) - dl), f32(dl - math.floor(dl))
            v = f32(w1 * f32(fine[l1])) + f32(w2 * f32(fine[l2]))
        out[k] = f32(v)
    return out


def interpolate_and_shift_trend(trend_long, z0_grid, dz_grid, z0_data, dz_fine, nz, nzp):
    """CommonData::InterpolateAndShiftTrend (commondata.cpp:6997-7050): as
    interpolate_grid_values on the fine trend, but the trend's first or last value outside it."""
    n_fine = len(trend_long)
    f32 = np.float32
    z0_shift = f32(f32(z0_grid) - f32(z0_data))
    inv = f32(f32(1.0) / f32(dz_fine))
    out = np.zeros(nzp, dtype=np.float32)
    for k in range(nzp):
        refk = z_simbox_index(k, nz, nzp)
        dl = f32(f32(z0_shift + f32(f32(refk) * f32(dz_grid))) * inv)
        l1, l2 = int(math.floor(dl)), int(math.ceil(dl))
        if l2 < 0 or l1 > n_fine - 1:
            v = trend_long[0] if l2 < 0 else trend_long[n_fine - 1]
        elif l1 < 0:
            v = trend_long[l2]
        elif l2 > n_fine - 1 or l1 == l2:
            v = trend_long[l1]
        else:
            w1, w2 = f32(math.ceil(dl) - dl), f32(dl - math.floor(dl))
            v = f32(w1 * f32(trend_long[l1])) + f32(w2 * f32(trend_long[l2]))
        out[k] = f32(v)
    return out


def trim_and_detrend(trace, z0_data, dz_data, shift=True):
    """FillInData's preparation of a trace for a parameter (not DATA) grid
    (commondata.cpp:6612-6644): where the first sample is zero, every zero sample is dropped and z0
    moves down one sample per leading zero (SEG-Y traces; shift=False: a STORM column's z0 stays,
    :6622-6628); then the line through the first and last samples is subtracted (float).  Returns
    (trace, z0, first, last) or None for an empty trace."""
    f32 = np.float32
    t = np.asarray(trace, dtype=np.float32)
    z0 = f32(z0_data)
    if len(t) and t[0] == 0:
        lead = 0
        while lead < len(t) and t[lead] == 0:
            lead += 1
        for _ in range(lead if shift else 0):
            z0 = f32(z0 + f32(dz_data))
        t = t[t != 0]
    n = len(t)
    if n == 0:
        return None
    first, last = f32(t[0]), f32(t[n - 1])
    inc = f32(f32(last - first) / f32(n - 1))
    k = np.arange(n).astype(np.float32)
    t = (t - (first + (k * inc).astype(np.float32)).astype(np.float32)).astype(np.float32)
    return t, z0, first, last


def fill_from_traces(traces, z0_data, dz_data, column_trace, z0_grid, dz_grid, nz, nzp, smooth_length,
                     parameter=False, trim_shift=True):
    """The columns of FillInData (commondata.cpp:6451-6775) from SEG-Y traces: column (j, i) of the
    padded real grid (nzp, nyp, rnxp) takes trace column_trace[j, i] with its first sample at
    z0_data[t] (ms); -1: outside the data (zeros), -2: a dead trace (zeros, RMISSING for a
    parameter grid).  Seismic (DATA) grids: tapered (smooth_trace_in_guard_zone), resampled
    (resample_trace with nt the closest factorable length of the longest trace, rounded to float
    as the reference's buffer is) and interpolated at z0_grid[j, i] + refk dz_grid[j, i]
    (interpolate_grid_values, dz_fine = dz_data / 10).  Parameter grids (parameter=True): trimmed
    and detrended (trim_and_detrend) instead of tapered, the trend resampled linearly ten times
    finer (res_fac (n - 1) + 1 samples) and added back (interpolate_and_shift_trend); an empty
    trace is dead; trim_shift=False: the trimmed zeros leave z0 (a STORM grid's columns).
    Returns (grid, nt)."""
    f32 = np.float32
    nyp, rnxp = column_trace.shape
    nt = lib().orc_factorable(max(len(t) for t in traces))
    dz_fine = f32(f32(dz_data) / f32(10.0))
    fine = {}
    out = np.zeros((nzp, nyp, rnxp), dtype=np.float32)
    for j in range(nyp):
        for i in range(rnxp):
            t = int(column_trace[j, i])
            if t == -2 and parameter:
                out[:, j, i] = f32(RMISSING)
            if t < 0:
                continue
            if t not in fine:
                if parameter:
                    r = trim_and_detrend(traces[t], z0_data[t], dz_data, trim_shift)
                    if r is None:
                        fine[t] = None
                    else:
                        tr, z0, first, last = r
                        n = len(tr)
                        nl = int(f32(f32(10.0) * f32(n - 1)) + f32(1))
                        inc2 = f32(f32(last - first) / f32(f32(10.0) * f32(n - 1)))
                        tl = (first + (np.arange(nl).astype(np.float32) * inc2).astype(np.float32)).astype(np.float32)
                        fine[t] = (resample_trace(tr, nt).astype(np.float32), z0, tl)
                else:
                    sm = smooth_trace_in_guard_zone(traces[t], dz_data, smooth_length)
                    fine[t] = (resample_trace(sm, nt).astype(np.float32), f32(z0_data[t]), None)
            if fine[t] is None:
                out[:, j, i] = f32(RMISSING)
                continue
            fn, z0, tl = fine[t]
            col = interpolate_grid_values(fn, z0_grid[j, i], dz_grid[j, i], z0, dz_fine, nz, nzp)
            if tl is not None:
                col = (col + interpolate_and_shift_trend(tl, z0_grid[j, i], dz_grid[j, i], z0, dz_fine, nz,
                                                         nzp)).astype(np.float32)
            out[:, j, i] = col
    return out, nt
