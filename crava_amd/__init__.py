"""crava_amd -- MI355X-native FFT-domain Bayesian AVO inversion (CRAVA's hot path).

Python mirror of the reference's interface for the path, over the C ABI of
include/cravagpu.h (libcravagpu.so, hand-written HIP for gfx950):

* ``FFTGrid``   -- src/fftgrid.h:22-285: padded cube in the FFTW in-place layout, fftInPlace /
                   invFFTInPlace with cube-type scaling, complex accessors, add,
                   multiplyByScalar, fillInComplexNoise (counter-based generator).
* ``posterior`` -- AVOInversion::computePostMeanResidAndFFTCov, per-wavenumber part
                   (src/avoinversion.cpp:944-1083).
* ``simulate``  -- one realisation of AVOInversion::simulate (src/avoinversion.cpp:1401-1511).
* ``build_kz_tables`` -- per-k_z tables of the constant-thickness branch (avoinversion.cpp:944-975).

The library is required: importing the GPU entry points without it, or on a machine
without a GPU, raises.  There is no CPU fallback.
"""
import ctypes
import os
import weakref

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcravagpu.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "cravagpu.h")

CG_OK, CG_EINVAL, CG_ENOMEM, CG_EHIP, CG_ESTATE, CG_EIO = 0, -1, -2, -3, -4, -5
DATA, PARAMETER, COVARIANCE = 1, 2, 3
PRIOR_GRIDS, PRIOR_SEPARABLE = 0, 1
PREC_REFERENCE, PREC_MIXED = 0, 1
NOISE_COUNTER, NOISE_GRIDS = 0, 1
K_FFT_X, K_FFT_Y, K_FFT_Z, K_POSTERIOR, K_NOISE, K_ELEMENTWISE = range(6)
FFT_AUTO, FFT_GENERIC = 0, 1
RMISSING = -99999.0

_vp = ctypes.c_void_p
_f_p = ctypes.POINTER(ctypes.c_float)
_d_p = ctypes.POINTER(ctypes.c_double)
_i_p = ctypes.POINTER(ctypes.c_int)
_u64_p = ctypes.POINTER(ctypes.c_uint64)


class PostArgs(ctypes.Structure):
    _fields_ = [
        ("ntheta", ctypes.c_int),
        ("K", _f_p), ("e1", _f_p), ("e2", _f_p), ("e3", _f_p), ("invert", _i_p),
        ("sigma0e", _d_p), ("wnc", ctypes.c_float), ("prior_var0", _d_p), ("prior_mode", ctypes.c_int),
        ("cov", _vp * 6), ("rxy", _f_p), ("rt", _f_p), ("err_corr", _vp), ("lam2d", _f_p),
        ("mean", _vp * 3), ("data", _vp * 8), ("want_residual", ctypes.c_int), ("precision", ctypes.c_int),
        ("cond_limit", ctypes.c_double), ("stats", _u64_p),
    ]


class SimArgs(ctypes.Structure):
    _fields_ = [
        ("cov", _vp * 6), ("mean", _vp * 3), ("out", _vp * 3), ("noise_source", ctypes.c_int),
        ("seed", ctypes.c_uint64), ("realisation", ctypes.c_uint32), ("stats", _u64_p),
    ]


class ScaleWaveletArgs(ctypes.Structure):
    _fields_ = [
        ("mode", ctypes.c_int), ("dt_ms", ctypes.c_double), ("wavelet", _f_p), ("rel_thick", _d_p),
        ("shift_ms", _f_p), ("gain", _f_p), ("stretch", _f_p), ("low_cut", ctypes.c_float),
        ("high_cut", ctypes.c_float), ("wnc", ctypes.c_float), ("error_var", ctypes.c_float), ("corr_t", _f_p),
        ("A", _f_p), ("prior_var0", _d_p),
    ]


SCALE_WAVELET_DIVIDE, SCALE_WAVELET_MULTIPLY = 0, 1

_lib = None


def lib():
    """Load libcravagpu.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("crava_amd: %s is missing; build it with `make -C crava_amd` "
                               "(or __graft_entry__.build())" % LIB_PATH)
        L = ctypes.CDLL(LIB_PATH)
        i, u64, f, d = ctypes.c_int, ctypes.c_uint64, ctypes.c_float, ctypes.c_double
        sig = {
            "cg_ctx_create": [ctypes.POINTER(_vp), i, u64],
            "cg_ctx_destroy": [_vp], "cg_ctx_sync": [_vp],
            "cg_ctx_memory": [_vp, _u64_p, _u64_p], "cg_ctx_stream": [_vp, ctypes.POINTER(_vp)],
            "cg_ctx_set_profiling": [_vp, i], "cg_ctx_kernel_time": [_vp, i, _u64_p, _d_p],
            "cg_ctx_reset_kernel_time": [_vp],
            "cg_grid_create": [_vp, i, i, i, i, i, i, i, ctypes.POINTER(_vp)],
            "cg_grid_destroy": [_vp], "cg_grid_set_type": [_vp, i], "cg_grid_is_transformed": [_vp, _i_p],
            "cg_grid_dims": [_vp, _i_p], "cg_grid_device_ptr": [_vp, ctypes.POINTER(_vp)],
            "cg_grid_upload_real": [_vp, _f_p], "cg_grid_upload_complex": [_vp, _f_p],
            "cg_grid_download_real": [_vp, _f_p], "cg_grid_download_complex": [_vp, _f_p],
            "cg_grid_get_complex": [_vp, i, i, i, i, _f_p], "cg_grid_set_complex": [_vp, i, i, i, i, _f_p],
            "cg_grid_get_real_value": [_vp, i, i, i, i, _f_p], "cg_grid_get_real_trace": [_vp, i, i, i, _f_p],
            "cg_fft_r2c_inplace": [_vp], "cg_fft_c2r_inplace": [_vp],
            "cg_grid_add": [_vp, _vp], "cg_grid_multiply_by_scalar": [_vp, f], "cg_grid_fill_constant": [_vp, f],
            "cg_posterior": [_vp, ctypes.POINTER(PostArgs)], "cg_simulate": [_vp, ctypes.POINTER(SimArgs)],
            "cg_posterior_noalloc": [_vp, ctypes.POINTER(PostArgs)],
            "cg_simulate_pair": [_vp, ctypes.POINTER(SimArgs), _vp * 3],
            "cg_grid_fill_noise_counter": [_vp, u64, ctypes.c_uint32],
            "cg_grid_fill_noise_reference": [_vp, ctypes.POINTER(ctypes.c_uint32)],
            "cg_build_kz_tables": [i, i, d, i, _f_p, _f_p, f, f, f, _f_p, _f_p, _f_p, _f_p, _i_p],
            "cg_build_kz_tables_variable_thickness": [i, i, d, i, _f_p, _f_p, f, f, f, _f_p, _f_p, _f_p, _f_p, _i_p],
            "cg_factorable": [i],
            "cg_make_circ_auto_cov_pos_def": [i, _f_p],
            "cg_fft_r2c": [_vp, _vp],
            "cg_forward_model": [_vp, i, _f_p, _vp * 3, _vp * 8],
            "cg_grid_add_white_noise": [_vp, f, u64, ctypes.c_uint32],
            "cg_grid_sum_squares": [_vp, i, _d_p], "cg_grid_data_variance": [_vp, _f_p, _u64_p],
            "cg_ctx_set_fft_engine": [_vp, i], "cg_ctx_set_fft_variant": [_vp, i],
            "cg_ctx_set_fused_colour": [_vp, i],
            "cg_autotune": [_vp, i, i, i],
            "cg_arena_create": [_vp, u64, ctypes.POINTER(_vp)], "cg_arena_destroy": [_vp],
            "cg_grid_create_in": [_vp, i, i, i, i, i, i, i, ctypes.POINTER(_vp)],
            "cg_grid_fill_padded": [_vp, _vp, i],
            "cg_grid_interpolate_seismic": [_vp, f, _i_p, _i_p],
            "cg_grid_extrapolate_seismic": [_vp, i, i, i, i],
            "cg_grid_fill_param_corr": [_vp, _f_p, _f_p, f, f, f], "cg_grid_fill_trace_taper": [_vp, _vp, i], "cg_grid_fill_err_corr": [_vp, _f_p, f, f],
            "cg_grid_write_crava": [_vp, ctypes.c_char_p, ctypes.POINTER(CravaGeometry)],
            "cg_grid_scale_wavelet": [_vp, ctypes.POINTER(ScaleWaveletArgs)],
            "cg_grid_write_storm": [_vp, ctypes.c_char_p, ctypes.POINTER(StormGeometry), i],
            "cg_grid_read_storm": [_vp, ctypes.c_char_p, ctypes.POINTER(StormGeometry)],
            "cg_grid_write_segy": [_vp, ctypes.c_char_p, ctypes.POINTER(SegyGeometry)],
            "cg_grid_fill_from_traces": [_vp, ctypes.POINTER(TraceFillArgs)],
            "cg_grid_read_crava": [_vp, ctypes.c_char_p, ctypes.POINTER(CravaGeometry)],
            "cg_local_noise_create": [_vp, ctypes.POINTER(LocalNoiseArgs), ctypes.POINTER(_vp)],
            "cg_local_noise_matrices": [_vp, _d_p, _d_p],
            "cg_local_noise_correct": [_vp, _vp * 3, _vp * 3],
            "cg_local_noise_correct_realisation": [_vp, _vp * 3, _vp * 3],
        }
        for name, args in sig.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = ctypes.c_int
        L.cg_local_noise_destroy.argtypes = [_vp]
        L.cg_local_noise_destroy.restype = None
        L.cg_arena_grid_bytes.argtypes = [i, i, i]
        L.cg_arena_grid_bytes.restype = ctypes.c_uint64
        L.cg_last_error.argtypes = [_vp]
        L.cg_last_error.restype = ctypes.c_char_p
        _lib = L
    return _lib


class CravaError(RuntimeError):
    pass


class StormGeometry(ctypes.Structure):
    """cg_storm_geometry: the simbox volume of a STORM grid header."""
    _fields_ = [("x_min", ctypes.c_double), ("lx", ctypes.c_double), ("y_min", ctypes.c_double),
                ("ly", ctypes.c_double), ("lz", ctypes.c_double), ("angle", ctypes.c_double),
                ("top_name", ctypes.c_char_p), ("bot_name", ctypes.c_char_p)]


class LocalNoiseArgs(ctypes.Structure):
    """cg_local_noise_args."""
    _fields_ = [("nx", ctypes.c_int), ("ny", ctypes.c_int), ("ntheta", ctypes.c_int),
                ("G", ctypes.POINTER(ctypes.c_double)), ("prior_var0", ctypes.POINTER(ctypes.c_double)),
                ("post_var0", ctypes.POINTER(ctypes.c_double)), ("err_theta_cov", ctypes.POINTER(ctypes.c_double)),
                ("noise_scale", ctypes.POINTER(ctypes.POINTER(ctypes.c_double)))]


class TraceFillArgs(ctypes.Structure):
    _fields_ = [("samples", ctypes.POINTER(ctypes.c_float)), ("length", ctypes.POINTER(ctypes.c_int)),
                ("z0_data", ctypes.POINTER(ctypes.c_float)), ("ntraces", ctypes.c_int), ("stride", ctypes.c_int),
                ("dz_data", ctypes.c_float), ("smooth_length", ctypes.c_float),
                ("column_trace", ctypes.POINTER(ctypes.c_int)), ("z0_grid", ctypes.POINTER(ctypes.c_float)),
                ("dz_grid", ctypes.POINTER(ctypes.c_float)), ("grid_type", ctypes.c_int),
                ("trim_keeps_z0", ctypes.c_int)]


class SegyGeometry(ctypes.Structure):
    """cg_segy_geometry: the simbox geometry FFTGrid::writeSegyFile reads."""
    _fields_ = [("x0", ctypes.c_double), ("y0", ctypes.c_double), ("dx", ctypes.c_double), ("dy", ctypes.c_double),
                ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("angle", ctypes.c_double),
                ("il0", ctypes.c_double), ("xl0", ctypes.c_double), ("il_step_x", ctypes.c_double),
                ("il_step_y", ctypes.c_double), ("xl_step_x", ctypes.c_double), ("xl_step_y", ctypes.c_double),
                ("dz", ctypes.c_double), ("z0", ctypes.c_float), ("zmax", ctypes.c_double),
                ("top", ctypes.POINTER(ctypes.c_double)), ("rel_thick", ctypes.POINTER(ctypes.c_double)),
                ("text", ctypes.POINTER(ctypes.c_char_p)), ("ntext", ctypes.c_int)]


class CravaGeometry(ctypes.Structure):
    """cg_crava_geometry: the survey geometry in a CRAVA binary grid file's header."""
    _fields_ = [("x0", ctypes.c_double), ("y0", ctypes.c_double), ("dx", ctypes.c_double), ("dy", ctypes.c_double),
                ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("il0", ctypes.c_double), ("xl0", ctypes.c_double),
                ("il_step_x", ctypes.c_double), ("il_step_y", ctypes.c_double), ("xl_step_x", ctypes.c_double),
                ("xl_step_y", ctypes.c_double), ("angle", ctypes.c_double)]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


def _ptr(a, ctype):
    return None if a is None else a.ctypes.data_as(ctypes.POINTER(ctype))


class Context:
    """One device, one HIP stream, an HBM budget (0 = unlimited)."""

    def __init__(self, device=0, hbm_budget=0):
        self._h = _vp()
        self._grids = weakref.WeakSet()
        self._batches = weakref.WeakSet()
        self._arenas = weakref.WeakSet()
        rc = lib().cg_ctx_create(ctypes.byref(self._h), device, hbm_budget)
        if rc != CG_OK:
            raise CravaError("cg_ctx_create failed (%d): no usable GPU %d" % (rc, device))

    def check(self, rc, what=""):
        if rc != CG_OK:
            raise CravaError("%s failed (%d): %s" % (what, rc, lib().cg_last_error(self._h).decode()))
        return rc

    def sync(self):
        self.check(lib().cg_ctx_sync(self._h), "cg_ctx_sync")

    def memory(self):
        a, b = ctypes.c_uint64(), ctypes.c_uint64()
        lib().cg_ctx_memory(self._h, ctypes.byref(a), ctypes.byref(b))
        return a.value, b.value

    def autotune(self, nxp, nyp, nzp):
        """Time both tile variants of every FFT pass of an nxp x nyp x nzp grid and keep the
        faster for that geometry (every context of the device)."""
        self.check(lib().cg_autotune(self._h, nxp, nyp, nzp), "cg_autotune")

    def set_fft_variant(self, variant):
        """Force tile variant 0 or 1 of the compile-time-plan passes, or -1 for the tuned choice."""
        self.check(lib().cg_ctx_set_fft_variant(self._h, variant), "cg_ctx_set_fft_variant")

    def set_fused_colour(self, mode):
        """Realisations with counter noise: -1 automatic (autotuned, else by occupancy), 0 the
        colouring kernel then three z passes, 1 colouring fused into the inverse z pass."""
        self.check(lib().cg_ctx_set_fused_colour(self._h, mode), "cg_ctx_set_fused_colour")

    def set_fft_engine(self, engine):
        """FFT_AUTO (compile-time plans where available) or FFT_GENERIC (runtime plans only)."""
        self.check(lib().cg_ctx_set_fft_engine(self._h, engine), "set_fft_engine")

    def set_profiling(self, on):
        self.check(lib().cg_ctx_set_profiling(self._h, 1 if on else 0), "cg_ctx_set_profiling")

    def kernel_time(self, kclass):
        n, ms = ctypes.c_uint64(), ctypes.c_double()
        self.check(lib().cg_ctx_kernel_time(self._h, kclass, ctypes.byref(n), ctypes.byref(ms)), "kernel_time")
        return n.value, ms.value

    def reset_kernel_time(self):
        self.check(lib().cg_ctx_reset_kernel_time(self._h), "reset_kernel_time")

    def close(self):
        if self._h:
            for b in list(self._batches):
                b.close()
            for g in list(self._grids):
                g.close()
            for a in list(self._arenas):
                a.close()
            lib().cg_ctx_destroy(self._h)
            self._h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


class Arena:
    """One device allocation that grids are carved from (cg_arena): a request's grids cost one
    allocation and one release.  Close every grid of the arena before the arena."""

    def __init__(self, ctx, nbytes):
        self.ctx = ctx
        self._h = _vp()
        ctx.check(lib().cg_arena_create(ctx._h, int(nbytes), ctypes.byref(self._h)), "cg_arena_create")
        ctx._arenas.add(self)

    @staticmethod
    def grid_bytes(nxp, nyp, nzp):
        return int(lib().cg_arena_grid_bytes(nxp, nyp, nzp))

    def close(self):
        if self._h and self.ctx._h:
            self.ctx.check(lib().cg_arena_destroy(self._h), "cg_arena_destroy")
        self._h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class FFTGrid:
    """Device-resident FFTGrid (src/fftgrid.h:22-285) in the reference's in-place layout."""
    DATA, PARAMETER, COVARIANCE = DATA, PARAMETER, COVARIANCE

    def __init__(self, ctx, nx, ny, nz, nxp, nyp, nzp, cubetype=PARAMETER, arena=None):
        self.ctx = ctx
        self.nx, self.ny, self.nz, self.nxp, self.nyp, self.nzp = nx, ny, nz, nxp, nyp, nzp
        self.cnxp = nxp // 2 + 1
        self.rnxp = 2 * self.cnxp
        self._h = _vp()
        if arena is None:
            ctx.check(lib().cg_grid_create(ctx._h, nx, ny, nz, nxp, nyp, nzp, cubetype, ctypes.byref(self._h)),
                      "cg_grid_create")
        else:
            ctx.check(lib().cg_grid_create_in(arena._h, nx, ny, nz, nxp, nyp, nzp, cubetype, ctypes.byref(self._h)),
                      "cg_grid_create_in")
        ctx._grids.add(self)

    # -- state
    def setType(self, cubetype):
        self.ctx.check(lib().cg_grid_set_type(self._h, cubetype), "setType")

    def getIsTransformed(self):
        t = ctypes.c_int()
        lib().cg_grid_is_transformed(self._h, ctypes.byref(t))
        return bool(t.value)

    def device_ptr(self):
        p = _vp()
        lib().cg_grid_device_ptr(self._h, ctypes.byref(p))
        return p.value

    # -- transforms and elementwise ops
    def fftInPlace(self):
        self.ctx.check(lib().cg_fft_r2c_inplace(self._h), "fftInPlace")

    def invFFTInPlace(self):
        self.ctx.check(lib().cg_fft_c2r_inplace(self._h), "invFFTInPlace")

    def fft_from(self, src):
        """Out-of-place forward transform: this grid receives src's spectrum, src is kept."""
        self.ctx.check(lib().cg_fft_r2c(src._h, self._h), "cg_fft_r2c")

    def add_white_noise(self, sigma, seed, stream=0):
        self.ctx.check(lib().cg_grid_add_white_noise(self._h, sigma, seed, stream), "add_white_noise")

    def sum_squares(self, extSimbox=False):
        v = ctypes.c_double()
        self.ctx.check(lib().cg_grid_sum_squares(self._h, int(extSimbox), ctypes.byref(v)), "sum_squares")
        return v.value

    def data_variance(self):
        """ModelAVODynamic::ComputeDataVariance for this stack: (variance as float, non-zero samples)."""
        v, n = ctypes.c_float(), ctypes.c_uint64()
        self.ctx.check(lib().cg_grid_data_variance(self._h, ctypes.byref(v), ctypes.byref(n)), "data_variance")
        return v.value, n.value

    def add(self, other):
        self.ctx.check(lib().cg_grid_add(self._h, other._h), "add")

    def multiplyByScalar(self, s):
        self.ctx.check(lib().cg_grid_multiply_by_scalar(self._h, s), "multiplyByScalar")

    def fillInConstant(self, value):
        """FFTGrid::fillInConstant: value on the cube, RMISSING on the row padding (real domain)."""
        self.ctx.check(lib().cg_grid_fill_constant(self._h, value), "fillInConstant")

    def fill_padded(self, cube):
        """FFTGrid(const NRLib::Grid<float> *, nxp, nyp, nzp): cube (nz, ny, nx), a host array, or
        an integer device address of nz*ny*nx floats; the padding is filled by linear edge
        interpolation and the grid becomes a PARAMETER grid in the real domain."""
        if isinstance(cube, int):
            self.ctx.check(lib().cg_grid_fill_padded(self._h, ctypes.c_void_p(cube), 1), "fill_padded")
            return
        a = np.ascontiguousarray(cube, dtype=np.float32)
        if a.shape != (self.nz, self.ny, self.nx):
            raise ValueError("cube must be (nz, ny, nx) = %s" % ((self.nz, self.ny, self.nx),))
        self.ctx.check(lib().cg_grid_fill_padded(self._h, a.ctypes.data_as(_vp), 0), "fill_padded")

    def getRealValue(self, i, j, k, extSimbox=False):
        v = ctypes.c_float()
        self.ctx.check(lib().cg_grid_get_real_value(self._h, i, j, k, 1 if extSimbox else 0, ctypes.byref(v)),
                       "getRealValue")
        return v.value

    def getRealTrace(self, i, j, add_padding=False):
        out = np.empty(self.nzp if add_padding else self.nz, dtype=np.float32)
        self.ctx.check(lib().cg_grid_get_real_trace(self._h, i, j, 1 if add_padding else 0, _ptr(out, ctypes.c_float)),
                       "getRealTrace")
        return out

    def fill_trace_taper(self, cube):
        """ComputeSyntSeismic's trace taper: cube (nz, ny, nx) host array, or a device address."""
        if isinstance(cube, int):
            self.ctx.check(lib().cg_grid_fill_trace_taper(self._h, ctypes.c_void_p(cube), 1), "fill_trace_taper")
            return
        a = np.ascontiguousarray(cube, dtype=np.float32)
        if a.shape != (self.nz, self.ny, self.nx):
            raise ValueError("cube must be (nz, ny, nx) = %s" % ((self.nz, self.ny, self.nx),))
        self.ctx.check(lib().cg_grid_fill_trace_taper(self._h, a.ctypes.data_as(_vp), 0), "fill_trace_taper")

    def fillInParamCorr(self, corr_xy, corr_t, grad_i=0.0, grad_j=0.0, scale=1.0):
        """FFTGrid::fillInParamCorr / FillInLateralCorr, then multiplyByScalar(scale): corr_xy
        (nyp, nxp), corr_t (nzp,)."""
        cxy = np.ascontiguousarray(corr_xy, dtype=np.float32)
        ct = np.ascontiguousarray(corr_t, dtype=np.float32)
        if cxy.shape != (self.nyp, self.nxp) or ct.shape != (self.nzp,):
            raise ValueError("corr_xy must be (nyp, nxp) and corr_t (nzp,)")
        self.ctx.check(lib().cg_grid_fill_param_corr(self._h, _ptr(cxy, ctypes.c_float), _ptr(ct, ctypes.c_float),
                                                     grad_i, grad_j, scale), "fillInParamCorr")

    def fillInErrCorr(self, corr_xy, grad_i=0.0, grad_j=0.0):
        """FFTGrid::fillInErrCorr: corr_xy (nyp, nxp)."""
        cxy = np.ascontiguousarray(corr_xy, dtype=np.float32)
        if cxy.shape != (self.nyp, self.nxp):
            raise ValueError("corr_xy must be (nyp, nxp)")
        self.ctx.check(lib().cg_grid_fill_err_corr(self._h, _ptr(cxy, ctypes.c_float), grad_i, grad_j),
                       "fillInErrCorr")

    def scale_wavelet(self, mode, dt_ms, wavelet, rel_thick, low_cut=5.0, high_cut=55.0, wnc=0.1, error_var=1.0,
                      corr_t=None, A=None, prior_var0=None, shift_ms=None, gain=None, stretch=None):
        """The non-constant-thickness data rescaling on this real-domain grid, trace by trace:
        SCALE_WAVELET_DIVIDE (divideDataByScaleWavelet with computeAdjustmentFactor) or
        SCALE_WAVELET_MULTIPLY (multiplyDataByScaleWaveletAndWriteToFile before the write).
        wavelet (nzp,) in FFT order; rel_thick (nyp, nxp); shift_ms, gain, stretch (ny, nx) or None;
        corr_t (nzp,) the filtered prior temporal correlation, A (3,), prior_var0 (3, 3) (divide)."""
        a = ScaleWaveletArgs()
        keep = []

        def arr(x, dt, shape=None):
            if x is None:
                return None
            y = np.ascontiguousarray(x, dtype=dt)
            if shape is not None and y.shape != shape:
                raise ValueError("expected shape %s, got %s" % (shape, y.shape))
            keep.append(y)
            return y

        a.mode = mode
        a.dt_ms = dt_ms
        a.wavelet = _ptr(arr(wavelet, np.float32, (self.nzp,)), ctypes.c_float)
        a.rel_thick = _ptr(arr(rel_thick, np.float64, (self.nyp, self.nxp)), ctypes.c_double)
        a.shift_ms = _ptr(arr(shift_ms, np.float32, (self.ny, self.nx)), ctypes.c_float)
        a.gain = _ptr(arr(gain, np.float32, (self.ny, self.nx)), ctypes.c_float)
        a.stretch = _ptr(arr(stretch, np.float32, (self.ny, self.nx)), ctypes.c_float)
        a.low_cut, a.high_cut, a.wnc, a.error_var = low_cut, high_cut, wnc, error_var
        a.corr_t = _ptr(arr(corr_t, np.float32, (self.nzp,)), ctypes.c_float)
        a.A = _ptr(arr(A, np.float32, (3,)), ctypes.c_float)
        a.prior_var0 = _ptr(arr(prior_var0, np.float64, (3, 3)), ctypes.c_double)
        self.ctx.check(lib().cg_grid_scale_wavelet(self._h, ctypes.byref(a)), "scale_wavelet")

    def interpolateSeismic(self, energy_threshold):
        """FFTGrid::interpolateSeismic; returns (bad traces, (imin, imax, jmin, jmax))."""
        nbad, rect = ctypes.c_int(), (ctypes.c_int * 4)()
        self.ctx.check(lib().cg_grid_interpolate_seismic(self._h, energy_threshold, ctypes.byref(nbad), rect),
                       "interpolateSeismic")
        return nbad.value, tuple(rect)

    def extrapolateSeismic(self, imin, imax, jmin, jmax):
        """FFTGrid::extrapolateSeismic with the rectangle of good box traces."""
        self.ctx.check(lib().cg_grid_extrapolate_seismic(self._h, imin, imax, jmin, jmax), "extrapolateSeismic")

    def writeCravaFile(self, file_name, geometry, suffix=".crava"):
        """FFTGrid::writeCravaFile: file_name + suffix (IO::SuffixCrava); geometry: a dict of the
        CravaGeometry fields."""
        geo = CravaGeometry(**geometry)
        self.ctx.check(lib().cg_grid_write_crava(self._h, (file_name + suffix).encode(), ctypes.byref(geo)),
                       "writeCravaFile")

    def writeStormFile(self, file_name, geometry, padding=False, suffix=".storm"):
        """FFTGrid::writeStormFile, binary: file_name + suffix (IO::SuffixStormBinary); geometry: a
        dict of x_min, lx, y_min, ly, lz, angle (radians) and optionally top_name, bot_name."""
        g = StormGeometry(geometry["x_min"], geometry["lx"], geometry["y_min"], geometry["ly"], geometry["lz"],
                          geometry["angle"], (geometry.get("top_name") or "").encode() or None,
                          (geometry.get("bot_name") or "").encode() or None)
        self.ctx.check(lib().cg_grid_write_storm(self._h, (file_name + suffix).encode(), ctypes.byref(g),
                                                 1 if padding else 0), "writeStormFile")

    def writeSegyFile(self, file_name, geometry):
        """FFTGrid::writeSegyFile (fftgrid.cpp:1960-2050) to file_name as given; geometry: a dict of
        x0, y0, dx, dy, angle, il0, xl0, il_step_x, il_step_y, xl_step_x, xl_step_y, dz, z0, zmax,
        top (ny, nx), rel_thick (ny, nx) and optionally text (header lines from the second on)."""
        g = geometry
        top = np.ascontiguousarray(g["top"], dtype=np.float64)
        rel = np.ascontiguousarray(g["rel_thick"], dtype=np.float64)
        if top.shape != (self.ny, self.nx) or rel.shape != (self.ny, self.nx):
            raise ValueError("writeSegyFile: top and rel_thick must be (ny, nx) = (%d, %d)" % (self.ny, self.nx))
        text = [t.encode() for t in g.get("text", [])]
        tarr = (ctypes.c_char_p * max(len(text), 1))(*text)
        geo = SegyGeometry(g["x0"], g["y0"], g["dx"], g["dy"], self.nx, self.ny, g["angle"], g["il0"], g["xl0"],
                           g["il_step_x"], g["il_step_y"], g["xl_step_x"], g["xl_step_y"], g["dz"], float(g["z0"]),
                           g["zmax"], _ptr(top, ctypes.c_double), _ptr(rel, ctypes.c_double), tarr, len(text))
        self.ctx.check(lib().cg_grid_write_segy(self._h, file_name.encode(), ctypes.byref(geo)), "writeSegyFile")

    def fillInData(self, traces, z0_data, dz_data, column_trace, z0_grid, dz_grid, smooth_length,
                   grid_type=DATA, trim_keeps_z0=False):
        """CommonData::FillInData (commondata.cpp:6451-6775) from SEG-Y traces, with the columns
        resolved by the caller: traces (a list of 1-D arrays, at most 1024 samples), z0_data (time
        of each trace's first sample, ms), dz_data (ms), column_trace (nyp, rnxp) trace index per
        column, -1 outside the data, -2 a dead trace, z0_grid / dz_grid (nyp, rnxp) the centre of
        each column's first cell and its thickness (ms), smooth_length the guard-zone taper (ms,
        DATA grids).  grid_type DATA: the seismic branch (tapered traces, zeros where no trace);
        PARAMETER: the branch for other grids (zero samples trimmed, the linear trend removed and
        added back, RMISSING in dead columns; every box column must have a trace or be dead);
        trim_keeps_z0: the trimmed zeros do not move z0_data (traces from a STORM grid)."""
        n = len(traces)
        stride = max(len(t) for t in traces)
        samples = np.zeros((n, stride), dtype=np.float32)
        for t, tr in enumerate(traces):
            samples[t, :len(tr)] = np.asarray(tr, dtype=np.float32)
        length = np.array([len(t) for t in traces], dtype=np.int32)
        z0d = np.ascontiguousarray(z0_data, dtype=np.float32)
        shape = (self.nyp, self.rnxp)
        ct = np.ascontiguousarray(column_trace, dtype=np.int32)
        z0g = np.ascontiguousarray(z0_grid, dtype=np.float32)
        dzg = np.ascontiguousarray(dz_grid, dtype=np.float32)
        if ct.shape != shape or z0g.shape != shape or dzg.shape != shape or z0d.shape != (n,):
            raise ValueError("fillInData: columns must be (nyp, rnxp) = %s, z0_data (%d,)" % (shape, n))
        a = TraceFillArgs(_ptr(samples, ctypes.c_float), _ptr(length, ctypes.c_int), _ptr(z0d, ctypes.c_float), n,
                          stride, float(dz_data), float(smooth_length), _ptr(ct, ctypes.c_int),
                          _ptr(z0g, ctypes.c_float), _ptr(dzg, ctypes.c_float), int(grid_type),
                          1 if trim_keeps_z0 else 0)
        self.ctx.check(lib().cg_grid_fill_from_traces(self._h, ctypes.byref(a)), "fillInData")

    def readStormFile(self, file_name):
        """A STORM binary grid of this grid's box size, then the padding fill of
        FFTGrid(const NRLib::Grid *); returns the header's volume as a dict."""
        g = StormGeometry()
        self.ctx.check(lib().cg_grid_read_storm(self._h, file_name.encode(), ctypes.byref(g)), "readStormFile")
        return dict(x_min=g.x_min, lx=g.lx, y_min=g.y_min, ly=g.ly, lz=g.lz, angle=g.angle)

    def readCravaFile(self, file_name):
        """FFTGrid::readCravaFile; returns the header's geometry as a dict."""
        geo = CravaGeometry()
        self.ctx.check(lib().cg_grid_read_crava(self._h, file_name.encode(), ctypes.byref(geo)), "readCravaFile")
        return geo.as_dict()

    def fillInComplexNoiseReference(self, seed):
        """FFTGrid::fillInComplexNoise with the reference's RandomGen stream from state `seed`;
        returns the state after the fill (the reference's seed_ after the call)."""
        st = ctypes.c_uint32(seed)
        self.ctx.check(lib().cg_grid_fill_noise_reference(self._h, ctypes.byref(st)), "fillInComplexNoiseReference")
        return st.value

    def fillInComplexNoise(self, seed, stream=0):
        self.ctx.check(lib().cg_grid_fill_noise_counter(self._h, seed, stream), "fillInComplexNoise")

    # -- host mirror
    def upload_real(self, real):
        """real: (nzp, nyp, nxp) or padded (nzp, nyp, rnxp) float array."""
        a = np.asarray(real, dtype=np.float32)
        if a.shape[-1] == self.nxp and self.nxp != self.rnxp:
            buf = np.zeros((self.nzp, self.nyp, self.rnxp), dtype=np.float32)
            buf[:, :, :self.nxp] = a
        else:
            buf = np.ascontiguousarray(a)
        assert buf.size == self.rnxp * self.nyp * self.nzp
        self.ctx.check(lib().cg_grid_upload_real(self._h, _ptr(buf, ctypes.c_float)), "upload_real")

    def upload_complex(self, spec):
        a = np.ascontiguousarray(spec, dtype=np.complex64)
        assert a.size == self.cnxp * self.nyp * self.nzp
        self.ctx.check(lib().cg_grid_upload_complex(self._h, _ptr(a.view(np.float32), ctypes.c_float)),
                       "upload_complex")

    def download_real(self, padded=False):
        buf = np.empty((self.nzp, self.nyp, self.rnxp), dtype=np.float32)
        self.ctx.check(lib().cg_grid_download_real(self._h, _ptr(buf, ctypes.c_float)), "download_real")
        return buf if padded else buf[:, :, :self.nxp].copy()

    def download_complex(self):
        buf = np.empty((self.nzp, self.nyp, self.cnxp), dtype=np.complex64)
        self.ctx.check(lib().cg_grid_download_complex(self._h, _ptr(buf.view(np.float32), ctypes.c_float)),
                       "download_complex")
        return buf

    def getComplexValue(self, i, j, k, extSimbox=False):
        v = np.zeros(2, dtype=np.float32)
        self.ctx.check(lib().cg_grid_get_complex(self._h, i, j, k, int(extSimbox), _ptr(v, ctypes.c_float)),
                       "getComplexValue")
        return complex(v[0], v[1])

    def setComplexValue(self, i, j, k, value, extSimbox=False):
        v = np.array([value.real, value.imag], dtype=np.float32)
        rc = lib().cg_grid_set_complex(self._h, i, j, k, int(extSimbox), _ptr(v, ctypes.c_float))
        if rc < 0:
            self.ctx.check(rc, "setComplexValue")
        return rc

    def close(self):
        # a grid outlived by its context was released with it (cg_ctx_destroy)
        if self._h and self.ctx._h:
            lib().cg_grid_destroy(self._h)
        self._h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def build_kz_tables(nz, nzp, dt_ms, wavelets, A, low_cut=5.0, high_cut=55.0, min_rel_thick=1.0,
                    constant_thickness=True):
    """Host-side per-k_z tables (K, e1, e2, e3, invert) of avoinversion.cpp:944-975, or with
    constant_thickness=False those of the non-constant-thickness branch (:976-1014)."""
    w = np.ascontiguousarray(wavelets, dtype=np.float32)
    nt = w.shape[0]
    a = np.ascontiguousarray(A, dtype=np.float32)
    K = np.zeros((nzp, nt, 3), dtype=np.complex64)
    e1 = np.zeros((nzp, nt), dtype=np.complex64)
    e2 = np.zeros((nzp, nt), dtype=np.complex64)
    e3 = np.zeros((nzp, nt), dtype=np.complex64)
    inv = np.zeros(nzp, dtype=np.int32)
    f = ctypes.c_float
    fn = lib().cg_build_kz_tables if constant_thickness else lib().cg_build_kz_tables_variable_thickness
    rc = fn(nz, nzp, dt_ms, nt, _ptr(w, f), _ptr(a, f), low_cut, high_cut, min_rel_thick,
                                  _ptr(K.view(np.float32), f), _ptr(e1.view(np.float32), f),
                                  _ptr(e2.view(np.float32), f), _ptr(e3.view(np.float32), f),
                                  _ptr(inv, ctypes.c_int))
    if rc != CG_OK:
        raise CravaError("cg_build_kz_tables failed (%d)" % rc)
    return K, e1, e2, e3, inv


def forward_model(ctx, K, m, d):
    """d_l(k) = sum_j K_lj(kz) m_j(k) on Fourier-domain grids (synthetic seismic spectra)."""
    Kc = np.ascontiguousarray(K, dtype=np.complex64)
    mm = (_vp * 3)(*[g._h for g in m])
    dd = (_vp * 8)(*([g._h for g in d] + [None] * (8 - len(d))))
    ctx.check(lib().cg_forward_model(ctx._h, len(d), _ptr(Kc.view(np.float32), ctypes.c_float), mm, dd),
              "cg_forward_model")


def posterior(ctx, tables, mean, data, cov, prior_mode=PRIOR_SEPARABLE, rxy=None, rt=None, err_corr=None,
              lam2d=None, want_residual=True, precision=PREC_REFERENCE, cond_limit=1e6, stats=True, noalloc=False):
    """Per-wavenumber posterior on Fourier-domain FFTGrids (in place).

    tables: dict with K, e1, e2, e3, invert, sigma0e, wnc, prior_var0.  noalloc: the entry point
    that allocates no device memory (cg_posterior_noalloc).
    Returns (inverted, cholesky_failures, fp64_resolves)."""
    a = PostArgs()
    keep = []

    def arr(x, dt):
        y = np.ascontiguousarray(x, dtype=dt)
        keep.append(y)
        return y

    nt = len(data)
    a.ntheta = nt
    a.K = _ptr(arr(tables["K"], np.complex64).view(np.float32), ctypes.c_float)
    a.e1 = _ptr(arr(tables["e1"], np.complex64).view(np.float32), ctypes.c_float)
    a.e2 = _ptr(arr(tables["e2"], np.complex64).view(np.float32), ctypes.c_float)
    a.e3 = _ptr(arr(tables["e3"], np.complex64).view(np.float32), ctypes.c_float)
    a.invert = _ptr(arr(tables["invert"], np.int32), ctypes.c_int)
    a.sigma0e = _ptr(arr(tables["sigma0e"], np.float64), ctypes.c_double)
    a.wnc = float(tables["wnc"])
    a.prior_var0 = _ptr(arr(tables["prior_var0"], np.float64), ctypes.c_double)
    a.prior_mode = prior_mode
    for g in range(6):
        a.cov[g] = cov[g]._h
    if rxy is not None:
        a.rxy = _ptr(arr(rxy, np.float32), ctypes.c_float)
    if rt is not None:
        a.rt = _ptr(arr(rt, np.float32), ctypes.c_float)
    a.err_corr = err_corr._h if err_corr is not None else None
    if lam2d is not None:
        a.lam2d = _ptr(arr(lam2d, np.float32), ctypes.c_float)
    for g in range(3):
        a.mean[g] = mean[g]._h
    for l in range(nt):
        a.data[l] = data[l]._h
    a.want_residual = int(want_residual)
    a.precision = precision
    a.cond_limit = cond_limit
    want_stats = stats
    stats = np.zeros(3, dtype=np.uint64)
    # statistics cost a device-to-host copy and a stream synchronisation
    a.stats = _ptr(stats, ctypes.c_uint64) if want_stats else None
    if noalloc:
        ctx.check(lib().cg_posterior_noalloc(ctx._h, ctypes.byref(a)), "cg_posterior_noalloc")
    else:
        ctx.check(lib().cg_posterior(ctx._h, ctypes.byref(a)), "cg_posterior")
    return tuple(int(x) for x in stats)


def simulate(ctx, cov, mean, out, seed=0, realisation=0, noise_source=NOISE_COUNTER, stats=False, next_out=None):
    """One posterior realisation into out[3] (real domain), or with next_out two: realisations
    `realisation` (into out) and `realisation` + 1 (into next_out), coloured together where the
    fused colouring applies (cg_simulate_pair).  Returns Cholesky failures (of both)."""
    a = SimArgs()
    for g in range(6):
        a.cov[g] = cov[g]._h
    for g in range(3):
        a.mean[g] = mean[g]._h if mean is not None and mean[g] is not None else None
        a.out[g] = out[g]._h
    a.noise_source = noise_source
    a.seed = seed
    a.realisation = realisation
    st = np.zeros(1, dtype=np.uint64)
    a.stats = _ptr(st, ctypes.c_uint64) if stats else None
    if next_out is None:
        ctx.check(lib().cg_simulate(ctx._h, ctypes.byref(a)), "cg_simulate")
    else:
        nx = (_vp * 3)(*[g._h for g in next_out])
        ctx.check(lib().cg_simulate_pair(ctx._h, ctypes.byref(a), nx), "cg_simulate_pair")
    return int(st[0])


def header_symbols():
    """Function names declared in include/cravagpu.h."""
    import re
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(cg_\w+)\s*\(", txt, re.M)))


class InversionDesc(ctypes.Structure):
    _fields_ = [("post", PostArgs), ("mean_in", _vp * 3), ("seis_in", _vp * 8), ("nreal", ctypes.c_int),
                ("out", ctypes.POINTER(_vp)), ("seed", ctypes.c_uint64)]


def _post_args(tables, mean, data, cov, rxy, rt, lam2d, precision, cond_limit, want_residual, keep):
    a = PostArgs()

    def arr(x, dt):
        y = np.ascontiguousarray(x, dtype=dt)
        keep.append(y)
        return y

    a.ntheta = len(data)
    a.K = _ptr(arr(tables["K"], np.complex64).view(np.float32), ctypes.c_float)
    a.e1 = _ptr(arr(tables["e1"], np.complex64).view(np.float32), ctypes.c_float)
    a.e2 = _ptr(arr(tables["e2"], np.complex64).view(np.float32), ctypes.c_float)
    a.e3 = _ptr(arr(tables["e3"], np.complex64).view(np.float32), ctypes.c_float)
    a.invert = _ptr(arr(tables["invert"], np.int32), ctypes.c_int)
    a.sigma0e = _ptr(arr(tables["sigma0e"], np.float64), ctypes.c_double)
    a.wnc = float(tables["wnc"])
    a.prior_var0 = _ptr(arr(tables["prior_var0"], np.float64), ctypes.c_double)
    a.prior_mode = PRIOR_SEPARABLE
    for g in range(6):
        a.cov[g] = cov[g]._h
    a.rxy = _ptr(arr(rxy, np.float32), ctypes.c_float)
    a.rt = _ptr(arr(rt, np.float32), ctypes.c_float)
    a.lam2d = _ptr(arr(lam2d, np.float32), ctypes.c_float)
    for g in range(3):
        a.mean[g] = mean[g]._h
    for l in range(len(data)):
        a.data[l] = data[l]._h
    a.want_residual = int(want_residual)
    a.precision = precision
    a.cond_limit = cond_limit
    return a


def invert(ctx, item, precision=PREC_MIXED, cond_limit=1e6, want_residual=False, stats=False):
    """One whole inversion (cg_invert: AVOInversion::computePostMeanResidAndFFTCov followed by
    simulate): `item` as for Batch (tables, rxy, rt, lam2d, mean_in[3], seis_in[nt], mean[3],
    data[nt], cov[6], out, seed).  On return mean[] hold the real-domain posterior means, cov[]
    the covariance spectra, out[] the realisations.  With stats, returns (inverted, cholesky
    failures, FP64 re-solves)."""
    keep = []
    d = InversionDesc()
    d.post = _post_args(item["tables"], item["mean"], item["data"], item["cov"], item["rxy"], item["rt"],
                        item["lam2d"], precision, cond_limit, want_residual, keep)
    st = (ctypes.c_uint64 * 3)()
    if stats:
        d.post.stats = ctypes.cast(st, ctypes.POINTER(ctypes.c_uint64))
    for g in range(3):
        d.mean_in[g] = item["mean_in"][g]._h
    for l, s_ in enumerate(item["seis_in"]):
        d.seis_in[l] = s_._h
    outs = item.get("out", [])
    d.nreal = len(outs)
    arr = (_vp * max(1, 3 * len(outs)))(*[g._h for r in outs for g in r])
    d.out = ctypes.cast(arr, ctypes.POINTER(_vp))
    d.seed = int(item.get("seed", 0))
    L = lib()
    L.cg_invert.argtypes = [_vp, ctypes.POINTER(InversionDesc)]
    L.cg_invert.restype = ctypes.c_int
    ctx.check(L.cg_invert(ctx._h, ctypes.byref(d)), "cg_invert")
    return tuple(int(x) for x in st) if stats else None


class Batch:
    """Many inversions of any sizes planned together (cg_batch): every FFT pass, the posterior
    (per angle count) and the colouring of all of them share launches.  Each item is a dict
    with tables, rxy, rt, lam2d, mean_in[3], seis_in[nt] (real-domain inputs), mean[3],
    data[nt], cov[6] (work grids), out (list of [3] grids per realisation), seed."""

    def __init__(self, ctx, items, precision=PREC_MIXED, cond_limit=1e6, want_residual=False):
        self.ctx = ctx
        self._keep = []
        descs = (InversionDesc * len(items))()
        for d, it in zip(descs, items):
            d.post = _post_args(it["tables"], it["mean"], it["data"], it["cov"], it["rxy"], it["rt"], it["lam2d"],
                                precision, cond_limit, want_residual, self._keep)
            for g in range(3):
                d.mean_in[g] = it["mean_in"][g]._h
            for l, s in enumerate(it["seis_in"]):
                d.seis_in[l] = s._h
            outs = it.get("out", [])
            d.nreal = len(outs)
            arr = (_vp * max(1, 3 * len(outs)))(*[g._h for r in outs for g in r])
            self._keep.append(arr)
            d.out = ctypes.cast(arr, ctypes.POINTER(_vp))
            d.seed = int(it.get("seed", 0))
        self._h = _vp()
        L = lib()
        L.cg_batch_create.argtypes = [_vp, ctypes.c_int, ctypes.POINTER(InversionDesc), ctypes.POINTER(_vp)]
        L.cg_batch_create.restype = ctypes.c_int
        for name in ("cg_batch_run", "cg_batch_wait", "cg_batch_destroy"):
            getattr(L, name).argtypes = [_vp]
            getattr(L, name).restype = ctypes.c_int
        L.cg_batch_use_graph.argtypes = [_vp, ctypes.c_int]
        L.cg_batch_use_graph.restype = ctypes.c_int
        L.cg_batch_query.argtypes = [_vp, _i_p]
        L.cg_batch_query.restype = ctypes.c_int
        ctx.check(L.cg_batch_create(ctx._h, len(items), descs, ctypes.byref(self._h)), "cg_batch_create")
        ctx._batches.add(self)
        self._keep = None  # host tables were copied to the device

    def use_graph(self, on=True):
        self.ctx.check(lib().cg_batch_use_graph(self._h, 1 if on else 0), "cg_batch_use_graph")

    def run(self):
        self.ctx.check(lib().cg_batch_run(self._h), "cg_batch_run")

    def done(self):
        d = ctypes.c_int()
        self.ctx.check(lib().cg_batch_query(self._h, ctypes.byref(d)), "cg_batch_query")
        return bool(d.value)

    def wait(self):
        self.ctx.check(lib().cg_batch_wait(self._h), "cg_batch_wait")

    def close(self):
        if self._h and self.ctx._h:
            lib().cg_batch_destroy(self._h)
        self._h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def make_circ_auto_cov_pos_def(circ_auto_cov):
    """SeismicParametersHolder::MakeCircAutoCovPosDef on the nine circular auto- and
    cross-covariance series (3, 3, nzp); returns the adjusted series (host computation)."""
    a = np.array(circ_auto_cov, dtype=np.float32, order="C", copy=True)
    if a.ndim != 3 or a.shape[:2] != (3, 3):
        raise ValueError("circ_auto_cov must be (3, 3, nzp)")
    rc = lib().cg_make_circ_auto_cov_pos_def(a.shape[2], _ptr(a, ctypes.c_float))
    if rc != CG_OK:
        raise CravaError("cg_make_circ_auto_cov_pos_def failed (%d): a lag-0 variance is not positive" % rc)
    return a


class LocalNoise:
    """AVOInversion::correctVpVsRho (avoinversion.cpp:2146-2302) on the device: the per-trace
    3 x 3 matrices of newPosteriorCovPointwise from the local noise scale grids (n_theta arrays
    (ny, nx), Grid2D (i, j) at [j, i]) and the operator G of AVOInversion::computeG
    (crava_amd.model.compute_g), then applied to the real-domain posterior means (correct) and to
    realisations' noise fields before the posterior is added (correct_realisation)."""

    def __init__(self, ctx, G, prior_var0, post_var0, err_theta_cov, noise_scales):
        sc = [np.ascontiguousarray(a, dtype=np.float64) for a in noise_scales]
        ny, nx = sc[0].shape
        nt = len(sc)
        self.ctx, self.nx, self.ny = ctx, nx, ny
        keep = [np.ascontiguousarray(G, dtype=np.float64).reshape(nt, 3),
                np.ascontiguousarray(prior_var0, dtype=np.float64).reshape(3, 3),
                np.ascontiguousarray(post_var0, dtype=np.float64).reshape(3, 3),
                np.ascontiguousarray(err_theta_cov, dtype=np.float64).reshape(nt, nt)]
        ptrs = (ctypes.POINTER(ctypes.c_double) * nt)(*[_ptr(a, ctypes.c_double) for a in sc])
        args = LocalNoiseArgs(nx, ny, nt, *[_ptr(a, ctypes.c_double) for a in keep], ptrs)
        self._h = _vp()
        ctx.check(lib().cg_local_noise_create(ctx._h, ctypes.byref(args), ctypes.byref(self._h)),
                  "cg_local_noise_create")

    def matrices(self):
        """(sigmanew, kept), each (ny, nx, 3, 3)."""
        a = np.empty((self.ny, self.nx, 3, 3))
        b = np.empty((self.ny, self.nx, 3, 3))
        self.ctx.check(lib().cg_local_noise_matrices(self._h, _ptr(a, ctypes.c_double), _ptr(b, ctypes.c_double)),
                       "cg_local_noise_matrices")
        return a, b

    def correct(self, post, mean):
        self.ctx.check(lib().cg_local_noise_correct(self._h, (_vp * 3)(*[g._h for g in post]),
                                                    (_vp * 3)(*[g._h for g in mean])), "cg_local_noise_correct")

    def correct_realisation(self, real, post):
        self.ctx.check(lib().cg_local_noise_correct_realisation(self._h, (_vp * 3)(*[g._h for g in real]),
                                                                (_vp * 3)(*[g._h for g in post])),
                       "cg_local_noise_correct_realisation")

    def close(self):
        if self._h:
            lib().cg_local_noise_destroy(self._h)
            self._h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
