/*
 * cravagpu.h -- C ABI of the MI355X-native FFT-domain Bayesian AVO inversion.
 *
 * A drop-in for the hot path of CRAVA: FFTGrid's 3D FFTs and complex accessors, and
 * AVOInversion's posterior (computePostMeanResidAndFFTCov) and simulation (simulate) calls.
 * Host C++ calls hand-written HIP kernels built for gfx950 through this ABI; nothing here
 * exposes torch or HIP types.  Reference paths are relative to the CRAVA source tree.
 *
 * Conventions
 *  - Every function returns 0 on success or a negative CG_E* code; cg_last_error(ctx) holds
 *    the message of the last failure on that context.  No exception crosses the ABI.
 *  - One context owns one device and one HIP stream.  A context is not re-entrant: use one
 *    per host thread.  All calls enqueue on the context stream; cg_ctx_sync waits for them.
 *  - Grids keep CRAVA's in-place FFTW layout on the device (fftgrid.cpp:40-70): real cells at
 *    i + rnxp*j + rnxp*nyp*k with rnxp = 2*(nxp/2+1) floats; complex cells at
 *    i + cnxp*j + cnxp*nyp*k with cnxp = nxp/2+1, i = x being the half-spectrum axis.
 *  - Host buffers are borrowed for the duration of a call only.
 */
#ifndef CRAVAGPU_H
#define CRAVAGPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CG_ABI_VERSION 1

/* status codes */
#define CG_OK 0
#define CG_EINVAL (-1)   /* bad argument (sizes, null pointers, wrong grid state) */
#define CG_ENOMEM (-2)   /* device allocation failed or HBM budget exceeded */
#define CG_EHIP (-3)     /* HIP runtime error */
#define CG_ESTATE (-4)   /* operation not valid in the grid's current domain */
#define CG_EIO (-5)      /* a file could not be opened, read or written */

/* FFTGrid cube types (fftgrid.h:155); the value selects FFT scaling. */
#define CG_CT_DATA 1
#define CG_CT_PARAMETER 2
#define CG_CT_COVARIANCE 3

typedef struct cg_ctx cg_ctx;
typedef struct cg_grid cg_grid;
typedef struct cg_arena cg_arena;

/* ------------------------------------------------------------------ context */
/* Replaces the process-global FFTGrid statics (fftgrid.cpp:2661-2668): the budget is the
 * analogue of setMaxAllowedGrids/terminateOnMaxGrid, enforced per context (0 = no limit). */
int cg_ctx_create(cg_ctx **ctx, int device, uint64_t hbm_budget_bytes);
/* Waits for the stream, then releases the context and any grids still attached to it (their
 * handles become invalid). */
int cg_ctx_destroy(cg_ctx *ctx);
const char *cg_last_error(const cg_ctx *ctx);
int cg_ctx_sync(cg_ctx *ctx);
/* Bytes of device memory held by grids and work buffers of this context. */
int cg_ctx_memory(const cg_ctx *ctx, uint64_t *in_use, uint64_t *peak);
/* The context's HIP stream (hipStream_t) for callers that interoperate; opaque here. */
int cg_ctx_stream(const cg_ctx *ctx, void **stream);
/* Per-kernel timing with HIP events on the context stream: when enabled, every launch of the
 * named kernel class is bracketed by events; cg_ctx_kernel_time returns launch count and the
 * summed duration in ms (synchronises).  Classes: see CG_K_* below. */
int cg_ctx_set_profiling(cg_ctx *ctx, int enable);
int cg_ctx_kernel_time(cg_ctx *ctx, int kernel_class, uint64_t *launches, double *total_ms);
int cg_ctx_reset_kernel_time(cg_ctx *ctx);
#define CG_K_FFT_X 0     /* x-axis r2c / c2r passes */
#define CG_K_FFT_Y 1     /* y-axis complex passes */
#define CG_K_FFT_Z 2     /* z-axis complex passes */
#define CG_K_POSTERIOR 3 /* per-wavenumber posterior update */
#define CG_K_NOISE 4     /* noise generation and colouring */
#define CG_K_ELEMENTWISE 5
#define CG_K_COUNT 6
/* FFT engine selection: CG_FFT_AUTO uses the compile-time radix plans where a length has one
 * and the runtime-plan engine otherwise; CG_FFT_GENERIC forces the runtime-plan engine. */
#define CG_FFT_AUTO 0
#define CG_FFT_GENERIC 1
int cg_ctx_set_fft_engine(cg_ctx *ctx, int engine);
/* Autotuner (no reference counterpart; FFTW 2's planner plays this part in the reference): every
 * compiled length has two tile variants per pass (y/z: columns per workgroup; x: rows per
 * workgroup).  cg_autotune times both for every pass of a grid of this padded size on scratch
 * memory and keeps the faster for later transforms of that geometry, by any context of the
 * device (batched launches use the choice for their groups' first grid).  Without tuning,
 * variant 0 runs.  cg_ctx_set_fft_variant forces a variant (0 or 1) or restores the tuned
 * choice (-1). */
int cg_autotune(cg_ctx *ctx, int nxp, int nyp, int nzp);
int cg_ctx_set_fft_variant(cg_ctx *ctx, int variant);
/* Realisations with counter noise (cg_simulate, no reference counterpart: a schedule choice):
 * the colouring can run inside the inverse z pass of the three output grids (one kernel, the
 * coloured spectra never written to HBM) or as its own kernel followed by three z passes.  Mode
 * -1 (default): the autotuned choice for the geometry (cg_autotune times both), else fused when
 * the fused kernel keeps at least 12 wavefronts per CU resident; 0: never fused; 1: fused
 * whenever nzp has a compiled plan.  On grids with a compiled split of the z transform (see
 * cg_invert) modes -1 and 1 colour with the split kernels (the colouring carries the inverse z
 * transform's second stage, the x pass its first); 0 keeps the unsplit, unfused colouring.
 * Results agree to FP32 rounding either way. */
int cg_ctx_set_fused_colour(cg_ctx *ctx, int mode);

/* ------------------------------------------------------------------ grids (FFTGrid) */
/* FFTGrid(nx,ny,nz,nxp,nyp,nzp) + createRealGrid/createComplexGrid (fftgrid.cpp:40, 1048-1063),
 * device-resident.  The new grid is in the real (time) domain and zero-filled. */
int cg_grid_create(cg_ctx *ctx, int nx, int ny, int nz, int nxp, int nyp, int nzp, int cubetype,
                   cg_grid **grid);
int cg_grid_destroy(cg_grid *grid);
/* An arena: one device allocation from which grids are carved (256-byte aligned), so that a
 * request holding dozens of grids costs one allocation and one release.  No reference
 * counterpart (FFTGrid allocates per grid).  Grids of an arena are destroyed as usual (which
 * releases no memory); cg_arena_destroy needs every grid of the arena destroyed first and
 * frees the allocation after the work already enqueued on the context stream.
 * cg_arena_grid_bytes gives the bytes a grid of that padded size takes in an arena. */
int cg_arena_create(cg_ctx *ctx, uint64_t bytes, cg_arena **arena);
int cg_arena_destroy(cg_arena *arena);
uint64_t cg_arena_grid_bytes(int nxp, int nyp, int nzp);
int cg_grid_create_in(cg_arena *arena, int nx, int ny, int nz, int nxp, int nyp, int nzp, int cubetype,
                      cg_grid **grid);
/* setType (fftgrid.h) */
int cg_grid_set_type(cg_grid *grid, int cubetype);
/* getIsTransformed (fftgrid.h): 1 in the Fourier domain, 0 in the real domain. */
int cg_grid_is_transformed(const cg_grid *grid, int *transformed);
int cg_grid_dims(const cg_grid *grid, int dims[6]);
/* Raw device pointer to the rnxp*nyp*nzp floats (aliasing the complex view). */
int cg_grid_device_ptr(cg_grid *grid, void **ptr);
/* Host mirror for cold callers: whole-grid copies in the reference layout.  Upload of a real
 * grid sets the real domain; upload of a complex grid sets the Fourier domain
 * (createComplexGrid + setNextComplex loop). */
int cg_grid_upload_real(cg_grid *grid, const float *host_rvalue);
int cg_grid_upload_complex(cg_grid *grid, const float *host_cvalue);
int cg_grid_download_real(cg_grid *grid, float *host_rvalue);
int cg_grid_download_complex(cg_grid *grid, float *host_cvalue);
/* getComplexValue / setComplexValue (fftgrid.cpp:1335-1363, 1467-1487): random access; out of
 * range gives RMISSING (-99999) in both parts, resp. returns 1 and writes nothing.  With
 * ext_simbox = 0 the range is the unpadded nx*ny*nz box, else the padded box. */
int cg_grid_get_complex(cg_grid *grid, int i, int j, int k, int ext_simbox, float value[2]);
int cg_grid_set_complex(cg_grid *grid, int i, int j, int k, int ext_simbox, const float value[2]);
/* getRealValue (fftgrid.cpp:1237-1263) in the real domain: RMISSING outside the box (ext_simbox = 0)
 * or the padded box; SeismicParametersHolder::getOrigin reads (0, 0, 0) with it. */
int cg_grid_get_real_value(cg_grid *grid, int i, int j, int k, int ext_simbox, float *value);
/* getRealTrace (fftgrid.cpp:1266-1283): value[k] = getRealValue(i, j, k) for k < nz, or k < nzp
 * with add_padding (whose padding part then reads RMISSING, as getRealValue tests the unpadded
 * box); SeismicParametersHolder::createPostCov00 (seismicparametersholder.cpp:1056-1067) is the
 * trace at (0, 0).  Real domain. */
int cg_grid_get_real_trace(cg_grid *grid, int i, int j, int add_padding, float *value);

/* FFTGrid::fftInPlace (fftgrid.cpp:1587-1612): DATA/PARAMETER are scaled by 1/sqrt(N) before
 * the forward transform, COVARIANCE is not.  Sign -1, FFTW pass order x (r2c), y, z. */
int cg_fft_r2c_inplace(cg_grid *grid);
/* FFTGrid::invFFTInPlace (fftgrid.cpp:1615-1645): z, y complex, then x half-complex-to-real
 * that ignores Im of the DC and Nyquist bins; x 1/N (COVARIANCE) or 1/sqrt(N) afterwards. */
int cg_fft_c2r_inplace(cg_grid *grid);
/* Out-of-place form of fftInPlace for callers that keep the real cube: dst (same padded size)
 * receives src's spectrum with src's cube type; src is unchanged.  No reference counterpart
 * (FFTGrid only transforms in place); it lets resident inputs be transformed repeatedly. */
int cg_fft_r2c(const cg_grid *src, cg_grid *dst);
/* FFTGrid::add / multiplyByScalar (fftgrid.cpp:1659-1678, 1766-1773). */
int cg_grid_add(cg_grid *dst, const cg_grid *src);
int cg_grid_multiply_by_scalar(cg_grid *grid, float scalar);
/* FFTGrid::fillInConstant (fftgrid.cpp:634-651): real domain, value on every padded cell with
 * x < nxp, RMISSING (-99999) on the row padding. */
int cg_grid_fill_constant(cg_grid *grid, float value);

/* ------------------------------------------------------------------ padding fills (input side) */
/* FFTGrid(const NRLib::Grid<float> *grid, nxp, nyp, nzp) (fftgrid.cpp:125-213), into an
 * existing grid of those sizes: the unpadded cube (nz*ny*nx floats, i fastest as NRLib::Grid
 * stores it, grid.hpp:231-236) is copied in and the padding filled by linear interpolation
 * between the box edges (x first, then y from x-padded values, then z from both), the row
 * padding x >= nxp with 0.  cube_on_device: 1 if `cube` is device memory, 0 if host memory
 * (copied in; the call then returns after the fill).  Sets type PARAMETER, real domain. */
int cg_grid_fill_padded(cg_grid *grid, const float *cube, int cube_on_device);
/* FFTGrid::interpolateSeismic (fftgrid.cpp:2349-2437) on a DATA grid in the real domain: box
 * traces whose energy over the nz box samples is at most energy_threshold * mean energy are
 * replaced by the mean of their good 4-neighbours (interpolateTrace, fftgrid.cpp:2440-2481, in
 * the reference's order), then the padding is extrapolated (extrapolateSeismic below) from the
 * rectangle of good traces.  n_bad and rect = {imin, imax, jmin, jmax} may be null.  CG_EINVAL
 * when every trace is below the limit (the reference's sweep then reads cells it overwrites). */
int cg_grid_interpolate_seismic(cg_grid *grid, float energy_threshold, int *n_bad, int rect[4]);
/* FFTGrid::extrapolateSeismic (fftgrid.cpp:2484-2518): every column outside the box rectangle
 * [imin, imax] x [jmin, jmax] (real domain, x < nxp) becomes the column of its clamped fill
 * number (getFillNumber, fftgrid.cpp:1103-1140) times max(1 - dx^2 - dy^2 - dz^2, 0)^3, d the
 * distances into the padding (getDistToBoundary, fftgrid.cpp:1157-1200).  The rectangle must be
 * non-empty and inside the box (CG_EINVAL otherwise). */
int cg_grid_extrapolate_seismic(cg_grid *grid, int imin, int imax, int jmin, int jmax);
/* FFTGrid::fillInParamCorr / FillInLateralCorr (fftgrid.cpp:830-923) followed by
 * multiplyByScalar(scale) (seismicparametersholder.cpp:311-348 scales each covariance grid by its
 * priorVar0 entry): cell (i, j, k), i < nxp, gets the temporal correlation corr_t (nzp floats:
 * circCorrT or a circ_auto_cov entry) linearly interpolated, cyclically, at k + cycleI grad_i +
 * cycleJ grad_j (cycleI = i or i - nxp, whichever is nearer 0), times corr_xy[i + nxp j] (the
 * lateral correlation surface, nyp*nxp floats); the row padding gets RMISSING; then every cell is
 * multiplied by scale.  Host tables; real domain. */
int cg_grid_fill_param_corr(cg_grid *grid, const float *corr_xy, const float *corr_t, float grad_i, float grad_j,
                            float scale);
/* The trace taper of CravaResult::ComputeSyntSeismic (cravaresult.cpp:2338-2346): box columns of
 * the cube (nz*ny*nx floats, i fastest) keep their nz samples, and over the z padding run from the
 * last sample back to the first, fac ((k - nz) r_0 + (nzp - k - 1) r_{nz-1}) with
 * fac = 1 / (nzp - nz - 1); other columns and the row padding are 0.  With cg_fft_r2c,
 * cg_forward_model (K = A D conj(W), D the backward difference 1 - e^{-2 pi i k / nzp}) and
 * cg_fft_c2r_inplace it gives ComputeSyntSeismic's synthetic seismic (the operators act along z
 * only).  Needs nzp >= nz + 2.  cube_on_device as in cg_grid_fill_padded.  Real domain. */
int cg_grid_fill_trace_taper(cg_grid *grid, const float *cube, int cube_on_device);
/* FFTGrid::fillInErrCorr (fftgrid.cpp:777-828): corr_xy[i + nxp j] times 1 - (the distance of
 * k + cycleI grad_i + cycleJ grad_j from the k = 0 plane, cyclically) where that distance is
 * below one sample, 0 elsewhere, RMISSING on the row padding.  Real domain. */
int cg_grid_fill_err_corr(cg_grid *grid, const float *corr_xy, float grad_i, float grad_j);

/* ------------------------------------------------------------------ non-constant thickness */
/* The seismic data rescaling of the non-constant-thickness branch, trace by trace along z on a
 * real-domain DATA grid (every trace transformed, filtered and transformed back on the device):
 *  CG_SCALE_WAVELET_DIVIDE   AVOInversion::divideDataByScaleWavelet (avoinversion.cpp:485-579)
 *      with computeAdjustmentFactor (:582-719): every padded trace, the z padding tapered by
 *      max(1 - dist^2, 0) beyond nz (getDistToBoundary), the spectrum multiplied by the adjustment
 *      factor of its local wavelet inside lowCut < deltaF k < highCut (deltaF from the trace's
 *      relative thickness) and set to zero outside.  The reference then calls interpolateSeismic
 *      (cg_grid_interpolate_seismic);
 *  CG_SCALE_WAVELET_MULTIPLY multiplyDataByScaleWaveletAndWriteToFile (:721-790) before the write:
 *      every box trace's spectrum times the conjugate of its local wavelet stretched by relThick *
 *      localStretch; the padding traces are left (up to FP32 rounding of the transform pair).
 * The local wavelet of a trace is the angle's wavelet shifted by shift_ms and scaled by gain
 * (Wavelet1D::shiftAndScale, wavelet1D.cpp:580-602) and stretched (Wavelet::getCAmp(k, scale),
 * wavelet.cpp:337-394); divide looks it up at the box trace the reference mirrors a padded trace
 * to (avoinversion.cpp:516-533).  No reference counterpart in the signature: the arrays replace
 * Simbox and Wavelet queries. */
#define CG_SCALE_WAVELET_DIVIDE 0
#define CG_SCALE_WAVELET_MULTIPLY 1
typedef struct {
  int mode;                  /* CG_SCALE_WAVELET_* */
  double dt_ms;              /* sample interval (lz = nz dt) */
  const float *wavelet;      /* [nzp] the angle's wavelet, time samples in FFT order */
  const double *rel_thick;   /* [nyp][nxp] Simbox::getRelThick at every padded trace */
  const float *shift_ms;     /* [ny][nx] local time shifts (Wavelet::getLocalTimeshift) or null */
  const float *gain;         /* [ny][nx] local gain factors (getLocalGainFactor) or null */
  const float *stretch;      /* [ny][nx] local stretch (getLocalStretch) or null */
  float low_cut, high_cut;   /* Hz */
  float wnc;                 /* white-noise component */
  float error_var;           /* errThetaCov[l][l] of the angle */
  const float *corr_t;       /* [nzp] SeismicParametersHolder::getPriorCorrTFiltered(nz, nzp) */
  const float *A;            /* [3] the angle's reflection coefficients */
  const double *prior_var0;  /* [3][3] */
} cg_scale_wavelet_args;
int cg_grid_scale_wavelet(cg_grid *grid, const cg_scale_wavelet_args *args);

/* ------------------------------------------------------------------ CRAVA binary grid files */
/* Survey geometry stored in the file header (Simbox getters written by writeCravaFile). */
typedef struct {
  double x0, y0, dx, dy;
  int nx, ny;
  double il0, xl0, il_step_x, il_step_y, xl_step_x, xl_step_y, angle;
} cg_crava_geometry;
/* FFTGrid::writeCravaFile (fftgrid.cpp:2183-2221) to `path` as given (the reference appends
 * IO::SuffixCrava()): the tag line "crava_fftgrid_binary", the geometry, rnxp, nyp, nzp and the
 * grid's rnxp*nyp*nzp floats as stored (either domain), all big-endian.  CG_EIO if the file
 * cannot be written. */
int cg_grid_write_crava(cg_grid *grid, const char *path, const cg_crava_geometry *geometry);
/* FFTGrid::readCravaFile (fftgrid.cpp:2224-2271): the file's floats into the grid, which must have
 * the file's rnxp, nyp and nzp (CG_EINVAL, "Grid dimension is wrong", otherwise); the tag line is
 * not checked.  Sets the real domain; geometry (may be null) receives the header.  CG_EIO if the
 * file cannot be opened or ends early (the grid's contents are then undefined). */
int cg_grid_read_crava(cg_grid *grid, const char *path, cg_crava_geometry *geometry);

/* ------------------------------------------------------------------ STORM grid files */
/* Simbox geometry in a STORM header (Simbox::getStormHeader, simbox.cpp:1012-1034). */
typedef struct {
  double x_min, lx, y_min, ly;  /* GetXMin, GetLX, GetYMin, GetLY */
  double lz;                    /* GetLZ */
  double angle;                 /* GetAngle, radians (the header stores degrees) */
  const char *top_name;         /* surface names of a non-flat simbox, or null for a flat one */
  const char *bot_name;
} cg_storm_geometry;
/* FFTGrid::writeStormFile (fftgrid.cpp:1889-1957), binary form, to `path` as given (the reference
 * appends IO::SuffixStormBinary()): the header of getStormHeader with the grid's cube type, then
 * the box (padding = 0) or the padded cube (padding = 1) of the real grid, k, j, i order, i fastest,
 * big-endian floats (packed and byte-swapped on the device), then "0\n".  Real domain; CG_EIO if the
 * file cannot be written. */
int cg_grid_write_storm(cg_grid *grid, const char *path, const cg_storm_geometry *geometry, int padding);
/* StormContGrid::ReadFromFile (stormcontgrid.cpp:82-170) of a STORM binary grid whose nx, ny, nz are
 * the grid's box, followed by the padding fill of FFTGrid(const NRLib::Grid *) (as
 * cg_grid_fill_padded).  geometry (may be null) receives the volume (surface names are not
 * returned).  CG_EINVAL for another format or size, CG_EIO if the file cannot be read or ends early. */
int cg_grid_read_storm(cg_grid *grid, const char *path, cg_storm_geometry *geometry);

/* CommonData::FillInData (commondata.cpp:6451-6775), the SEG-Y seismic (DATA) branch, with the
 * columns resolved by the caller as FillInData resolves them (GetFillNumber, Simbox::getCoord /
 * getdz, SegyGeometry::IsInside, SegY::GetNearestTrace): column (i, j) of the padded grid,
 * i < rnxp, takes trace column_trace[j][i], or zeros where that is -1 (outside the data) or -2
 * (a dead trace).  Each trace is tapered over floor(smooth_length / dz_data) samples at each end
 * (SmoothTraceInGuardZone, :6847-6885), resampled ten times finer through its spectrum of length
 * nt = the closest 2^a 3^b 5^c 7^d >= the longest trace (ResampleTrace, commondata.h:136-190),
 * and interpolated at the cell centres z0_grid + refk dz_grid, refk = GetZSimboxIndex(k)
 * (InterpolateGridValues, :6947-6995).  The grid becomes a real-domain DATA grid.  Traces at most
 * 1024 samples, none shorter than its taper; host arrays, copied before the call returns.
 * grid_type CG_CT_PARAMETER: FillInData's branch for other grids (background and parameter cubes
 * from SEG-Y, :6613-6644 and :6684-6721): no taper; where a trace's first sample is zero, its zero
 * samples are dropped and z0_data moves down one dz_data per leading zero; the line through the
 * first and last samples is subtracted before the resampling and, resampled linearly ten times
 * finer, interpolated at the cells (InterpolateAndShiftTrend, :6997-7050, its end values outside)
 * and added back.  column_trace -2 marks a dead trace (or a missing cell column): RMISSING there
 * and for traces with no nonzero sample (zeros in the DATA branch).  Every box column (i < nx,
 * j < ny) must have a trace or be dead (the reference stops and reports the grid as not covering
 * the area): CG_EINVAL otherwise.  The grid becomes a real-domain PARAMETER grid; z0_data is then
 * GetNearestTrace's z0 itself, z0_grid the cell centre z0 (is_seismic false). */
typedef struct {
  const float *samples;      /* [ntraces][stride] */
  const int *length;         /* [ntraces] samples of each trace, 1..stride */
  const float *z0_data;      /* [ntraces] time of the first sample, ms: GetNearestTrace's z0
                              * less half a sample (commondata.cpp:6584-6585) */
  int ntraces, stride;
  float dz_data;             /* sample interval, ms */
  float smooth_length;       /* guard-zone taper, ms */
  const int *column_trace;   /* [nyp][rnxp] */
  const float *z0_grid;      /* [nyp][rnxp] centre of the column's first cell, ms (z0 + dz / 2) */
  const float *dz_grid;      /* [nyp][rnxp] */
  int grid_type;             /* CG_CT_DATA (or 0) / CG_CT_PARAMETER */
  int trim_keeps_z0;         /* 1: trimmed zeros leave z0_data where it is, as FillInData does for
                              * traces from a STORM grid (the shift is SEG-Y only, :6622-6628) */
} cg_trace_fill_args;
int cg_grid_fill_from_traces(cg_grid *grid, const cg_trace_fill_args *args);

/* ------------------------------------------------------------------ posterior (AVOInversion) */
#define CG_PRIOR_GRIDS 0     /* six prior covariance spectra, as SeismicParametersHolder holds */
#define CG_PRIOR_SEPARABLE 1 /* Sigma_m(k) = |rxy(kx,ky) * rt(kz)| * Sigma0 (no grids read) */

#define CG_PREC_REFERENCE 0 /* FP32 throughout, the reference's decisions (prior on failure) */
#define CG_PREC_MIXED 1     /* FP32, FP64 re-solve where cond(M) > cond_limit or FP32 fails */

typedef struct {
  int ntheta;               /* number of angle gathers, 1..8 */
  /* per-k_z tables (avoinversion.cpp:944-975), host memory, complex interleaved:
   * K [nzp][ntheta][3], e1/e2/e3 [nzp][ntheta]; invert [nzp] (0/1) */
  const float *K;
  const float *e1;
  const float *e2;
  const float *e3;
  const int *invert;
  const double *sigma0e;    /* [ntheta][ntheta] errThetaCov (modelavodynamic.cpp:623-651) */
  float wnc;                /* white-noise component (modelsettings.cpp:98) */
  const double *prior_var0; /* [3][3] priorVar0_ */
  int prior_mode;           /* CG_PRIOR_* */
  /* CG_PRIOR_GRIDS: six COVARIANCE grids in the Fourier domain, order Vp, Vs, Rho, VpVs,
   * VpRho, VsRho; they receive the posterior covariance spectra.
   * CG_PRIOR_SEPARABLE: host tables rxy [nyp][cnxp] and rt [nzp]; the posterior covariance
   * spectra are written to the six grids in cov[] (which must exist, any content). */
  cg_grid *cov[6];
  const float *rxy;
  const float *rt;
  /* error-correlation spectrum lambda: either the errCorr COVARIANCE grid in the Fourier
   * domain (avoinversion.cpp:917, 1248) or a host table lam2d [nyp][cnxp] (fillInErrCorr puts
   * the lateral correlation in the k = 0 plane only, so its spectrum does not depend on kz). */
  cg_grid *err_corr;
  const float *lam2d;
  cg_grid *mean[3];         /* Fourier-domain prior means in, posterior means out */
  cg_grid *data[8];         /* Fourier-domain seismic in, residual out (if want_residual) */
  int want_residual;
  int precision;            /* CG_PREC_* */
  double cond_limit;        /* CG_PREC_MIXED threshold, e.g. 1e6 */
  uint64_t *stats;          /* optional host [3]: inverted, cholesky failures, FP64 re-solves */
} cg_post_args;

/* AVOInversion::computePostMeanResidAndFFTCov, per-wavenumber part (avoinversion.cpp:944-1083).
 * All grids in the Fourier domain; the caller transforms before and after, as the reference
 * does at :229-244, :908-917 and :1112-1114. */
int cg_posterior(cg_ctx *ctx, const cg_post_args *args);
/* cg_posterior without any device allocation.  No counterpart in the reference; it is for callers
 * that cannot allocate device memory (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  Same arguments and results as cg_posterior (the same kernels; the FP64
 * re-solves are found through a bit mask instead of a list).  Its tables, counters and deferral
 * mask live in one fixed buffer of 55 MB per device, a file-scope device array that every
 * GPU the library runs on holds until the process ends; grids beyond nzp <= 1536,
 * (nxp/2+1)*nyp <= 2^21 and (nxp/2+1)*nyp*nzp <= 2^28 complex cells are refused with CG_EINVAL.
 * Every call on a device shares that buffer: calls that overlap on one GPU (other contexts or
 * streams) race and must not run concurrently.  The host tables are copied straight from host
 * memory and the call waits for the copies (one stream synchronisation). */
int cg_posterior_noalloc(cg_ctx *ctx, const cg_post_args *args);

/* ------------------------------------------------------------------ simulation */
/* Counter noise: stream s, complex cell c draws the Philox-4x32-10 block with counter
 * (c, s >> 1, 0) and key seed, and takes a normal pair by Box-Muller from words 0-1 (even s)
 * or 2-3 (odd s); the fillInComplexNoise layout then places the pairs (Hermitian cells). */
#define CG_NOISE_COUNTER 0   /* counter-based Philox-4x32-10 + Box-Muller, keyed by seed */
#define CG_NOISE_GRIDS 1     /* caller-supplied noise spectra in seed[] (e.g. the reference stream) */

typedef struct {
  cg_grid *cov[6];          /* posterior covariance spectra (Fourier domain, order as above) */
  cg_grid *mean[3];         /* posterior means (optional), either in the real domain (after
                             * cg_fft_c2r_inplace: added after the inverse transform) or as
                             * PARAMETER spectra (before it: added to the coloured spectra, which
                             * saves one read of each mean; all three in the same domain) */
  cg_grid *out[3];          /* realisation grids: receive Vp, Vs, Rho real fields */
  int noise_source;         /* CG_NOISE_* */
  uint64_t seed;            /* CG_NOISE_COUNTER key */
  uint32_t realisation;     /* CG_NOISE_COUNTER realisation number (streams 3r, 3r+1, 3r+2) */
  /* CG_NOISE_GRIDS: out[] already hold the noise spectra (fillInComplexNoise result) */
  uint64_t *stats;          /* optional host [1]: Cholesky failures (noise zeroed) */
} cg_sim_args;

/* One realisation of AVOInversion::simulate (avoinversion.cpp:1401-1511): noise,
 * z <- chol(Sigma_post) z per wavenumber (zero where the 3x3 Cholesky fails), unitary C2R,
 * + posterior mean. */
int cg_simulate(cg_ctx *ctx, const cg_sim_args *args);
/* Realisations args->realisation and args->realisation + 1 of the same posterior (no reference
 * counterpart: AVOInversion::simulate loops over realisations one at a time): the first into
 * args->out, the second into next_out.  With counter noise one colouring launch (fused with the
 * inverse z pass where cg_simulate would fuse it) reads and factorises each wavenumber's
 * covariance once for both; results equal two cg_simulate calls bit for bit.  With noise grids it
 * makes those two calls (next_out then holds the second noise).  stats receives the failures of
 * both. */
int cg_simulate_pair(cg_ctx *ctx, const cg_sim_args *args, cg_grid *next_out[3]);

/* FFTGrid::fillInComplexNoise layout (fftgrid.cpp:997-1045) with the counter-based generator:
 * the grid becomes a PARAMETER grid in the Fourier domain. */
int cg_grid_fill_noise_counter(cg_grid *grid, uint64_t seed, uint32_t stream);
/* FFTGrid::fillInComplexNoise (fftgrid.cpp:997-1045) with RandomGen (random.cpp:38-118,
 * random.h:11-14): the reference's own stream, bit for bit, generated in parallel.  *seed is the
 * reference's process-wide RandomGen seed before the fill and is advanced exactly as the
 * reference advances it, so consecutive calls continue one stream (three calls = the three seed
 * grids of one realisation, as AVOInversion::simulate draws them).  The grid becomes a
 * PARAMETER grid in the Fourier domain.  Synchronises the context stream. */
int cg_grid_fill_noise_reference(cg_grid *grid, uint32_t *seed);

/* ------------------------------------------------------------------ batched inversions */
/* Many independent inversions of any sizes (the reference runs one per process) planned
 * together: each pass of all of them shares a few launches (one per compiled FFT length, run
 * concurrently on side streams; one posterior launch per angle count; colouring), so small
 * cubes no longer pay one launch per grid and pass.  No reference counterpart; it serves
 * batched and continuously batched requests.  Results are identical to running each inversion
 * alone in a batch of one.  Grids may belong to any context on the same device (a server
 * generates inputs on one stream and runs batches on others); the batch runs on the stream of
 * `ctx`, so work that other contexts enqueued on its grids must have completed (cg_ctx_sync)
 * before cg_batch_run. */
typedef struct cg_batch cg_batch;
typedef struct {
  /* tables, prior, precision and work grids as for cg_posterior: post.mean[3] and post.data[]
   * receive the spectra of mean_in / seis_in; post.cov[6] the posterior covariance spectra;
   * after a run post.mean[] hold the real-domain posterior means. */
  cg_post_args post;
  cg_grid *mean_in[3]; /* real-domain prior means (PARAMETER), left unchanged */
  cg_grid *seis_in[8]; /* real-domain seismic cubes (DATA), left unchanged */
  int nreal;           /* posterior realisations */
  cg_grid **out;       /* [nreal][3] realisation grids (real domain after a run) */
  uint64_t seed;       /* counter-noise key; realisation r uses streams 3r, 3r+1, 3r+2 */
} cg_inversion;

int cg_batch_create(cg_ctx *ctx, int n, const cg_inversion *inv, cg_batch **batch);
/* Enqueue one run of the whole batch on the context stream (forward FFTs from the inputs,
 * posterior, inverse FFTs of the means, realisations). */
int cg_batch_run(cg_batch *batch);
/* With enable = 1 the next run captures its launches into a hipGraph, later runs replay it. */
int cg_batch_use_graph(cg_batch *batch, int enable);
/* done = 1 once the last enqueued run has completed (non-blocking). */
int cg_batch_query(cg_batch *batch, int *done);
int cg_batch_wait(cg_batch *batch);
int cg_batch_destroy(cg_batch *batch);

/* AVOInversion::computePostMeanResidAndFFTCov (avoinversion.cpp:229-244, 908-1114) followed by
 * AVOInversion::simulate (:1401-1511), for one inversion in one call on the context stream:
 * forward transforms of the real-domain inputs into post.mean[] / post.data[], the
 * per-wavenumber posterior, inverse transforms of the posterior means, then inv->nreal
 * realisations (counter noise, streams 3r..3r+2 for realisation r; consecutive pairs as
 * cg_simulate_pair; the posterior means added in the real domain).  Where the grid has a compiled
 * split nzp = 16 N2 for its x length and the call uses the separable prior with host tables
 * (rxy, rt, lam2d) and no residual, the z transform runs in two stages carried by the x passes
 * and the posterior pass (fewer passes over HBM, same results to FP32 rounding); otherwise the
 * drop-in sequence (cg_fft_r2c, cg_posterior, cg_fft_c2r_inplace) runs.  On return post.mean[]
 * hold the real-domain posterior means and post.cov[] the posterior covariance spectra;
 * post.data[] are work grids whose contents are unspecified unless want_residual (then the
 * residual spectra).  post.stats as for cg_posterior. */
int cg_invert(cg_ctx *ctx, const cg_inversion *inv);

/* ------------------------------------------------------------------ forward modelling */
/* Synthetic seismic spectra d_l(k) = sum_j K_lj(kz) m_j(k) from Fourier-domain elastic grids
 * m[3], with the per-k_z operator K of cg_post_args (host table [nzp][ntheta][3] complex).
 * The spectral counterpart of CravaResult::ComputeSyntSeismic (cravaresult.cpp:2312-2377),
 * which applies the same wavelet-times-reflectivity operator trace by trace. */
int cg_forward_model(cg_ctx *ctx, int ntheta, const float *K, cg_grid *m[3], cg_grid *d[]);
/* Real-domain white noise x += sigma * N(0,1) on every padded cell (counter-based generator). */
int cg_grid_add_white_noise(cg_grid *grid, float sigma, uint64_t seed, uint32_t stream);
/* Sum of squares of a real-domain grid over the unpadded box (ext_simbox = 0) or the padded
 * box, sum over cells of double(x)^2 in a fixed order (rows in x order, the rows of a plane in y
 * order, the planes in z order), so that equal grids give equal sums on every run.  Synchronises.
 * No reference counterpart (the synthetic inputs' signal variance). */
int cg_grid_sum_squares(cg_grid *grid, int ext_simbox, double *sum);
/* ModelAVODynamic::ComputeDataVariance (modelavodynamic.cpp:575-620) for one stack: the mean of
 * float(x * x) over the samples of the unpadded box with x != 0, summed in double in the
 * reference's order (each row in x order, the rows of a plane in y order, the planes in z order),
 * returned as float (0 when every sample is zero); count (may be null) receives the number of
 * non-zero samples.  Bit-identical to the reference's sequential sweep.  Real domain;
 * synchronises. */
int cg_grid_data_variance(cg_grid *grid, float *variance, uint64_t *count);

/* ------------------------------------------------------------------ SEG-Y grid files */
/* The simbox geometry FFTGrid::writeSegyFile reads (fftgrid.cpp:1960-2050). */
typedef struct {
  double x0, y0, dx, dy;      /* Simbox::getx0, gety0, getdx, getdy */
  int nx, ny;                 /* the box's traces (the grid's nx, ny) */
  double angle;               /* Simbox::getAngle, radians */
  double il0, xl0, il_step_x, il_step_y, xl_step_x, xl_step_y;  /* getIL0 ... getXLStepY */
  double dz;                  /* Simbox::getdz, ms */
  float z0;                   /* time of the file's first sample */
  double zmax;                /* Simbox::getMinMaxZ's maximum */
  const double *top;          /* [ny][nx] Simbox::getTop at each trace centre; -99999 (missing): trace not written */
  const double *rel_thick;    /* [ny][nx] Simbox::getRelThick(i, j) */
  const char *const *text;    /* textual header lines (from the second line on), or null */
  int ntext;                  /* 0..38 */
} cg_segy_geometry;
/* FFTGrid::writeSegyFile (fftgrid.cpp:1960-2050) with SegY's writer (libs/nrlib/segy: BinaryHeader,
 * SeisWorks trace headers, SegY::WriteAllTracesToFile with scalco 1 and offset 0), to `path` as
 * given: sample interval floor(dz + 0.5) ms (1 if 0), ceil((zmax - z0) / that) samples per trace;
 * trace (i, j) holds, from the sample nearest its top to the one nearest its bottom, the grid's
 * trace interpolated linearly at the sample times (getRegularZInterpolatedRealValue) and 0 above
 * and below; traces sorted by inline then crossline (of the float cell-centre coordinates), big-endian
 * IBM floats (truncated).  The 3200-byte textual header is "C" + line number + the given text, in
 * EBCDIC (not the reference's standard header text).  Real domain; samples computed on the device. */
int cg_grid_write_segy(cg_grid *grid, const char *path, const cg_segy_geometry *geometry);

/* NRLib's SegY reading (libs/nrlib/segy: BinaryHeader::Update, TraceHeader, SegY::ReadTrace) for
 * SegY::GetNearestTrace's callers such as CommonData::FillInData: every trace of a SEG-Y file in
 * file order.  The binary header gives the sample interval (bytes 3217-3218, microseconds), the
 * samples per trace (3221-3222) and the format (3225-3226: 1 IBM or 5 IEEE floats, big-endian);
 * each trace header gives inline, crossline, x, y (times scalco's factor as TraceHeader::Read sets
 * it: 10..10000 multiply, -10..-10000 divide, any other value 1) and the start time (ms).  Header byte positions (1-based; 0 or a null format: the SeisWorks
 * positions the SEG-Y writer uses, 9, 21, 73, 77, 71, 109).  Call with null buffers to learn
 * ntraces and nsamples, then with buffers of that size (any may be null).  CG_EIO if the file
 * cannot be read, CG_EINVAL for another format or a size that is not whole traces. */
typedef struct {
  int il_pos, xl_pos, x_pos, y_pos, scalco_pos, start_time_pos;
} cg_segy_format;
typedef struct {
  int ntraces, nsamples; /* from the file (first call); must match on the second */
  float dz;              /* sample interval, ms */
  int format_code;
  float *samples;        /* [ntraces][nsamples] */
  int *il, *xl;
  double *x, *y;
  float *start_time;
} cg_segy_traces;
int cg_segy_read(const char *path, const cg_segy_format *format, cg_segy_traces *traces);

/* ------------------------------------------------------------------ local noise */
/* Laterally varying data noise (AVOInversion::correctVpVsRho, avoinversion.cpp:2146-2302): per
 * trace (i, j) the 3 x 3 matrices of newPosteriorCovPointwise (:2028-2124) from the angle scales
 * s_l = noise_scale_l(i, j) / min noise_scale_l (Grid2D::FindMin, -99999 marking missing values),
 * computed on the device in FP64 by cg_local_noise_create (synchronises):
 *   sigmanew = (Sigma_m G^T M^+ G Sigma_m)^1/2 (Sigma_m - postVar0)^-1/2,
 *   kept     = (Sigma_m - Sigma_m G^T M^+ G Sigma_m)^1/2 postVar0^-1/2   (sigmamdnew_, :2232-2241),
 * M = G Sigma_m G^T + D Sigma_e D, D = diag(sqrt s), M^+ without eigenvalues <= 1e-8.  G is the
 * caller's AVOInversion::computeG (:1943-2026), whose LAPACK eigenvectors it keeps. */
typedef struct cg_local_noise cg_local_noise;
typedef struct {
  int nx, ny, ntheta;                 /* the box's traces; 1 to 8 angles */
  const double *G;                    /* ntheta x 3, row-major */
  const double *prior_var0;           /* Sigma_m, 3 x 3 row-major (priorVar0_) */
  const double *post_var0;            /* 3 x 3 (postVar0_, the lag-0 posterior covariance) */
  const double *err_theta_cov;        /* Sigma_e, ntheta x ntheta (errThetaCov_) */
  const double *const *noise_scale;   /* ntheta host arrays of nx*ny, (i, j) at i + nx*j */
} cg_local_noise_args;
int cg_local_noise_create(cg_ctx *ctx, const cg_local_noise_args *args, cg_local_noise **ln);
/* Host copies of the per-trace matrices, trace t = i + nx*j at [9t, 9t + 9) row-major; either
 * pointer may be null.  Synchronises. */
int cg_local_noise_matrices(const cg_local_noise *ln, double *sigmanew, double *kept);
/* The posterior update of correctVpVsRho (:2243-2263) in place on the real-domain posterior
 * means post[3] with the prior means mean[3] (meanVp2_ ...): inside the box
 * p = float(m + s_a0 d0 + s_a1 d1 + s_a2 d2), d = p - m in float, the sum in double. */
int cg_local_noise_correct(cg_local_noise *ln, cg_grid *const post[3], cg_grid *const mean[3]);
/* simulate's local-noise scaling (:1485-1504) and the posterior added after it: inside the box
 * the real-domain noise fields real[3] become float(k_a0 r0 + k_a1 r1 + k_a2 r2) (double), then
 * every padded cell gets post[3] added in float (FFTGrid::add). */
int cg_local_noise_correct_realisation(cg_local_noise *ln, cg_grid *const real[3], cg_grid *const post[3]);
void cg_local_noise_destroy(cg_local_noise *ln);

/* ------------------------------------------------------------------ host-side setup */
/* Per-k_z tables of the constant-thickness branch (avoinversion.cpp:944-975; Wavelet1D
 * difference operators wavelet1D.cpp:389-500; findNorm wavelet.cpp:882-900).  wavelets
 * [ntheta][nzp] real in FFT order, A [ntheta][3].  Outputs as in cg_post_args. */
int cg_build_kz_tables(int nz, int nzp, double dt_ms, int ntheta, const float *wavelets, const float *A,
                       float low_cut, float high_cut, float min_rel_thick, float *K, float *e1, float *e2,
                       float *e3, int *invert);
/* The same tables for the non-constant-thickness branch (avoinversion.cpp:976-1014, used when the
 * simbox is not of constant thickness and the seismic data have been divided by the scale wavelet,
 * cg_grid_divide_by_scale_wavelet): K = D A with D the forward difference operator, e1 = 1 / |w_l|,
 * e2 = D3 / |errorSmooth3_l| with D3 the central, backward and forward differences, and
 * e3 = fillInverseAbskWRobust (:1296-1322), 1 / max(|W_l(k)|, |W_l(k / 0.999)| / 0.999) with the
 * stretched spectrum of Wavelet::getCAmp(k, scale) (wavelet.cpp:337-394).  min_rel_thick: the
 * simbox's minimum relative thickness (the band test is f > low_cut min_rel_thick). */
int cg_build_kz_tables_variable_thickness(int nz, int nzp, double dt_ms, int ntheta, const float *wavelets,
                                          const float *A, float low_cut, float high_cut, float min_rel_thick,
                                          float *K, float *e1, float *e2, float *e3, int *invert);
/* FFTGrid::findClosestFactorableNumber (fftgrid.cpp:2630-2658). */
int cg_factorable(int leastint);
/* SeismicParametersHolder::MakeCircAutoCovPosDef (seismicparametersholder.cpp:661-767), in place on
 * the nine circular auto- and cross-covariance series circ_auto_cov[3][3][nzp] (host memory):
 * per frequency the 3 x 3 spectrum, scaled by the lag-0 standard deviations, has its eigenvalues
 * replaced by their absolute values, then the series are transformed back (x 1/nzp).  The
 * cross-covariances come in Hermitian pairs (c_jk(t) = c_kj(-t)), as the reference builds them,
 * so each spectrum is Hermitian.  CG_EINVAL if a lag-0 variance is not positive (the
 * reference's result is then NaN).  Host only. */
int cg_make_circ_auto_cov_pos_def(int nzp, float *circ_auto_cov);

#ifdef __cplusplus
}
#endif
#endif
