// cg_trace.h -- parameters of the per-trace z filters (cg_trace.hip), filled by cg_api.hip.
#pragma once

#include <hip/hip_runtime.h>

#include "cg_internal.h"

namespace cg {

struct ScaleWaveletParams {
  float2 *data;  // the real grid viewed as complex columns (two traces per column)
  int mode;      // CG_SCALE_WAVELET_*
  int nx, ny, nz, nxp, nyp, nzp, cnxp;
  double lz;                 // nz dt (ms)
  float dz;                  // dt (ms), the wavelet's sample interval
  float low_cut, high_cut, wnc, error_var;
  float wnorm;               // |w| in the time domain
  float norm_f3;             // band norm of the global wavelet's errorSmooth3 spectrum
  float modw;                // |w|^2 (nzp/2) 1000 nz / (lz nzp) / highCut
  const float2 *W;           // [nzp/2 + 1] wavelet spectrum (unnormalised, sign -1)
  const float2 *D3;          // [nzp/2 + 1] central(backward(forward)) difference spectrum
  const float *tau2;         // [nzp/2 + 1] |Re| of the reflection-coefficient covariance spectrum
  const double *rel_thick;   // [nyp][nxp]
  const float *shift;        // [ny][nx] or null
  const float *gain;         // [ny][nx] or null
  const float *stretch;      // [ny][nx] or null
};

// Taper (divide), forward z pass, per-trace factors, inverse z pass; taper_dev holds
// max(1 - dist^2, 0) for k = nz + 1 .. nzp - 1.
int scale_wavelet_filter(cg_ctx *ctx, const ScaleWaveletParams &p, const float *taper_dev);

// CommonData::FillInData (commondata.cpp:6451-6775) from SEG-Y traces: traces tapered in the
// guard zone (DATA grids, k_pack_traces) or trimmed of zeros and detrended (other grids,
// k_pack_detrend) and zero-padded, their length-nt spectra (the x pass), spread into length-mt
// spectra (k_spread_spectra), fine traces (the x pass inverse, scale 1 / nt), the grid's columns
// interpolated from them, the trend added back for other grids (k_fill_columns).
struct TraceInfo {  // per trace of a detrended (non-DATA) fill
  float z0;         // time of the first kept sample
  float first;      // trend at the first kept sample
  float inc2;       // trend increment per fine sample
  int n_trend;      // fine trend samples, 10 (n - 1) + 1; 0: no nonzero sample (a dead trace)
};
struct TraceFillParams {
  const float *samples;    // [ntraces][stride]
  const int *length;       // [ntraces]
  long stride;
  int ntraces;
  float dz_data;
  int ns;                  // samples tapered at each end: floor(smooth_length / dz_data)
  int nt, rnt, mt, rmt;    // rnt = 2 (nt / 2 + 1), mt = 10 nt, rmt = 2 (mt / 2 + 1)
  float *a;                // [ntraces][rnt]: tapered traces, then their spectra
  float *b;                // [ntraces][rmt]: fine spectra, then fine traces
  const int *column_trace;  // [nyp][rnxp]
  const float *z0_grid;     // [nyp][rnxp]
  const float *dz_grid;     // [nyp][rnxp]
  const float *z0_data;     // [ntraces]
  float dz_fine;
  float *grid;             // real grid, in-place layout
  int nz, nyp, nzp, rnxp;
  int detrend;             // non-DATA grid: k_pack_detrend's trimming and trend, RMISSING when dead
  int trim_shift;          // leading zeros trimmed move z0 (SEG-Y traces; not STORM columns)
  TraceInfo *info;         // [ntraces] when detrend
};
int pack_traces(cg_ctx *ctx, const TraceFillParams &p);
int pack_detrend(cg_ctx *ctx, const TraceFillParams &p);  // stride <= 1024
int spread_spectra(cg_ctx *ctx, const TraceFillParams &p);
int fill_columns(cg_ctx *ctx, const TraceFillParams &p);

}  // namespace cg
