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

}  // namespace cg
