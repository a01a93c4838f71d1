// cg_post.h -- parameters of the posterior kernels (cg_post.hip), filled by cg_api.hip.
#pragma once

#include <hip/hip_runtime.h>

#include "cg_internal.h"

#include <cstdint>

namespace cg {

// Per-k_z plane tables, lower triangles at t = i (i + 1) / 2 + j.
template <typename R>
struct PlaneTab {
  struct C {
    R re, im;
  };
  C K[8][3];  // K = diag(kD conj W_l) A                (avoinversion.cpp:947-965)
  C S[8][3];  // K Sigma0
  C G[36];    // K Sigma0 K^H
  C Q[36];    // 0.5 (1 - wnc) Sigma0e_lm (e1_l e1_m* + e2_l e2_m*)   (avoinversion.cpp:1254-1266)
  R D[8];     // wnc Sigma0e_ll |e3_l|^2
};

struct PostParams {
  int nxp, nyp, nzp, cnxp, nt;
  long csize;
  float2 *mean[3];
  float2 *data[8];
  float2 *cov[6];
  const float2 *lam_grid;  // errCorr spectrum grid, or null
  const float *lam2d;      // [nyp][cnxp] when lam_grid is null
  const float *rxy;        // [nyp][cnxp] separable prior
  const float *rt;         // [nzp]
  const int *invert;       // [nzp]
  const PlaneTab<float> *tab32;
  const PlaneTab<double> *tab64;
  float s6f[6];            // priorVar0 entries 00 11 22 01 02 12 as float
  double s6d[6];
  int prior_mode;
  int want_residual;
  int precision;
  float cond_limit;
  unsigned long long *list_len;   // MIXED: number of wavenumbers deferred to FP64
  unsigned long long *stats;      // optional: [0,64) inverted, [64,128) Cholesky failures (slots summed on the host)
  unsigned *list;                // wavenumbers deferred to FP64 (capacity: inverted cells)
  // Instead of the list (when set): one bit per complex cell, bit idx of word idx / 64 set for a
  // deferred wavenumber; all zero on entry.  The FP64 pass clears what it consumes.
  unsigned long long *mask;
};

struct PostItem {
  PostParams p;
  unsigned block0;  // first block of this inversion in a batched launch
  unsigned bpp;     // blocks per k plane
};

int launch_posterior(cg_ctx *ctx, const PostParams &p);
// The allocation-free variant's fixed device buffer (cg_posterior_noalloc): one per device, a
// file-scope array of kNoAllocBytes; noalloc_scratch_init looks up its address on the current
// device once (at context creation), noalloc_scratch returns it (null before the lookup or if it
// failed).  Layout: counters [0, kNoAllocTables), per-call tables [kNoAllocTables, kNoAllocMask),
// the FP64 deferral bit mask [kNoAllocMask, kNoAllocBytes).
constexpr size_t kNoAllocTables = 4096;
constexpr size_t kNoAllocMaxNzp = 1536;             // k planes
constexpr size_t kNoAllocMaxPlane = size_t(1) << 21;  // complex cells per k plane
constexpr size_t kNoAllocMaxCells = size_t(1) << 28;  // complex cells per grid (mask bits)
constexpr size_t kNoAllocTableBytes =
    kNoAllocMaxNzp * (sizeof(PlaneTab<float>) + sizeof(PlaneTab<double>) + 2 * sizeof(int)) +
    2 * kNoAllocMaxPlane * sizeof(float) + 8 * 256;
constexpr size_t kNoAllocMask = kNoAllocTables + ((kNoAllocTableBytes + 255) / 256) * 256;
constexpr size_t kNoAllocBytes = kNoAllocMask + kNoAllocMaxCells / 8;
int noalloc_scratch_init(int device);
void *noalloc_scratch(int device);
// Clears n 64-bit words of device memory with a kernel.
int zero_words(cg_ctx *ctx, unsigned long long *d, size_t n);
int launch_posterior_batch(cg_ctx *ctx, int nt, bool sep, bool mixed, const PostItem *dev_items, int n,
                           unsigned blocks);

}  // namespace cg
