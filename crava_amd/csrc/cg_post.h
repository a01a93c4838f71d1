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

// The split-z posterior pass (cg_invert.hip): mean[] and data[] hold the spectra with the z
// transform's first stage done (plane n1 + N1 k2 of frequency index k2), cov[] receive the
// natural-layout covariance spectra, mean[] the posterior means with the inverse z transform's
// first stage done.
struct PostZArgs {
  PostParams p;
  int n2;                          // N2 (nzp = N1 N2)
  unsigned ctiles;                 // 64-cell tiles per plane
  unsigned *redo;                  // MIXED: tiles left to the FP64 pass (capacity: all tiles)
  unsigned long long *redo_len;    // MIXED: their number (zero on entry)
};
int launch_posterior_z(cg_ctx *ctx, const PostZArgs &z, int n1, bool mixed);
// A compiled split nzp = n1 n2 for this x length (even nxp only).
bool split_z_supported(int nxp, int nzp, int *n1, int *n2);
// x pass with z stage 1 (forward: src real grid -> grid) or inverse z stage 1 with the x c2r
// (grid in place, + add).
int fft_xz(cg_ctx *ctx, int forward, const float *src, float *grid, const float *add, int nxp, int nyp, int nzp,
           float scale);
struct ColourParams;
// Colouring (one realisation, or a pair: out2 / stream1) with the inverse z transform's stage 2,
// into the output grids' split layout (cg_invert.hip); the y pass and fft_xz follow.  mean_split
// (optional, [3]): posterior means in that layout, added.
int colour_split(cg_ctx *ctx, const ColourParams &p, int n1, int n2, bool pair,
                 const float2 *const *mean_split = nullptr);
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
