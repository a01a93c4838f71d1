// cg_internal.h -- shared declarations of the HIP implementation behind include/cravagpu.h.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <set>
#include <string>
#include <vector>

#include "cravagpu.h"

namespace cg {

constexpr int kMaxStages = 24;
constexpr int kMaxTheta = 8;
constexpr int kThreads = 256;

// A 1D transform length factored into the radices the LDS engine implements.
struct FftPlan {
  int n;
  int nstages;
  int radix[kMaxStages];
};

bool make_plan(int n, FftPlan *plan);  // false if n has a prime factor above 7

// One grid pass in a batched launch (cg_fft.hip: k_fft_batch).  Item i owns the blocks
// [block0, block0 + nblocks) of the launch.
enum { FFT_ITEM_X_R2C = 0, FFT_ITEM_X_C2R = 1, FFT_ITEM_STRIDED = 2 };
struct FftItem {
  const float *src;  // x forward: real source (may equal grid)
  float *grid;       // x passes: the grid
  float2 *data;      // strided passes: complex view
  const float *add;  // x inverse: optional real grid added after scaling
  long S, inner, OS, rows, outer;
  unsigned tiles_per_outer, block0, nblocks;
  int nxp, n, tile;
  float scale;
  FftPlan plan;
  const float2 *tw, *twr;
};

// Descriptors of the compile-time-plan passes (cg_fft_fixed.h): one grid of a y/z pass, or one
// grid of an x pass.  In a list, item i owns the blocks [block0_i, block0_{i+1}) of the launch.
struct StridedArgs {
  float2 *data;
  long S, OS;             // element stride, slab stride (complex units)
  unsigned inner, ncols;  // contiguous columns per slab, columns in all
  unsigned block0;
  float scale;
};
struct XArgs {
  const float *src;  // forward: real source (may equal dst)
  float *dst;        // the grid
  const float *add;  // inverse: optional real grid added after scaling
  long rows;
  unsigned block0;
  float scale;
};

// Registry of compiled lengths: kind STRIDED (y/z, n = axis length), XE (even nxp, n = nxp/2),
// XO (odd nxp, n = nxp); `tile` is TI columns (STRIDED) or RW rows (XE/XO) per workgroup.
enum { FIX_STRIDED = 0, FIX_XE = 1, FIX_XO = 2, FIX_KINDS = 3 };
typedef int (*StridedFn)(cg_ctx *, int dir, const StridedArgs *one, const StridedArgs *items, int nitems,
                         unsigned nblocks, const float2 *tw);
typedef int (*XFn)(cg_ctx *, int forward, int add, const XArgs *one, const XArgs *items, int nitems,
                   unsigned nblocks, const float2 *tw, const float2 *twr);
struct ColourParams;
struct ColourItem;
typedef int (*ColourZFn)(cg_ctx *, const ColourParams *one, const ColourItem *items, int nitems, unsigned nblocks,
                         const float2 *tw);
typedef int (*ColourZ2Fn)(cg_ctx *, const ColourParams *one, const ColourItem *items, int nitems, unsigned nblocks,
                          const float2 *tw);
struct FixedEntry {
  int kind, n;
  int tile[2];  // the two tile variants (equal when the length has only one)
  StridedFn s[2];
  XFn x[2];
  ColourZFn cz;  // strided lengths: colouring fused with the inverse z pass
  int cz_waves;  // wavefronts of the fused kernel resident per CU
  int cz_tile;   // its columns per workgroup
  int tune;      // 0: variant 0 is kept (the autotuner's timing on scratch grids misjudges it)
  ColourZ2Fn cz2;  // strided lengths: two realisations per launch (fused colouring)
  int cz2_tile;    // its columns per workgroup
};
// Tile variant for a pass geometry: the autotuner's choice (cg_autotune) for this exact
// geometry, or the context's forced variant, else 0.  `geom` identifies the pass: strided
// (S, inner, ncols), x (rows, forward, add).
int fixed_variant(cg_ctx *ctx, const FixedEntry *e, long g0, long g1, long g2);

struct TimedLaunch {
  int kclass;
  hipEvent_t start, stop;
};

}  // namespace cg

struct cg_ctx {
  int device = 0;
  int num_cu = 1;  // compute units of the device
  hipStream_t stream = nullptr;
  std::string err;
  uint64_t budget = 0, in_use = 0, peak = 0;
  // per-call device staging for host tables, grown on demand
  void *tables = nullptr;
  size_t tables_bytes = 0;
  uint64_t *counters = nullptr;  // device counters for posterior/simulation statistics
  // pinned host staging of small tables; `stage_done` is recorded after the copy out of it
  void *stage_host = nullptr;
  size_t stage_host_bytes = 0;
  hipEvent_t stage_done = nullptr;
  // wavenumbers deferred to the FP64 posterior pass
  unsigned *list = nullptr;
  size_t list_bytes = 0;
  std::set<cg_grid *> grids;  // live grids; cg_ctx_destroy releases any left over
  // Grid and batch memory comes from the device's stream-ordered pool when it has one: frees
  // then wait for nothing but the context stream (a plain hipFree waits for the whole device,
  // which would serialise concurrently running batches of a server).
  bool async_alloc = false;
  std::set<cg_arena *> arenas;
  // side streams (default priority) and fork/join events of batched launches, created on the
  // first batch and reused
  static constexpr int kSide = 4;
  hipStream_t side[kSide] = {};
  hipEvent_t fork = nullptr;
  hipEvent_t join[kSide] = {};
  int fft_engine = CG_FFT_AUTO;
  int fft_variant = -1;  // -1: autotuned choice or 0; 0/1: forced tile variant
  int fused_colour = -1;  // -1: autotuned or occupancy rule; 0: never; 1: always (compiled nzp)
  bool profiling = false;
  std::vector<cg::TimedLaunch> timed;
  std::vector<hipEvent_t> event_pool;
  uint64_t launches[CG_K_COUNT] = {0};
  double kernel_ms[CG_K_COUNT] = {0};
};

struct cg_arena {
  cg_ctx *ctx = nullptr;
  char *base = nullptr;
  size_t bytes = 0, used = 0;
  int live = 0;  // grids carved and not yet destroyed
};

struct cg_grid {
  cg_ctx *ctx = nullptr;
  cg_arena *arena = nullptr;  // memory owner when carved from an arena
  int nx = 0, ny = 0, nz = 0, nxp = 0, nyp = 0, nzp = 0;
  int cnxp = 0, rnxp = 0;
  int cubetype = 0;
  int transformed = 0;
  size_t rsize = 0, csize = 0;  // floats, complex cells
  float *d = nullptr;
};

namespace cg {

int fail(cg_ctx *ctx, int code, const std::string &msg);
int hip_fail(cg_ctx *ctx, hipError_t e, const char *what);
const float2 *twiddle_table(cg_ctx *ctx, int n);
hipError_t dev_alloc(cg_ctx *ctx, void **p, size_t bytes);  // stream-ordered on ctx->stream when possible
void dev_free(cg_ctx *ctx, void *p);
void *table_space(cg_ctx *ctx, size_t bytes);
void launch_begin(cg_ctx *ctx, int kclass, hipEvent_t *start);
void launch_end(cg_ctx *ctx, int kclass, hipEvent_t start);

// FFT passes (cg_fft.hip).  All enqueue on ctx->stream and return a CG_* status.
int fft_x_r2c(cg_ctx *ctx, const float *src, float *grid, int nxp, long rows, float scale);
int fft_x_c2r(cg_ctx *ctx, float *grid, int nxp, long rows, float scale, const float *add);
int fft_strided(cg_ctx *ctx, float2 *data, int n, long elem_stride, long inner, long outer, long outer_stride,
                int dir, float scale);
// The same, timed as kernel class kclass (CG_K_*).
int fft_strided(cg_ctx *ctx, float2 *data, int n, long elem_stride, long inner, long outer, long outer_stride,
                int dir, float scale, int kclass);
int grid_fft_forward(cg_grid *g, const float *src);
// Batched passes: plan_fft_item completes an item whose geometry fields are set (kind x:
// src/grid/add/nxp/rows/scale; strided: data/n/S/inner/outer/OS/scale) and returns the LDS
// bytes it needs; launch_fft_batch runs items already copied to device memory.
int plan_fft_item(cg_ctx *ctx, int kind, FftItem *it, size_t *lds);
int launch_fft_batch(cg_ctx *ctx, int kind, int dir, const FftItem *dev_items, int nitems, unsigned total_blocks,
                     size_t lds, int kclass);
// Compile-time-plan passes (cg_fft_fixed.hip): return 1 and set *rc if the length has a plan.
int fixed_strided(cg_ctx *ctx, float2 *data, int n, long S, long inner, long outer, long OS, int dir, float scale,
                  int *rc);
int fixed_x(cg_ctx *ctx, int forward, const float *src, float *grid, int nxp, long rows, float scale, const float *add,
            int *rc);
const FixedEntry *fixed_find(int kind, int n);  // null if the length was not compiled
// Lists of one kind and length (device copies of the descriptors, block0 filled in).
int fixed_strided_list(cg_ctx *ctx, const FixedEntry *e, int v, int dir, const StridedArgs *dev_items, int nitems,
                       unsigned nblocks);
int fixed_x_list(cg_ctx *ctx, const FixedEntry *e, int v, int forward, int add, const XArgs *dev_items, int nitems,
                 unsigned nblocks, int nxp);
int autotune_geometry(cg_ctx *ctx, int nxp, int nyp, int nzp);
// Colouring of one realisation fused with the inverse z pass of its three grids (counter noise
// only); returns 1 and sets *rc if it ran: nzp has a compiled plan and the fused kernel is the
// faster choice for this geometry (autotuned, else when it keeps at least 12 wavefronts per CU
// resident; with fewer the colouring's loads are poorly hidden).
int fixed_colour_zinv(cg_ctx *ctx, const ColourParams &p, int nzp, long plane, int *rc);
// The same for two realisations of one posterior (p.out2 / p.stream1 set): one launch.
int fixed_colour_zinv2(cg_ctx *ctx, const ColourParams &p, int nzp, long plane, int *rc);
// The paired kernel's entry for an nzp x plane geometry (null where pairs are not fused), and a
// list launch of pairs (even first realisations) of that entry's length.
const FixedEntry *fixed_colour2_entry(cg_ctx *ctx, int nzp, long plane);
int fixed_colour_zinv2_list(cg_ctx *ctx, const FixedEntry *e, const ColourItem *dev_items, int nitems,
                            unsigned nblocks);
// The compiled entry whose fused colouring a realisation of nzp planes of `plane` cells would use
// (same choice as fixed_colour_zinv), or null.
const FixedEntry *fixed_colour_entry(cg_ctx *ctx, int nzp, long plane);
// Fused colouring of a list of realisations of one nzp (descriptors in device memory).
int fixed_colour_zinv_list(cg_ctx *ctx, const FixedEntry *e, const ColourItem *dev_items, int nitems,
                           unsigned nblocks);
// grid_fft_inverse without the z pass (it was fused into the producer).
int grid_fft_inverse_xy(cg_grid *g, const float *add_real);
int grid_fft_inverse(cg_grid *g, const float *add_real);

// Elementwise kernels (cg_noise.hip).
int scale_real(cg_ctx *ctx, float *d, size_t n, float s);
// The reference's noise stream (cg_refrng.hip).
int fill_noise_reference(cg_ctx *ctx, float2 *out, int nxp, int nyp, int nzp, uint32_t *seed);
int fill_constant(cg_ctx *ctx, float *d, size_t n, int nxp, int rnxp, float v);
int add_real(cg_ctx *ctx, float *d, const float *s, size_t n);
// Padding fills (cg_pad.hip): the background cube into the padded grid; per-trace energies;
// trace averages (op = trace << 4 | neighbour mask: 1 left, 2 right, 4 up, 8 down), either ops
// of one dependency level (one workgroup each) or a run to apply in order (one workgroup);
// the tapered extrapolation outside a rectangle of box traces (tables per axis).
int fill_padded(cg_ctx *ctx, float *rv, const float *cube, int nx, int ny, int nz, int nxp, int nyp, int nzp);
int trace_energy(cg_ctx *ctx, const float *rv, float *energy, int nx, int ny, int nz, int rnxp, int nyp);
int trace_ops(cg_ctx *ctx, float *rv, const unsigned *ops, int nops, bool one_level, int nx, int rnxp, int nyp,
              int nzp);
int extrapolate(cg_ctx *ctx, float *rv, const int *ref_i, const float *dist_x, const int *ref_j, const float *dist_y,
                const float *dist_z, int nxp, int nyp, int nzp, int imin, int imax, int jmin, int jmax);
// Prior / error correlation grids (cg_pad.hip); cxy [nyp][nxp], ct [nzp] in device memory.
int fill_param_corr(cg_ctx *ctx, float *rv, const float *cxy, const float *ct, int nxp, int nyp, int nzp, float gi,
                    float gj, float scale);
int fill_err_corr(cg_ctx *ctx, float *rv, const float *cxy, int nxp, int nyp, int nzp, float gi, float gj);
// The forward model's trace taper (cg_pad.hip); cube [nz][ny][nx] in device memory.
int fill_trace_taper(cg_ctx *ctx, float *rv, const float *cube, int nx, int ny, int nz, int nxp, int nyp, int nzp);

}  // namespace cg

// XCD-aware block order: MI355X dispatches consecutive workgroups round-robin over its 8
// XCDs, each with its own L2.  This bijection gives every XCD a contiguous range of logical
// tiles, so tiles that share cache lines stay in one L2.
__device__ inline unsigned cg_xcd_remap(unsigned b, unsigned nb)
{
  const unsigned q = nb / 8u, r = nb % 8u;
  const unsigned xcd = b % 8u, local = b / 8u;
  return xcd * q + (xcd < r ? xcd : r) + local;
}
