// cg_api.hip -- the C ABI of include/cravagpu.h on top of the HIP kernels.
//
// Contexts replace CRAVA's process-global FFTGrid statics (fftgrid.cpp:2661-2668): each owns
// a device, a HIP stream, a device-memory budget, twiddle tables and staging buffers.  Grids
// are device-resident in the reference's in-place FFTW layout.
#include <algorithm>
#include <mutex>
#include <cmath>
#include <complex>
#include <cstring>
#include <new>

#include "cg_internal.h"
#include "cg_noise.h"
#include "cg_post.h"
#include "cg_trace.h"

namespace {

constexpr float kRMissing = -99999.0f;  // RMISSING, definitions.h:35

size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

// Offsets of the parts of one device allocation (256-byte aligned).
struct Layout {
  size_t size = 0;
  size_t add(size_t bytes)
  {
    const size_t off = size;
    size += align_up(bytes ? bytes : 1, 256);
    return off;
  }
};

struct DeviceGuard {
  explicit DeviceGuard(int dev) { (void)hipSetDevice(dev); }
};


// Pinned staging of small host tables into one device buffer on the context stream.
int stage_tables(cg_ctx *ctx, const std::vector<std::pair<const void *, size_t>> &parts, std::vector<void *> &dptrs)
{
  size_t total = 0;
  for (auto &p : parts) total += align_up(p.second, 256);
  if (total == 0) total = 256;
  if (!ctx->stage_done) {
    hipError_t e = hipEventCreateWithFlags(&ctx->stage_done, hipEventDisableTiming);
    if (e != hipSuccess) return cg::hip_fail(ctx, e, "hipEventCreate");
  } else {
    hipError_t e = hipEventSynchronize(ctx->stage_done);  // previous copy has left the buffer
    if (e != hipSuccess) return cg::hip_fail(ctx, e, "hipEventSynchronize");
  }
  if (ctx->stage_host_bytes < total) {
    if (ctx->stage_host) (void)hipHostFree(ctx->stage_host);
    ctx->stage_host = nullptr;
    ctx->stage_host_bytes = 0;
    hipError_t e = hipHostMalloc(&ctx->stage_host, total, hipHostMallocDefault);
    if (e != hipSuccess) return cg::hip_fail(ctx, e, "hipHostMalloc(staging)");
    ctx->stage_host_bytes = total;
  }
  char *dev = static_cast<char *>(cg::table_space(ctx, total));
  if (!dev) return CG_ENOMEM;
  char *h = static_cast<char *>(ctx->stage_host);
  size_t off = 0;
  dptrs.clear();
  for (auto &p : parts) {
    if (p.second) std::memcpy(h + off, p.first, p.second);
    dptrs.push_back(p.second ? dev + off : nullptr);
    off += align_up(p.second, 256);
  }
  hipError_t e = hipMemcpyAsync(dev, h, off ? off : 256, hipMemcpyHostToDevice, ctx->stream);
  if (e != hipSuccess) return cg::hip_fail(ctx, e, "hipMemcpyAsync(tables)");
  e = hipEventRecord(ctx->stage_done, ctx->stream);
  return cg::hip_fail(ctx, e, "hipEventRecord");
}

bool valid_type(int t) { return t == CG_CT_DATA || t == CG_CT_PARAMETER || t == CG_CT_COVARIANCE; }

bool same_shape(const cg_grid *a, const cg_grid *b)
{
  return a->nxp == b->nxp && a->nyp == b->nyp && a->nzp == b->nzp;
}

// Per-plane tables of the posterior (see cg_post.hip for the identities).
template <typename R>
void build_plane_tables(const cg_post_args *a, int nzp, std::vector<cg::PlaneTab<R>> &tab)
{
  typedef std::complex<double> C;
  const int nt = a->ntheta;
  const double *pv = a->prior_var0;
  const double S0[3][3] = {{pv[0], pv[1], pv[2]}, {pv[3], pv[4], pv[5]}, {pv[6], pv[7], pv[8]}};
  const double wnc = a->wnc;
  tab.assign(nzp, cg::PlaneTab<R>());
  for (int k = 0; k < nzp; ++k) {
    cg::PlaneTab<R> &T = tab[k];
    std::memset(&T, 0, sizeof(T));
    C K[8][3], S[8][3], e1[8], e2[8];
    for (int l = 0; l < nt; ++l) {
      for (int j = 0; j < 3; ++j) {
        const float *kp = a->K + 2 * ((k * nt + l) * 3 + j);
        K[l][j] = C(kp[0], kp[1]);
        T.K[l][j].re = (R)kp[0];
        T.K[l][j].im = (R)kp[1];
      }
      e1[l] = C(a->e1[2 * (k * nt + l)], a->e1[2 * (k * nt + l) + 1]);
      e2[l] = C(a->e2[2 * (k * nt + l)], a->e2[2 * (k * nt + l) + 1]);
      const double e3r = a->e3[2 * (k * nt + l)];
      T.D[l] = (R)(wnc * a->sigma0e[l * nt + l] * e3r * e3r);
    }
    for (int l = 0; l < nt; ++l)
      for (int b = 0; b < 3; ++b) {
        C x = 0;
        for (int c = 0; c < 3; ++c) x += K[l][c] * S0[c][b];
        S[l][b] = x;
        T.S[l][b].re = (R)x.real();
        T.S[l][b].im = (R)x.imag();
      }
    for (int l = 0; l < nt; ++l)
      for (int m = 0; m <= l; ++m) {
        C g = 0;
        for (int c = 0; c < 3; ++c) g += S[l][c] * std::conj(K[m][c]);
        const int t = l * (l + 1) / 2 + m;
        T.G[t].re = (R)g.real();
        T.G[t].im = (R)g.imag();
        C q = 0.5 * (1.0 - wnc) * a->sigma0e[l * nt + m] * (e1[l] * std::conj(e1[m]) + e2[l] * std::conj(e2[m]));
        T.Q[t].re = (R)q.real();
        T.Q[t].im = (l == m) ? (R)0 : (R)q.imag();
      }
  }
}

// host DFT bin of a real sequence via a cosine/sine table (no trig in the loop)
void dft_bin(const std::vector<double> &x, const std::vector<double> &ct, const std::vector<double> &st, int k,
             double &re, double &im)
{
  const int n = (int)x.size();
  double sr = 0, si = 0;
  long e = 0;
  for (int t = 0; t < n; ++t) {
    sr += x[t] * ct[e];
    si -= x[t] * st[e];
    e += k;
    if (e >= n) e -= n;
  }
  re = sr;
  im = si;
}

float find_norm(const std::vector<float> &r)
{
  double n2 = 0;
  for (float v : r) n2 += (double)(v * v);
  return (float)std::sqrt(n2);
}

}  // namespace

namespace cg {

int fail(cg_ctx *ctx, int code, const std::string &msg)
{
  if (ctx) ctx->err = msg;
  return code;
}

int hip_fail(cg_ctx *ctx, hipError_t e, const char *what)
{
  if (e == hipSuccess) return CG_OK;
  if (ctx) ctx->err = std::string(what) + ": " + hipGetErrorString(e);
  return CG_EHIP;
}

hipError_t dev_alloc(cg_ctx *ctx, void **p, size_t bytes)
{
  return ctx->async_alloc ? hipMallocAsync(p, bytes, ctx->stream) : hipMalloc(p, bytes);
}

void dev_free(cg_ctx *ctx, void *p)
{
  if (!p) return;
  if (ctx->async_alloc) (void)hipFreeAsync(p, ctx->stream);
  else (void)hipFree(p);
}

// Twiddle tables exp(-2 pi i m / n), m < n, shared by every context of a device (immutable,
// built once per length and kept until the process ends).
const float2 *twiddle_table(cg_ctx *ctx, int n)
{
  static std::mutex mu;
  static std::map<std::pair<int, int>, float2 *> tables;
  std::lock_guard<std::mutex> lock(mu);
  const auto key = std::make_pair(ctx->device, n);
  auto it = tables.find(key);
  if (it != tables.end()) return it->second;
  std::vector<float2> h(n);
  for (int m = 0; m < n; ++m) {
    const double a = -2.0 * M_PI * (double)m / (double)n;
    h[m] = make_float2((float)std::cos(a), (float)std::sin(a));
  }
  float2 *d = nullptr;
  hipError_t e = hipMalloc(&d, sizeof(float2) * n);
  if (e != hipSuccess) {
    hip_fail(ctx, e, "hipMalloc(twiddles)");
    return nullptr;
  }
  e = hipMemcpy(d, h.data(), sizeof(float2) * n, hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    hip_fail(ctx, e, "hipMemcpy(twiddles)");
    (void)hipFree(d);
    return nullptr;
  }
  tables[key] = d;
  return d;
}

void *table_space(cg_ctx *ctx, size_t bytes)
{
  if (ctx->tables_bytes >= bytes) return ctx->tables;
  if (ctx->tables) {
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipFree(ctx->tables);
    ctx->in_use -= ctx->tables_bytes;
    ctx->tables = nullptr;
    ctx->tables_bytes = 0;
  }
  const size_t want = align_up(bytes, 1 << 20);
  hipError_t e = hipMalloc(&ctx->tables, want);
  if (e != hipSuccess) {
    hip_fail(ctx, e, "hipMalloc(tables)");
    ctx->tables = nullptr;
    return nullptr;
  }
  ctx->tables_bytes = want;
  ctx->in_use += want;
  ctx->peak = std::max(ctx->peak, ctx->in_use);
  return ctx->tables;
}

void launch_begin(cg_ctx *ctx, int kclass, hipEvent_t *start)
{
  *start = nullptr;
  if (!ctx->profiling) return;
  if (hipEventCreate(start) != hipSuccess) {
    *start = nullptr;
    return;
  }
  (void)hipEventRecord(*start, ctx->stream);
  (void)kclass;
}

void launch_end(cg_ctx *ctx, int kclass, hipEvent_t start)
{
  if (!ctx->profiling || !start) return;
  hipEvent_t stop;
  if (hipEventCreate(&stop) != hipSuccess) return;
  (void)hipEventRecord(stop, ctx->stream);
  ctx->timed.push_back(TimedLaunch{kclass, start, stop});
}

}  // namespace cg

// ============================================================================ ABI
extern "C" {

int cg_ctx_create(cg_ctx **out, int device, uint64_t hbm_budget_bytes)
{
  if (!out) return CG_EINVAL;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return CG_EHIP;
  if (device < 0 || device >= ndev) return CG_EINVAL;
  cg_ctx *ctx = new (std::nothrow) cg_ctx();
  if (!ctx) return CG_ENOMEM;
  ctx->device = device;
  ctx->budget = hbm_budget_bytes;
  DeviceGuard g(device);
  hipError_t e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipMalloc(&ctx->counters, 256 * sizeof(uint64_t));
  if (e == hipSuccess) e = hipDeviceGetAttribute(&ctx->num_cu, hipDeviceAttributeMultiprocessorCount, device);
  // the allocation-free posterior's buffer: its address on this device, looked up once per device
  // (a failure is reported by cg_posterior_noalloc, the only user)
  if (e == hipSuccess) (void)cg::noalloc_scratch_init(device);
  if (e == hipSuccess) {
    int pools = 0;
    if (hipDeviceGetAttribute(&pools, hipDeviceAttributeMemoryPoolsSupported, device) == hipSuccess && pools) {
      hipMemPool_t pool;
      if (hipDeviceGetDefaultMemPool(&pool, device) == hipSuccess) {
        // keep freed blocks in the pool instead of unmapping them at every synchronisation
        uint64_t keep = UINT64_MAX;
        ctx->async_alloc = hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep) == hipSuccess;
      }
    }
  }
  if (e != hipSuccess) {
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return CG_EHIP;
  }
  *out = ctx;
  return CG_OK;
}

int cg_ctx_destroy(cg_ctx *ctx)
{
  if (!ctx) return CG_EINVAL;
  DeviceGuard g(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  for (auto &t : ctx->timed) {
    (void)hipEventDestroy(t.start);
    (void)hipEventDestroy(t.stop);
  }
  for (cg_grid *gr : ctx->grids) {
    if (!gr->arena) cg::dev_free(ctx, gr->d);
    delete gr;
  }
  ctx->grids.clear();
  for (cg_arena *a : ctx->arenas) {
    cg::dev_free(ctx, a->base);
    delete a;
  }
  ctx->arenas.clear();
  (void)hipStreamSynchronize(ctx->stream);
  for (int k = 0; k < cg_ctx::kSide; ++k) {
    if (ctx->side[k]) {
      (void)hipStreamSynchronize(ctx->side[k]);
      (void)hipStreamDestroy(ctx->side[k]);
    }
    if (ctx->join[k]) (void)hipEventDestroy(ctx->join[k]);
  }
  if (ctx->fork) (void)hipEventDestroy(ctx->fork);
  if (ctx->tables) (void)hipFree(ctx->tables);
  if (ctx->counters) (void)hipFree(ctx->counters);
  if (ctx->stage_host) (void)hipHostFree(ctx->stage_host);
  if (ctx->stage_done) (void)hipEventDestroy(ctx->stage_done);
  if (ctx->list) (void)hipFree(ctx->list);
  (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return CG_OK;
}

const char *cg_last_error(const cg_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int cg_ctx_sync(cg_ctx *ctx)
{
  if (!ctx) return CG_EINVAL;
  DeviceGuard g(ctx->device);
  return cg::hip_fail(ctx, hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
}

int cg_ctx_memory(const cg_ctx *ctx, uint64_t *in_use, uint64_t *peak)
{
  if (!ctx) return CG_EINVAL;
  if (in_use) *in_use = ctx->in_use;
  if (peak) *peak = ctx->peak;
  return CG_OK;
}

int cg_ctx_stream(const cg_ctx *ctx, void **stream)
{
  if (!ctx || !stream) return CG_EINVAL;
  *stream = (void *)ctx->stream;
  return CG_OK;
}

int cg_ctx_set_fft_variant(cg_ctx *ctx, int variant)
{
  if (!ctx || variant < -1 || variant > 1) return CG_EINVAL;
  ctx->fft_variant = variant;
  return CG_OK;
}

int cg_ctx_set_fused_colour(cg_ctx *ctx, int mode)
{
  if (!ctx || mode < -1 || mode > 1) return CG_EINVAL;
  ctx->fused_colour = mode;
  return CG_OK;
}

int cg_autotune(cg_ctx *ctx, int nxp, int nyp, int nzp)
{
  if (!ctx || nxp < 1 || nyp < 1 || nzp < 1) return CG_EINVAL;
  DeviceGuard dg(ctx->device);
  return cg::autotune_geometry(ctx, nxp, nyp, nzp);
}

int cg_ctx_set_fft_engine(cg_ctx *ctx, int engine)
{
  if (!ctx || (engine != CG_FFT_AUTO && engine != CG_FFT_GENERIC)) return CG_EINVAL;
  ctx->fft_engine = engine;
  return CG_OK;
}

int cg_ctx_set_profiling(cg_ctx *ctx, int enable)
{
  if (!ctx) return CG_EINVAL;
  ctx->profiling = enable != 0;
  return CG_OK;
}

static int collect_times(cg_ctx *ctx)
{
  DeviceGuard g(ctx->device);
  hipError_t e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) return cg::hip_fail(ctx, e, "hipStreamSynchronize");
  for (auto &t : ctx->timed) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, t.start, t.stop) == hipSuccess) {
      ctx->launches[t.kclass] += 1;
      ctx->kernel_ms[t.kclass] += ms;
    }
    (void)hipEventDestroy(t.start);
    (void)hipEventDestroy(t.stop);
  }
  ctx->timed.clear();
  return CG_OK;
}

int cg_ctx_kernel_time(cg_ctx *ctx, int kclass, uint64_t *launches, double *total_ms)
{
  if (!ctx || kclass < 0 || kclass >= CG_K_COUNT) return CG_EINVAL;
  int rc = collect_times(ctx);
  if (rc) return rc;
  if (launches) *launches = ctx->launches[kclass];
  if (total_ms) *total_ms = ctx->kernel_ms[kclass];
  return CG_OK;
}

int cg_ctx_reset_kernel_time(cg_ctx *ctx)
{
  if (!ctx) return CG_EINVAL;
  int rc = collect_times(ctx);
  for (int i = 0; i < CG_K_COUNT; ++i) {
    ctx->launches[i] = 0;
    ctx->kernel_ms[i] = 0;
  }
  return rc;
}

// ---------------------------------------------------------------------------- grids
namespace {

size_t grid_alloc_bytes(int nxp, int nyp, int nzp)
{
  return align_up((size_t)2 * (size_t)(nxp / 2 + 1) * (size_t)nyp * (size_t)nzp * sizeof(float), 256);
}

int grid_create(cg_ctx *ctx, cg_arena *arena, int nx, int ny, int nz, int nxp, int nyp, int nzp, int cubetype,
                cg_grid **out)
{
  if (!ctx || !out) return CG_EINVAL;
  *out = nullptr;
  if (nxp < 1 || nyp < 1 || nzp < 1 || nx < 1 || ny < 1 || nz < 1 || nx > nxp || ny > nyp || nz > nzp)
    return cg::fail(ctx, CG_EINVAL, "cg_grid_create: need 1 <= n <= n_padded on every axis");
  if (!valid_type(cubetype)) return cg::fail(ctx, CG_EINVAL, "cg_grid_create: unknown cube type");
  if ((long)nxp * nyp * nzp > (1L << 33)) return cg::fail(ctx, CG_EINVAL, "cg_grid_create: grid too large");
  cg_grid *g = new (std::nothrow) cg_grid();
  if (!g) return CG_ENOMEM;
  g->ctx = ctx;
  g->nx = nx; g->ny = ny; g->nz = nz;
  g->nxp = nxp; g->nyp = nyp; g->nzp = nzp;
  g->cnxp = nxp / 2 + 1;
  g->rnxp = 2 * g->cnxp;
  g->cubetype = cubetype;
  g->csize = (size_t)g->cnxp * nyp * nzp;
  g->rsize = 2 * g->csize;
  const size_t bytes = g->rsize * sizeof(float);
  DeviceGuard dg(ctx->device);
  hipError_t e;
  if (arena) {
    const size_t need = grid_alloc_bytes(nxp, nyp, nzp);
    if (arena->used + need > arena->bytes) {
      delete g;
      return cg::fail(ctx, CG_ENOMEM, "cg_grid_create_in: arena full");
    }
    g->d = reinterpret_cast<float *>(arena->base + arena->used);
    g->arena = arena;
    arena->used += need;
    arena->live += 1;
  } else {
    if (ctx->budget && ctx->in_use + bytes > ctx->budget) {
      delete g;
      return cg::fail(ctx, CG_ENOMEM, "cg_grid_create: HBM budget of the context exceeded");
    }
    e = cg::dev_alloc(ctx, reinterpret_cast<void **>(&g->d), bytes);
    if (e != hipSuccess) {
      delete g;
      return cg::fail(ctx, CG_ENOMEM, std::string("cg_grid_create: allocation: ") + hipGetErrorString(e));
    }
    ctx->in_use += bytes;
    ctx->peak = std::max(ctx->peak, ctx->in_use);
  }
  e = hipMemsetAsync(g->d, 0, bytes, ctx->stream);
  if (e != hipSuccess) {
    if (arena) {
      arena->live -= 1;
    } else {
      cg::dev_free(ctx, g->d);
      ctx->in_use -= bytes;
    }
    delete g;
    return cg::hip_fail(ctx, e, "hipMemsetAsync");
  }
  ctx->grids.insert(g);
  *out = g;
  return CG_OK;
}

}  // namespace

int cg_grid_create(cg_ctx *ctx, int nx, int ny, int nz, int nxp, int nyp, int nzp, int cubetype, cg_grid **out)
{
  return grid_create(ctx, nullptr, nx, ny, nz, nxp, nyp, nzp, cubetype, out);
}

int cg_grid_create_in(cg_arena *arena, int nx, int ny, int nz, int nxp, int nyp, int nzp, int cubetype,
                      cg_grid **out)
{
  if (!arena) return CG_EINVAL;
  return grid_create(arena->ctx, arena, nx, ny, nz, nxp, nyp, nzp, cubetype, out);
}

uint64_t cg_arena_grid_bytes(int nxp, int nyp, int nzp)
{
  if (nxp < 1 || nyp < 1 || nzp < 1) return 0;
  return grid_alloc_bytes(nxp, nyp, nzp);
}

int cg_arena_create(cg_ctx *ctx, uint64_t bytes, cg_arena **out)
{
  if (!ctx || !out || bytes == 0) return CG_EINVAL;
  *out = nullptr;
  if (ctx->budget && ctx->in_use + bytes > ctx->budget)
    return cg::fail(ctx, CG_ENOMEM, "cg_arena_create: HBM budget of the context exceeded");
  DeviceGuard dg(ctx->device);
  cg_arena *a = new (std::nothrow) cg_arena();
  if (!a) return CG_ENOMEM;
  a->ctx = ctx;
  a->bytes = bytes;
  hipError_t e = cg::dev_alloc(ctx, reinterpret_cast<void **>(&a->base), bytes);
  if (e != hipSuccess) {
    delete a;
    return cg::fail(ctx, CG_ENOMEM, std::string("cg_arena_create: allocation: ") + hipGetErrorString(e));
  }
  ctx->in_use += bytes;
  ctx->peak = std::max(ctx->peak, ctx->in_use);
  ctx->arenas.insert(a);
  *out = a;
  return CG_OK;
}

int cg_arena_destroy(cg_arena *a)
{
  if (!a) return CG_EINVAL;
  cg_ctx *ctx = a->ctx;
  if (a->live) return cg::fail(ctx, CG_ESTATE, "cg_arena_destroy: grids of the arena are still alive");
  DeviceGuard dg(ctx->device);
  cg::dev_free(ctx, a->base);
  ctx->in_use -= a->bytes;
  ctx->arenas.erase(a);
  delete a;
  return CG_OK;
}

int cg_grid_destroy(cg_grid *g)
{
  if (!g) return CG_EINVAL;
  cg_ctx *ctx = g->ctx;
  DeviceGuard dg(ctx->device);
  if (g->arena) {
    g->arena->live -= 1;  // the arena owns the memory
  } else {
    if (!ctx->async_alloc) (void)hipStreamSynchronize(ctx->stream);
    cg::dev_free(ctx, g->d);  // stream-ordered: after the work already enqueued on ctx->stream
    ctx->in_use -= g->rsize * sizeof(float);
  }
  ctx->grids.erase(g);
  delete g;
  return CG_OK;
}

int cg_grid_set_type(cg_grid *g, int cubetype)
{
  if (!g) return CG_EINVAL;
  if (!valid_type(cubetype)) return cg::fail(g->ctx, CG_EINVAL, "unknown cube type");
  g->cubetype = cubetype;
  return CG_OK;
}

int cg_grid_is_transformed(const cg_grid *g, int *t)
{
  if (!g || !t) return CG_EINVAL;
  *t = g->transformed;
  return CG_OK;
}

int cg_grid_dims(const cg_grid *g, int dims[6])
{
  if (!g || !dims) return CG_EINVAL;
  dims[0] = g->nx; dims[1] = g->ny; dims[2] = g->nz;
  dims[3] = g->nxp; dims[4] = g->nyp; dims[5] = g->nzp;
  return CG_OK;
}

int cg_grid_device_ptr(cg_grid *g, void **p)
{
  if (!g || !p) return CG_EINVAL;
  *p = g->d;
  return CG_OK;
}

static int copy_sync(cg_grid *g, void *dst, const void *src, size_t bytes, hipMemcpyKind kind)
{
  DeviceGuard dg(g->ctx->device);
  hipError_t e = hipMemcpyAsync(dst, src, bytes, kind, g->ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(g->ctx->stream);
  return cg::hip_fail(g->ctx, e, "grid copy");
}

int cg_grid_upload_real(cg_grid *g, const float *h)
{
  if (!g || !h) return CG_EINVAL;
  int rc = copy_sync(g, g->d, h, g->rsize * sizeof(float), hipMemcpyHostToDevice);
  if (rc == CG_OK) g->transformed = 0;
  return rc;
}

int cg_grid_upload_complex(cg_grid *g, const float *h)
{
  if (!g || !h) return CG_EINVAL;
  int rc = copy_sync(g, g->d, h, g->rsize * sizeof(float), hipMemcpyHostToDevice);
  if (rc == CG_OK) g->transformed = 1;
  return rc;
}

int cg_grid_download_real(cg_grid *g, float *h)
{
  if (!g || !h) return CG_EINVAL;
  if (g->transformed) return cg::fail(g->ctx, CG_ESTATE, "download_real: grid is in the Fourier domain");
  return copy_sync(g, h, g->d, g->rsize * sizeof(float), hipMemcpyDeviceToHost);
}

int cg_grid_download_complex(cg_grid *g, float *h)
{
  if (!g || !h) return CG_EINVAL;
  if (!g->transformed) return cg::fail(g->ctx, CG_ESTATE, "download_complex: grid is in the real domain");
  return copy_sync(g, h, g->d, g->rsize * sizeof(float), hipMemcpyDeviceToHost);
}

static bool in_box(const cg_grid *g, int i, int j, int k, int ext)
{
  if (i < 0 || j < 0 || k < 0) return false;
  return ext ? (i < g->nxp && j < g->nyp && k < g->nzp) : (i < g->nx && j < g->ny && k < g->nz);
}

int cg_grid_get_complex(cg_grid *g, int i, int j, int k, int ext, float v[2])
{
  if (!g || !v) return CG_EINVAL;
  if (!g->transformed) return cg::fail(g->ctx, CG_ESTATE, "get_complex: grid is in the real domain");
  // getComplexValue tests i < nx (or nxp), not i < cnxp, and then reads index
  // i + cnxp*j + cnxp*nyp*k: for cnxp <= i it reads into the next row, as the reference
  // does.  Indexes past the end of the buffer give RMISSING.
  const size_t idx = (size_t)i + (size_t)g->cnxp * ((size_t)j + (size_t)g->nyp * k);
  if (!in_box(g, i, j, k, ext) || idx >= g->csize) {
    v[0] = v[1] = kRMissing;
    return CG_OK;
  }
  return copy_sync(g, v, g->d + 2 * idx, 2 * sizeof(float), hipMemcpyDeviceToHost);
}

int cg_grid_set_complex(cg_grid *g, int i, int j, int k, int ext, const float v[2])
{
  if (!g || !v) return CG_EINVAL;
  if (!g->transformed) return cg::fail(g->ctx, CG_ESTATE, "set_complex: grid is in the real domain");
  const size_t idx = (size_t)i + (size_t)g->cnxp * ((size_t)j + (size_t)g->nyp * k);
  if (!in_box(g, i, j, k, ext) || idx >= g->csize) return 1;
  return copy_sync(g, g->d + 2 * idx, v, 2 * sizeof(float), hipMemcpyHostToDevice);
}

int cg_grid_get_real_value(cg_grid *g, int i, int j, int k, int ext, float *v)
{
  if (!g || !v) return CG_EINVAL;
  if (g->transformed) return cg::fail(g->ctx, CG_ESTATE, "get_real_value: grid is in the Fourier domain");
  if (!in_box(g, i, j, k, ext)) {
    *v = kRMissing;
    return CG_OK;
  }
  const size_t idx = (size_t)i + (size_t)g->rnxp * ((size_t)j + (size_t)g->nyp * k);
  return copy_sync(g, v, g->d + idx, sizeof(float), hipMemcpyDeviceToHost);
}

int cg_grid_get_real_trace(cg_grid *g, int i, int j, int add_padding, float *v)
{
  if (!g || !v) return CG_EINVAL;
  if (g->transformed) return cg::fail(g->ctx, CG_ESTATE, "get_real_trace: grid is in the Fourier domain");
  const int n = add_padding ? g->nzp : g->nz;
  // getRealValue(i, j, k) tests the unpadded box, so the padding part of a padded trace and every
  // sample of a column outside the box read RMISSING
  for (int k = 0; k < n; ++k) v[k] = kRMissing;
  if (!in_box(g, i, j, 0, 0)) return CG_OK;
  DeviceGuard dg(g->ctx->device);
  const size_t pitch = sizeof(float) * (size_t)g->rnxp * g->nyp;
  hipError_t e = hipMemcpy2DAsync(v, sizeof(float), g->d + (size_t)i + (size_t)g->rnxp * j, pitch, sizeof(float),
                                  (size_t)g->nz, hipMemcpyDeviceToHost, g->ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(g->ctx->stream);
  return cg::hip_fail(g->ctx, e, "get_real_trace");
}

int cg_fft_r2c_inplace(cg_grid *g)
{
  if (!g) return CG_EINVAL;
  if (g->transformed) return cg::fail(g->ctx, CG_ESTATE, "fftInPlace: grid already in the Fourier domain");
  DeviceGuard dg(g->ctx->device);
  int rc = cg::grid_fft_forward(g, nullptr);
  if (rc == CG_OK) g->transformed = 1;
  return rc;
}

int cg_fft_r2c(const cg_grid *src, cg_grid *dst)
{
  if (!src || !dst) return CG_EINVAL;
  if (src->ctx != dst->ctx || !same_shape(src, dst)) return cg::fail(dst->ctx, CG_EINVAL, "fft_r2c: grids differ");
  if (src->transformed) return cg::fail(dst->ctx, CG_ESTATE, "fft_r2c: source is in the Fourier domain");
  DeviceGuard dg(dst->ctx->device);
  dst->cubetype = src->cubetype;
  int rc = cg::grid_fft_forward(dst, src == dst ? nullptr : src->d);
  if (rc == CG_OK) dst->transformed = 1;
  return rc;
}

int cg_fft_c2r_inplace(cg_grid *g)
{
  if (!g) return CG_EINVAL;
  if (!g->transformed) return cg::fail(g->ctx, CG_ESTATE, "invFFTInPlace: grid already in the real domain");
  DeviceGuard dg(g->ctx->device);
  int rc = cg::grid_fft_inverse(g, nullptr);
  if (rc == CG_OK) g->transformed = 0;
  return rc;
}

int cg_grid_add(cg_grid *dst, const cg_grid *src)
{
  if (!dst || !src) return CG_EINVAL;
  if (!same_shape(dst, src)) return cg::fail(dst->ctx, CG_EINVAL, "add: grids differ in padded size");
  if (dst->transformed != src->transformed) return cg::fail(dst->ctx, CG_ESTATE, "add: grids in different domains");
  DeviceGuard dg(dst->ctx->device);
  return cg::add_real(dst->ctx, dst->d, src->d, dst->rsize);
}

int cg_grid_multiply_by_scalar(cg_grid *g, float s)
{
  if (!g) return CG_EINVAL;
  if (g->transformed) return cg::fail(g->ctx, CG_ESTATE, "multiplyByScalar: grid is in the Fourier domain");
  DeviceGuard dg(g->ctx->device);
  return cg::scale_real(g->ctx, g->d, g->rsize, s);
}

int cg_grid_fill_constant(cg_grid *g, float value)
{
  if (!g) return CG_EINVAL;
  DeviceGuard dg(g->ctx->device);
  const int rc = cg::fill_constant(g->ctx, g->d, g->rsize, g->nxp, g->rnxp, value);
  if (rc == CG_OK) g->transformed = 0;
  return rc;
}

int cg_grid_fill_padded(cg_grid *g, const float *cube, int cube_on_device)
{
  if (!g || !cube) return CG_EINVAL;
  cg_ctx *ctx = g->ctx;
  DeviceGuard dg(ctx->device);
  const float *src = cube;
  void *tmp = nullptr;
  if (!cube_on_device) {
    const size_t bytes = sizeof(float) * (size_t)g->nx * g->ny * g->nz;
    hipError_t e = cg::dev_alloc(ctx, &tmp, bytes);
    if (e != hipSuccess) return cg::hip_fail(ctx, e, "fill_padded: staging the cube");
    e = hipMemcpyAsync(tmp, cube, bytes, hipMemcpyHostToDevice, ctx->stream);
    if (e != hipSuccess) {
      cg::dev_free(ctx, tmp);
      return cg::hip_fail(ctx, e, "fill_padded: copying the cube");
    }
    src = static_cast<const float *>(tmp);
  }
  int rc = cg::fill_padded(ctx, g->d, src, g->nx, g->ny, g->nz, g->nxp, g->nyp, g->nzp);
  if (tmp) {
    cg::dev_free(ctx, tmp);
    // the caller may release its host cube on return
    const hipError_t e = hipStreamSynchronize(ctx->stream);
    if (rc == CG_OK) rc = cg::hip_fail(ctx, e, "fill_padded");
  }
  if (rc == CG_OK) {
    g->cubetype = CG_CT_PARAMETER;
    g->transformed = 0;
  }
  return rc;
}

}  // extern "C"

namespace {

// FFTGrid::getFillNumber (fftgrid.cpp:1103-1140) for 0 <= i < np: the box index whose trace a
// padding index copies, mirrored towards the nearer end and clamped at n / 2.
int fill_number(int i, int n, int np)
{
  if (i < n) return i;
  const int to_np = np - i - 1, past_n = i - n + 1;
  return past_n < to_np ? std::max(n - past_n, n / 2) : std::min(to_np, n / 2);
}

// FFTGrid::getDistToBoundary (fftgrid.cpp:1157-1200) for 0 <= i < np: distance into the padding
// in taper lengths min(n, np - n) / 2.1, with the reference's float rounding.
float dist_to_boundary(int i, int n, int np)
{
  if (i < n) return 0.0f;
  const float taper = static_cast<float>(std::min(n, np - n) / 2.1);
  const int to_np = np - i, past_n = i - (n - 1);
  return static_cast<float>(past_n < to_np ? past_n : to_np) / taper;
}

struct ExtrapTables {
  std::vector<int> ref_i, ref_j;
  std::vector<float> dx, dy, dz;
};

ExtrapTables extrap_tables(const cg_grid *g, int imin, int imax, int jmin, int jmax)
{
  ExtrapTables t;
  for (int i = 0; i < g->nxp; ++i) {
    t.ref_i.push_back(std::min(std::max(fill_number(i, g->nx, g->nxp), imin), imax));
    t.dx.push_back(dist_to_boundary(i, g->nx, g->nxp));
  }
  for (int j = 0; j < g->nyp; ++j) {
    t.ref_j.push_back(std::min(std::max(fill_number(j, g->ny, g->nyp), jmin), jmax));
    t.dy.push_back(dist_to_boundary(j, g->ny, g->nyp));
  }
  for (int k = 0; k < g->nzp; ++k) t.dz.push_back(dist_to_boundary(k, g->nz, g->nzp));
  return t;
}

// A range of the level-sorted trace averages: one dependency level (one workgroup per op) or
// a run of small levels applied in order by one workgroup.
struct OpRun {
  int o0, o1;
  bool one_level;
};

// Orders the trace averages by dependency level: an op comes after every earlier op that wrote
// one of its neighbours or read or wrote its own trace, so the ops of a level touch no column
// another of them writes and each reads what the sequential order gives it.  Levels of at
// least kWideLevel ops become one launch each; the smaller ones between them one serial run.
void level_ops(std::vector<unsigned> &ops, int nx, int ntr, std::vector<OpRun> &runs)
{
  constexpr int kWideLevel = 8;
  std::vector<int> wrote(ntr, -1), read(ntr, -1), lev(ops.size());
  int nlev = 0;
  for (size_t o = 0; o < ops.size(); ++o) {
    const int t = (int)(ops[o] >> 4);
    const unsigned m = ops[o] & 15u;
    const int nb[4] = {t - 1, t + 1, t - nx, t + nx};
    int L = std::max(wrote[t], read[t]);
    for (int b = 0; b < 4; ++b)
      if (m >> b & 1u) L = std::max(L, wrote[nb[b]]);
    lev[o] = ++L;
    wrote[t] = L;
    for (int b = 0; b < 4; ++b)
      if (m >> b & 1u) read[nb[b]] = std::max(read[nb[b]], L);
    nlev = std::max(nlev, L + 1);
  }
  std::vector<int> start(nlev + 1, 0);
  for (int l : lev) ++start[l + 1];
  for (int l = 0; l < nlev; ++l) start[l + 1] += start[l];
  std::vector<unsigned> sorted(ops.size());
  std::vector<int> fill(start.begin(), start.end() - 1);
  for (size_t o = 0; o < ops.size(); ++o) sorted[fill[lev[o]]++] = ops[o];
  ops.swap(sorted);
  runs.clear();
  int pending = -1;  // first op of the serial run being gathered
  for (int l = 0; l < nlev; ++l) {
    const int o0 = start[l], o1 = start[l + 1];
    if (o1 - o0 >= kWideLevel) {
      if (pending >= 0) runs.push_back({pending, o0, false});
      pending = -1;
      runs.push_back({o0, o1, true});
    } else if (pending < 0) {
      pending = o0;
    }
  }
  if (pending >= 0) runs.push_back({pending, (int)ops.size(), false});
}

// Stages the extrapolation tables (and, when given, the trace averages before them) and
// enqueues both passes.
int seismic_passes(cg_grid *g, const std::vector<unsigned> &ops, const std::vector<OpRun> &runs, int imin, int imax,
                   int jmin, int jmax)
{
  cg_ctx *ctx = g->ctx;
  const ExtrapTables t = extrap_tables(g, imin, imax, jmin, jmax);
  std::vector<void *> dp;
  int rc = stage_tables(ctx,
                        {{ops.data(), sizeof(unsigned) * ops.size()},
                         {t.ref_i.data(), sizeof(int) * t.ref_i.size()},
                         {t.dx.data(), sizeof(float) * t.dx.size()},
                         {t.ref_j.data(), sizeof(int) * t.ref_j.size()},
                         {t.dy.data(), sizeof(float) * t.dy.size()},
                         {t.dz.data(), sizeof(float) * t.dz.size()}},
                        dp);
  for (const OpRun &r : runs) {
    if (rc != CG_OK) break;
    rc = cg::trace_ops(ctx, g->d, static_cast<const unsigned *>(dp[0]) + r.o0, r.o1 - r.o0, r.one_level, g->nx,
                       g->rnxp, g->nyp, g->nzp);
  }
  if (rc == CG_OK)
    rc = cg::extrapolate(ctx, g->d, static_cast<const int *>(dp[1]), static_cast<const float *>(dp[2]),
                         static_cast<const int *>(dp[3]), static_cast<const float *>(dp[4]),
                         static_cast<const float *>(dp[5]), g->nxp, g->nyp, g->nzp, imin, imax, jmin, jmax);
  return rc;
}

}  // namespace

extern "C" {

int cg_grid_extrapolate_seismic(cg_grid *g, int imin, int imax, int jmin, int jmax)
{
  if (!g) return CG_EINVAL;
  if (g->transformed) return cg::fail(g->ctx, CG_ESTATE, "extrapolate_seismic: grid is in the Fourier domain");
  // outside a non-empty box rectangle the reference's in-place sweep would read cells it has
  // already overwritten; such rectangles are refused
  if (imin < 0 || imin > imax || imax >= g->nx || jmin < 0 || jmin > jmax || jmax >= g->ny)
    return cg::fail(g->ctx, CG_EINVAL, "extrapolate_seismic: the rectangle must lie in the box and be non-empty");
  DeviceGuard dg(g->ctx->device);
  return seismic_passes(g, {}, {}, imin, imax, jmin, jmax);
}

int cg_grid_interpolate_seismic(cg_grid *g, float energy_threshold, int *n_bad, int rect[4])
{
  if (!g) return CG_EINVAL;
  cg_ctx *ctx = g->ctx;
  if (g->cubetype != CG_CT_DATA) return cg::fail(ctx, CG_ESTATE, "interpolate_seismic: grid is not of type DATA");
  if (g->transformed) return cg::fail(ctx, CG_ESTATE, "interpolate_seismic: grid is in the Fourier domain");
  const int nx = g->nx, ny = g->ny, ntr = nx * ny;
  if ((long)nx * ny >= (1L << 28)) return cg::fail(ctx, CG_EINVAL, "interpolate_seismic: more than 2^28 traces");
  DeviceGuard dg(ctx->device);
  // energies over the nz box samples (device), then the flag walk on the host
  void *de = nullptr;
  hipError_t e = cg::dev_alloc(ctx, &de, sizeof(float) * ntr);
  if (e != hipSuccess) return cg::hip_fail(ctx, e, "interpolate_seismic: energy buffer");
  std::vector<float> energy(ntr);
  int rc = cg::trace_energy(ctx, g->d, static_cast<float *>(de), nx, ny, g->nz, g->rnxp, g->nyp);
  if (rc == CG_OK) {
    e = hipMemcpyAsync(energy.data(), de, sizeof(float) * ntr, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    rc = cg::hip_fail(ctx, e, "interpolate_seismic: energies");
  }
  cg::dev_free(ctx, de);
  if (rc != CG_OK) return rc;
  float total = 0.0f;
  for (int t = 0; t < ntr; ++t) total += energy[t];
  const float limit = energy_threshold * total / static_cast<float>(ntr);
  // flag bit 0: this trace is bad; bit 1: a bad trace came earlier (fftgrid.cpp:2351, 2376-2399)
  std::vector<short> flags(ntr);
  int imin = nx, imax = 0, jmin = ny, jmax = 0, nbad = 0;
  short earlier = 0;
  for (int t = 0; t < ntr; ++t) {
    const bool bad = energy[t] <= limit;
    flags[t] = static_cast<short>(earlier + (bad ? 1 : 0));
    if (bad) {
      ++nbad;
      earlier = 2;
    } else {
      imin = std::min(imin, t % nx);
      imax = std::max(imax, t % nx);
      jmin = std::min(jmin, t / nx);
      jmax = std::max(jmax, t / nx);
    }
  }
  if (nbad == ntr) return cg::fail(ctx, CG_EINVAL, "interpolate_seismic: no trace above the energy limit");
  // A bad trace with at least two good 4-neighbours becomes their mean (interpolateTrace,
  // fftgrid.cpp:2440-2481); after a success the pending bad traces before it are retried back
  // to the last settled one, and the settled run is cleared of bit 1 (fftgrid.cpp:2407-2434).
  // The walk depends on the flags alone, so it is replayed here and only the averages run on
  // the device, level by level.
  std::vector<unsigned> ops;
  auto try_trace = [&](int t) {
    const int i = t % nx, j = t / nx;
    unsigned m = 0;
    if (i > 0 && flags[t - 1] % 2 == 0) m |= 1u;
    if (i < nx - 1 && flags[t + 1] % 2 == 0) m |= 2u;
    if (j > 0 && flags[t - nx] % 2 == 0) m |= 4u;
    if (j < ny - 1 && flags[t + nx] % 2 == 0) m |= 8u;
    if (__builtin_popcount(m) > 1) {
      ops.push_back(static_cast<unsigned>(t) << 4 | m);
      --flags[t];
    }
    return flags[t];
  };
  for (int cur = 0; cur < ntr; ++cur) {
    if (flags[cur] % 2 != 1 || try_trace(cur) % 2 != 0) continue;
    int t = cur - 1;
    while (t >= 0 && flags[t] > 1) {
      if (flags[t] % 2 == 1) try_trace(t);
      --t;
    }
    t = std::max(t, 0) + 1;
    while (flags[t - 1] == 0 && t <= cur) {
      if (flags[t] > 1) flags[t] -= 2;
      ++t;
    }
  }
  if (n_bad) *n_bad = nbad;
  if (rect) {
    rect[0] = imin;
    rect[1] = imax;
    rect[2] = jmin;
    rect[3] = jmax;
  }
  std::vector<OpRun> runs;
  level_ops(ops, nx, ntr, runs);
  return seismic_passes(g, ops, runs, imin, imax, jmin, jmax);
}

int cg_grid_fill_trace_taper(cg_grid *g, const float *cube, int cube_on_device)
{
  if (!g || !cube) return CG_EINVAL;
  cg_ctx *ctx = g->ctx;
  if (g->nzp < g->nz + 2)
    return cg::fail(ctx, CG_EINVAL, "fill_trace_taper: needs nzp >= nz + 2 (the taper divides by nzp - nz - 1)");
  DeviceGuard dg(ctx->device);
  const float *src = cube;
  void *tmp = nullptr;
  if (!cube_on_device) {
    const size_t bytes = sizeof(float) * (size_t)g->nx * g->ny * g->nz;
    hipError_t e = cg::dev_alloc(ctx, &tmp, bytes);
    if (e != hipSuccess) return cg::hip_fail(ctx, e, "fill_trace_taper: staging the cube");
    e = hipMemcpyAsync(tmp, cube, bytes, hipMemcpyHostToDevice, ctx->stream);
    if (e != hipSuccess) {
      cg::dev_free(ctx, tmp);
      return cg::hip_fail(ctx, e, "fill_trace_taper: copying the cube");
    }
    src = static_cast<const float *>(tmp);
  }
  int rc = cg::fill_trace_taper(ctx, g->d, src, g->nx, g->ny, g->nz, g->nxp, g->nyp, g->nzp);
  if (tmp) {
    cg::dev_free(ctx, tmp);
    const hipError_t e = hipStreamSynchronize(ctx->stream);
    if (rc == CG_OK) rc = cg::hip_fail(ctx, e, "fill_trace_taper");
  }
  if (rc == CG_OK) g->transformed = 0;
  return rc;
}

int cg_grid_fill_param_corr(cg_grid *g, const float *corr_xy, const float *corr_t, float grad_i, float grad_j,
                            float scale)
{
  if (!g || !corr_xy || !corr_t) return CG_EINVAL;
  cg_ctx *ctx = g->ctx;
  DeviceGuard dg(ctx->device);
  std::vector<void *> dp;
  int rc = stage_tables(ctx, {{corr_xy, sizeof(float) * g->nxp * g->nyp}, {corr_t, sizeof(float) * g->nzp}}, dp);
  if (rc == CG_OK)
    rc = cg::fill_param_corr(ctx, g->d, static_cast<const float *>(dp[0]), static_cast<const float *>(dp[1]), g->nxp,
                             g->nyp, g->nzp, grad_i, grad_j, scale);
  if (rc == CG_OK) g->transformed = 0;
  return rc;
}

int cg_grid_fill_err_corr(cg_grid *g, const float *corr_xy, float grad_i, float grad_j)
{
  if (!g || !corr_xy) return CG_EINVAL;
  cg_ctx *ctx = g->ctx;
  DeviceGuard dg(ctx->device);
  std::vector<void *> dp;
  int rc = stage_tables(ctx, {{corr_xy, sizeof(float) * g->nxp * g->nyp}}, dp);
  if (rc == CG_OK)
    rc = cg::fill_err_corr(ctx, g->d, static_cast<const float *>(dp[0]), g->nxp, g->nyp, g->nzp, grad_i, grad_j);
  if (rc == CG_OK) g->transformed = 0;
  return rc;
}

int cg_grid_fill_noise_reference(cg_grid *g, uint32_t *seed)
{
  if (!g || !seed) return CG_EINVAL;
  DeviceGuard dg(g->ctx->device);
  const int rc = cg::fill_noise_reference(g->ctx, reinterpret_cast<float2 *>(g->d), g->nxp, g->nyp, g->nzp, seed);
  if (rc == CG_OK) {
    g->transformed = 1;
    g->cubetype = CG_CT_PARAMETER;
  }
  return rc;
}

int cg_grid_fill_noise_counter(cg_grid *g, uint64_t seed, uint32_t stream)
{
  if (!g) return CG_EINVAL;
  DeviceGuard dg(g->ctx->device);
  int rc = cg::fill_noise(g->ctx, reinterpret_cast<float2 *>(g->d), (long)g->csize, g->cnxp, g->nyp, g->nzp, seed,
                          stream);
  if (rc == CG_OK) {
    g->transformed = 1;
    g->cubetype = CG_CT_PARAMETER;
  }
  return rc;
}

// ---------------------------------------------------------------------------- posterior
static int post_validate(cg_ctx *ctx, const cg_post_args *a)
{
  if (!ctx || !a) return CG_EINVAL;
  const int nt = a->ntheta;
  if (nt < 1 || nt > cg::kMaxTheta) return cg::fail(ctx, CG_EINVAL, "posterior: ntheta must be 1..8");
  if (!a->K || !a->e1 || !a->e2 || !a->e3 || !a->invert || !a->sigma0e || !a->prior_var0)
    return cg::fail(ctx, CG_EINVAL, "posterior: missing table");
  if (a->prior_mode != CG_PRIOR_GRIDS && a->prior_mode != CG_PRIOR_SEPARABLE)
    return cg::fail(ctx, CG_EINVAL, "posterior: unknown prior mode");
  if (a->precision != CG_PREC_REFERENCE && a->precision != CG_PREC_MIXED)
    return cg::fail(ctx, CG_EINVAL, "posterior: unknown precision");
  const cg_grid *ref = a->mean[0];
  if (!ref) return cg::fail(ctx, CG_EINVAL, "posterior: missing mean grid");
  std::vector<const cg_grid *> all;
  for (int g = 0; g < 3; ++g) all.push_back(a->mean[g]);
  for (int l = 0; l < nt; ++l) all.push_back(a->data[l]);
  for (int g = 0; g < 6; ++g) all.push_back(a->cov[g]);
  if (a->err_corr) all.push_back(a->err_corr);
  for (const cg_grid *g : all) {
    if (!g) return cg::fail(ctx, CG_EINVAL, "posterior: missing grid");
    if (g->ctx != ctx) return cg::fail(ctx, CG_EINVAL, "posterior: grid from another context");
    if (!same_shape(g, ref)) return cg::fail(ctx, CG_EINVAL, "posterior: grids differ in padded size");
  }
  for (int g = 0; g < 3; ++g)
    if (!a->mean[g]->transformed) return cg::fail(ctx, CG_ESTATE, "posterior: mean grids must be transformed");
  for (int l = 0; l < nt; ++l)
    if (!a->data[l]->transformed) return cg::fail(ctx, CG_ESTATE, "posterior: data grids must be transformed");
  if (a->prior_mode == CG_PRIOR_GRIDS) {
    for (int g = 0; g < 6; ++g)
      if (!a->cov[g]->transformed) return cg::fail(ctx, CG_ESTATE, "posterior: covariance grids must be transformed");
  } else if (!a->rxy || !a->rt) {
    return cg::fail(ctx, CG_EINVAL, "posterior: separable prior needs rxy and rt");
  }
  if (a->err_corr) {
    if (!a->err_corr->transformed) return cg::fail(ctx, CG_ESTATE, "posterior: errCorr grid must be transformed");
  } else if (!a->lam2d) {
    return cg::fail(ctx, CG_EINVAL, "posterior: need err_corr grid or lam2d table");
  }
  return CG_OK;
}

// Device parameters of a posterior call; the caller sets the table, list/mask and counter
// pointers.
static cg::PostParams post_params(const cg_post_args *a)
{
  const cg_grid *ref = a->mean[0];
  cg::PostParams p;
  std::memset(&p, 0, sizeof(p));
  p.nxp = ref->nxp; p.nyp = ref->nyp; p.nzp = ref->nzp; p.cnxp = ref->cnxp; p.nt = a->ntheta;
  p.csize = (long)ref->csize;
  for (int g = 0; g < 3; ++g) p.mean[g] = reinterpret_cast<float2 *>(a->mean[g]->d);
  for (int l = 0; l < a->ntheta; ++l) p.data[l] = reinterpret_cast<float2 *>(a->data[l]->d);
  for (int g = 0; g < 6; ++g) p.cov[g] = reinterpret_cast<float2 *>(a->cov[g]->d);
  p.lam_grid = a->err_corr ? reinterpret_cast<const float2 *>(a->err_corr->d) : nullptr;
  const double *pv = a->prior_var0;
  const double s6[6] = {pv[0], pv[4], pv[8], pv[1], pv[2], pv[5]};
  for (int g = 0; g < 6; ++g) {
    p.s6f[g] = (float)s6[g];
    p.s6d[g] = s6[g];
  }
  p.prior_mode = a->prior_mode;
  p.want_residual = a->want_residual;
  p.precision = a->precision;
  p.cond_limit = (float)(a->cond_limit > 0 ? a->cond_limit : 1e6);
  return p;
}

// Statistics of a finished posterior call from its counters (device: [0] deferred, [8, 72)
// inverted slots, [72, 136) failure slots).  Synchronises.
static int post_stats(cg_ctx *ctx, const unsigned long long *dev_counters, bool mixed, uint64_t *stats)
{
  uint64_t h[136];
  hipError_t e = hipMemcpyAsync(h, dev_counters, sizeof(h), hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) return cg::hip_fail(ctx, e, "posterior stats");
  uint64_t inv_n = 0, fail_n = 0;
  for (int s = 0; s < 64; ++s) {
    inv_n += h[8 + s];
    fail_n += h[8 + 64 + s];
  }
  stats[0] = inv_n;
  stats[1] = fail_n;
  stats[2] = mixed ? h[0] : 0;  // FP64 re-solves = deferred wavenumbers
  return CG_OK;
}

int cg_posterior(cg_ctx *ctx, const cg_post_args *a)
{
  int rc0 = post_validate(ctx, a);
  if (rc0) return rc0;
  const cg_grid *ref = a->mean[0];
  DeviceGuard dg(ctx->device);
  const int nyp = ref->nyp, nzp = ref->nzp, cnxp = ref->cnxp;
  const size_t plane = (size_t)cnxp * nyp;

  std::vector<cg::PlaneTab<float>> t32;
  std::vector<cg::PlaneTab<double>> t64;
  build_plane_tables<float>(a, nzp, t32);
  const bool mixed = a->precision == CG_PREC_MIXED;
  if (mixed) build_plane_tables<double>(a, nzp, t64);
  std::vector<int> inv(a->invert, a->invert + nzp);
  size_t ninv = 0;
  for (int v : inv) ninv += (v != 0);
  std::vector<std::pair<const void *, size_t>> parts = {
      {t32.data(), sizeof(cg::PlaneTab<float>) * nzp},
      {mixed ? (const void *)t64.data() : nullptr, mixed ? sizeof(cg::PlaneTab<double>) * nzp : 0},
      {inv.data(), sizeof(int) * nzp},
      {a->prior_mode == CG_PRIOR_SEPARABLE ? a->rxy : nullptr, a->prior_mode == CG_PRIOR_SEPARABLE ? sizeof(float) * plane : 0},
      {a->prior_mode == CG_PRIOR_SEPARABLE ? a->rt : nullptr, a->prior_mode == CG_PRIOR_SEPARABLE ? sizeof(float) * nzp : 0},
      {a->err_corr ? nullptr : a->lam2d, a->err_corr ? 0 : sizeof(float) * plane},
  };
  // FP64 deferral list: capacity = inverted cells
  const size_t list_bytes = mixed ? align_up(std::max<size_t>(ninv * plane, 1) * sizeof(unsigned), 256) : 0;
  std::vector<void *> dp;
  int rc = stage_tables(ctx, parts, dp);
  if (rc) return rc;
  unsigned *list = nullptr;
  if (mixed) {
    if (ctx->list_bytes < list_bytes) {
      (void)hipStreamSynchronize(ctx->stream);
      if (ctx->list) {
        (void)hipFree(ctx->list);
        ctx->in_use -= ctx->list_bytes;
      }
      ctx->list = nullptr;
      ctx->list_bytes = 0;
      hipError_t e = hipMalloc(&ctx->list, list_bytes);
      if (e != hipSuccess) return cg::hip_fail(ctx, e, "hipMalloc(fp64 list)");
      ctx->list_bytes = list_bytes;
      ctx->in_use += list_bytes;
      ctx->peak = std::max(ctx->peak, ctx->in_use);
    }
    list = ctx->list;
  }
  cg::PostParams p = post_params(a);
  p.tab32 = static_cast<const cg::PlaneTab<float> *>(dp[0]);
  p.tab64 = static_cast<const cg::PlaneTab<double> *>(dp[1]);
  p.invert = static_cast<const int *>(dp[2]);
  p.rxy = static_cast<const float *>(dp[3]);
  p.rt = static_cast<const float *>(dp[4]);
  p.lam2d = static_cast<const float *>(dp[5]);
  p.list = list;
  // counters[0]: FP64 list length; counters[8, 136): statistics slots
  unsigned long long *cnt = reinterpret_cast<unsigned long long *>(ctx->counters);
  p.list_len = mixed ? cnt : nullptr;
  p.stats = a->stats ? cnt + 8 : nullptr;
  if (mixed || a->stats) {
    rc = cg::zero_counters(ctx, cnt, 136);
    if (rc) return rc;
  }
  rc = cg::launch_posterior(ctx, p);
  if (rc) return rc;
  for (int g = 0; g < 6; ++g) {  // the covariance grids now hold posterior covariance spectra
    a->cov[g]->transformed = 1;
    a->cov[g]->cubetype = CG_CT_COVARIANCE;
  }
  if (a->stats) return post_stats(ctx, cnt, mixed, a->stats);
  return CG_OK;
}

int cg_posterior_noalloc(cg_ctx *ctx, const cg_post_args *a)
{
  int rc = post_validate(ctx, a);
  if (rc) return rc;
  const cg_grid *ref = a->mean[0];
  const int nzp = ref->nzp;
  const size_t plane = (size_t)ref->cnxp * ref->nyp;
  if ((size_t)nzp > cg::kNoAllocMaxNzp || plane > cg::kNoAllocMaxPlane || ref->csize > cg::kNoAllocMaxCells)
    return cg::fail(ctx, CG_EINVAL, "posterior_noalloc: grid larger than the fixed buffer allows (nzp <= 1536, "
                                    "(nxp/2+1)*nyp <= 2^21 and (nxp/2+1)*nyp*nzp <= 2^28 complex cells)");
  DeviceGuard dg(ctx->device);
  char *base = static_cast<char *>(cg::noalloc_scratch(ctx->device));
  if (!base) return cg::fail(ctx, CG_EHIP, "posterior_noalloc: the device buffer's address was not found");
  std::vector<cg::PlaneTab<float>> t32;
  std::vector<cg::PlaneTab<double>> t64;
  build_plane_tables<float>(a, nzp, t32);
  const bool mixed = a->precision == CG_PREC_MIXED;
  if (mixed) build_plane_tables<double>(a, nzp, t64);
  const bool sep = a->prior_mode == CG_PRIOR_SEPARABLE;
  const std::pair<const void *, size_t> parts[6] = {
      {t32.data(), sizeof(cg::PlaneTab<float>) * nzp},
      {mixed ? (const void *)t64.data() : nullptr, mixed ? sizeof(cg::PlaneTab<double>) * nzp : 0},
      {a->invert, sizeof(int) * nzp},
      {sep ? a->rxy : nullptr, sep ? sizeof(float) * plane : 0},
      {sep ? a->rt : nullptr, sep ? sizeof(float) * nzp : 0},
      {a->err_corr ? nullptr : a->lam2d, a->err_corr ? 0 : sizeof(float) * plane},
  };
  // the tables go to the buffer straight from host memory; the call then waits for the copies, so
  // that the host tables may go (they live on this call's stack and heap)
  char *tab = base + cg::kNoAllocTables;
  const void *dp[6];
  size_t off = 0;
  hipError_t e = hipSuccess;
  for (int i = 0; i < 6; ++i) {
    dp[i] = parts[i].second ? tab + off : nullptr;
    if (parts[i].second && e == hipSuccess)
      e = hipMemcpyAsync(tab + off, parts[i].first, parts[i].second, hipMemcpyHostToDevice, ctx->stream);
    off += align_up(parts[i].second, 256);
  }
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) return cg::hip_fail(ctx, e, "posterior_noalloc tables");
  cg::PostParams p = post_params(a);
  p.tab32 = static_cast<const cg::PlaneTab<float> *>(dp[0]);
  p.tab64 = static_cast<const cg::PlaneTab<double> *>(dp[1]);
  p.invert = static_cast<const int *>(dp[2]);
  p.rxy = static_cast<const float *>(dp[3]);
  p.rt = static_cast<const float *>(dp[4]);
  p.lam2d = static_cast<const float *>(dp[5]);
  // counters at the buffer's start: [0] deferred wavenumbers, [8, 136) statistics slots; the
  // deferral mask after the tables, cleared here (the previous call's contents remain otherwise)
  unsigned long long *cnt = reinterpret_cast<unsigned long long *>(base);
  p.list = nullptr;
  p.mask = mixed ? reinterpret_cast<unsigned long long *>(base + cg::kNoAllocMask) : nullptr;
  p.list_len = mixed ? cnt : nullptr;
  p.stats = a->stats ? cnt + 8 : nullptr;
  if (mixed || a->stats) {
    rc = cg::zero_counters(ctx, cnt, 136);
    if (rc) return rc;
  }
  if (mixed) {
    rc = cg::zero_words(ctx, p.mask, ((size_t)ref->csize + 63) / 64);
    if (rc) return rc;
  }
  rc = cg::launch_posterior(ctx, p);
  if (rc) return rc;
  for (int g = 0; g < 6; ++g) {
    a->cov[g]->transformed = 1;
    a->cov[g]->cubetype = CG_CT_COVARIANCE;
  }
  if (a->stats) return post_stats(ctx, cnt, mixed, a->stats);
  return CG_OK;
}

// ---------------------------------------------------------------------------- simulation
// One realisation (next null), or realisations r and r + 1 of one posterior (next: the second's
// output grids), coloured in one fused launch where the length has one, else one at a time.
static int simulate_impl(cg_ctx *ctx, const cg_sim_args *a, cg_grid *const *next)
{
  if (!ctx || !a) return CG_EINVAL;
  const cg_grid *ref = a->cov[0];
  if (!ref) return cg::fail(ctx, CG_EINVAL, "simulate: missing covariance grid");
  for (int g = 0; g < 6; ++g) {
    if (!a->cov[g] || !same_shape(a->cov[g], ref)) return cg::fail(ctx, CG_EINVAL, "simulate: bad covariance grid");
    if (!a->cov[g]->transformed) return cg::fail(ctx, CG_ESTATE, "simulate: covariance grids must be transformed");
  }
  // Means in the Fourier domain are added to the coloured spectra (the inverse transform is
  // linear, and a PARAMETER grid's spectrum carries the same unitary scaling as the
  // realisation's), which saves reading them again in the x pass; real-domain means are added
  // after the inverse transform.
  int spectral = -1;  // -1: no mean, 0: real domain, 1: Fourier domain
  for (int g = 0; g < 3; ++g) {
    if (!a->out[g] || !same_shape(a->out[g], ref)) return cg::fail(ctx, CG_EINVAL, "simulate: bad output grid");
    if (a->mean[g]) {
      if (!same_shape(a->mean[g], ref)) return cg::fail(ctx, CG_EINVAL, "simulate: bad mean grid");
      const int dom = a->mean[g]->transformed ? 1 : 0;
      if (spectral >= 0 && dom != spectral)
        return cg::fail(ctx, CG_ESTATE, "simulate: mean grids must all be in the same domain");
      spectral = dom;
      if (dom && a->mean[g]->cubetype == CG_CT_COVARIANCE)
        return cg::fail(ctx, CG_EINVAL, "simulate: a Fourier-domain mean must be a PARAMETER grid");
    }
    if (a->noise_source == CG_NOISE_GRIDS && !a->out[g]->transformed)
      return cg::fail(ctx, CG_ESTATE, "simulate: noise grids must be in the Fourier domain");
  }
  if (a->noise_source != CG_NOISE_COUNTER && a->noise_source != CG_NOISE_GRIDS)
    return cg::fail(ctx, CG_EINVAL, "simulate: unknown noise source");
  if (next) {
    for (int g = 0; g < 3; ++g) {
      if (!next[g] || !same_shape(next[g], ref)) return cg::fail(ctx, CG_EINVAL, "simulate_pair: bad output grid");
      for (int h = 0; h < 3; ++h)
        if (next[g] == a->out[h]) return cg::fail(ctx, CG_EINVAL, "simulate_pair: output grids overlap");
    }
  }
  DeviceGuard dg(ctx->device);
  cg::ColourParams p;
  std::memset(&p, 0, sizeof(p));
  p.csize = (long)ref->csize;
  p.cnxp = ref->cnxp;
  p.nyp = ref->nyp;
  p.nzp = ref->nzp;
  cg::colour_dims(p);
  for (int g = 0; g < 6; ++g) p.cov[g] = reinterpret_cast<const float2 *>(a->cov[g]->d);
  for (int g = 0; g < 3; ++g) p.out[g] = reinterpret_cast<float2 *>(a->out[g]->d);
  for (int g = 0; g < 3; ++g)
    p.mean[g] = spectral == 1 && a->mean[g] ? reinterpret_cast<const float2 *>(a->mean[g]->d) : nullptr;
  p.from_grids = a->noise_source == CG_NOISE_GRIDS;
  p.seed = a->seed;
  // streams 3r, 3r + 1, 3r + 2: a pair of consecutive realisations from an even one takes three
  // whole Philox blocks per cell
  p.stream0 = 3u * a->realisation;
  p.fails = a->stats ? reinterpret_cast<unsigned long long *>(ctx->counters) : nullptr;
  int rc = CG_OK;
  if (p.fails) rc = cg::zero_counters(ctx, p.fails, 1);
  if (rc) return rc;
  // Counter noise with a compiled z plan: the colouring runs inside the inverse z pass, so the
  // coloured spectra never make a round trip through HBM.  Where the grid has a compiled split
  // of the z transform, the colouring carries the inverse's stage 2 and the x pass its stage 1.
  bool z_done = false, split = false;
  const bool fusable = !p.from_grids && ctx->fft_engine == CG_FFT_AUTO && p.nzp > 1;
  int sn1 = 0, sn2 = 0;
  if (fusable && ctx->fused_colour != 0 && cg::split_z_supported(ref->nxp, ref->nzp, &sn1, &sn2)) {
    if (next) {
      for (int g = 0; g < 3; ++g) p.out2[g] = reinterpret_cast<float2 *>(next[g]->d);
      p.stream1 = 3u * (a->realisation + 1u);
    }
    rc = cg::colour_split(ctx, p, sn1, sn2, next != nullptr);
    if (rc) return rc;
    split = true;
  } else if (next) {
    // both realisations in one launch, or one after the other
    for (int g = 0; g < 3; ++g) p.out2[g] = reinterpret_cast<float2 *>(next[g]->d);
    p.stream1 = 3u * (a->realisation + 1u);
    if (fusable && cg::fixed_colour_zinv2(ctx, p, p.nzp, (long)p.cnxp * p.nyp, &rc) != 0) {
      z_done = true;
    } else if (!p.from_grids) {
      rc = cg::colour2(ctx, p);  // unfused, then the z passes of all six grids
    } else {
      uint64_t s1 = 0, s2 = 0;
      cg_sim_args b = *a;
      b.stats = a->stats ? &s1 : nullptr;
      rc = simulate_impl(ctx, &b, nullptr);
      if (rc) return rc;
      for (int g = 0; g < 3; ++g) b.out[g] = next[g];
      b.realisation = a->realisation + 1u;
      b.stats = a->stats ? &s2 : nullptr;
      rc = simulate_impl(ctx, &b, nullptr);
      if (!rc && a->stats) a->stats[0] = s1 + s2;
      return rc;
    }
    if (rc) return rc;
  } else {
    if (fusable) z_done = cg::fixed_colour_zinv(ctx, p, p.nzp, (long)p.cnxp * p.nyp, &rc) != 0;
    if (!z_done) rc = cg::colour(ctx, p);
    if (rc) return rc;
  }
  for (int s = 0; s < (next ? 2 : 1); ++s) {
    for (int g = 0; g < 3; ++g) {
      cg_grid *o = s ? next[g] : a->out[g];
      o->cubetype = CG_CT_PARAMETER;
      o->transformed = 1;
      const float *add = spectral == 0 && a->mean[g] ? a->mean[g]->d : nullptr;
      if (split) {  // y pass, then the inverse stage 1 with the x c2r (PARAMETER: unitary scale)
        const long plane = (long)o->cnxp * o->nyp;
        if (o->nyp > 1)
          rc = cg::fft_strided(ctx, reinterpret_cast<float2 *>(o->d), o->nyp, o->cnxp, o->cnxp, o->nzp, plane, +1,
                               1.0f, CG_K_FFT_Y);
        const size_t n = (size_t)o->nxp * o->nyp * o->nzp;
        const float sc = float(1.0 / sqrtf(float(n)));  // as grid_fft_inverse
        if (!rc) rc = cg::fft_xz(ctx, 0, nullptr, o->d, add, o->nxp, o->nyp, o->nzp, sc);
      } else {
        rc = z_done ? cg::grid_fft_inverse_xy(o, add) : cg::grid_fft_inverse(o, add);
      }
      if (rc) return rc;
      o->transformed = 0;
    }
  }
  if (a->stats) {
    uint64_t h = 0;
    hipError_t e = hipMemcpyAsync(&h, ctx->counters, sizeof(h), hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return cg::hip_fail(ctx, e, "simulate stats");
    a->stats[0] = h;
  }
  return CG_OK;
}

int cg_simulate(cg_ctx *ctx, const cg_sim_args *a) { return simulate_impl(ctx, a, nullptr); }

int cg_simulate_pair(cg_ctx *ctx, const cg_sim_args *a, cg_grid *next_out[3])
{
  if (!next_out) return CG_EINVAL;
  return simulate_impl(ctx, a, next_out);
}

// ---------------------------------------------------------------------------- one inversion
// AVOInversion::computePostMeanResidAndFFTCov (avoinversion.cpp:229-244, 908-1114) and simulate
// (:1401-1511) for one inversion.  With a compiled split of the z transform (cg_invert.hip) the
// transform's stages ride in the x passes and the posterior pass; otherwise the drop-in
// sequence (forward transforms, cg_posterior, inverse transforms) runs.
int cg_invert(cg_ctx *ctx, const cg_inversion *v)
{
  if (!ctx || !v) return CG_EINVAL;
  const cg_post_args &a = v->post;
  const int nt = a.ntheta;
  if (nt < 1 || nt > cg::kMaxTheta) return cg::fail(ctx, CG_EINVAL, "invert: ntheta must be 1..8");
  const cg_grid *ref = v->mean_in[0];
  if (!ref) return cg::fail(ctx, CG_EINVAL, "invert: missing input grid");
  {
    std::vector<const cg_grid *> all;
    for (int g = 0; g < 3; ++g) all.insert(all.end(), {v->mean_in[g], a.mean[g]});
    for (int l = 0; l < nt; ++l) all.insert(all.end(), {v->seis_in[l], a.data[l]});
    for (int g = 0; g < 6; ++g) all.push_back(a.cov[g]);
    if (v->nreal < 0 || (v->nreal > 0 && !v->out)) return cg::fail(ctx, CG_EINVAL, "invert: bad realisations");
    for (int r = 0; r < 3 * v->nreal; ++r) all.push_back(v->out[r]);
    for (const cg_grid *g : all)
      if (!g || g->ctx != ctx || !same_shape(g, ref)) return cg::fail(ctx, CG_EINVAL, "invert: bad grid");
    for (int g = 0; g < 3; ++g)
      if (v->mean_in[g]->transformed) return cg::fail(ctx, CG_ESTATE, "invert: inputs must be in the real domain");
    for (int l = 0; l < nt; ++l)
      if (v->seis_in[l]->transformed) return cg::fail(ctx, CG_ESTATE, "invert: inputs must be in the real domain");
    for (int g = 0; g < 3; ++g)
      for (int l = 0; l < nt; ++l)
        if (a.mean[g] == a.data[l]) return cg::fail(ctx, CG_EINVAL, "invert: work grids overlap");
  }
  DeviceGuard dg(ctx->device);
  const int nxp = ref->nxp, nyp = ref->nyp, nzp = ref->nzp, cnxp = ref->cnxp;
  const size_t plane = (size_t)cnxp * nyp;
  int n1 = 0, n2 = 0;
  const bool split = ctx->fft_engine == CG_FFT_AUTO && !a.want_residual && a.prior_mode == CG_PRIOR_SEPARABLE &&
                     !a.err_corr && a.rxy && a.rt && a.lam2d && cg::split_z_supported(nxp, nzp, &n1, &n2);
  int rc = CG_OK;
  std::vector<cg_grid *> work;
  for (int g = 0; g < 3; ++g) work.push_back(a.mean[g]);
  for (int l = 0; l < nt; ++l) work.push_back(a.data[l]);
  auto input = [&](int i) { return i < 3 ? v->mean_in[i] : v->seis_in[i - 3]; };
  if (!split) {
    for (int i = 0; i < 3 + nt && !rc; ++i) {
      work[i]->cubetype = input(i)->cubetype;
      rc = cg::grid_fft_forward(work[i], input(i)->d);
      if (!rc) work[i]->transformed = 1;
    }
    if (!rc) rc = cg_posterior(ctx, &a);
    for (int g = 0; g < 3 && !rc; ++g) {
      rc = cg::grid_fft_inverse(a.mean[g], nullptr);
      if (!rc) a.mean[g]->transformed = 0;
    }
  } else {
    if (a.precision != CG_PREC_REFERENCE && a.precision != CG_PREC_MIXED)
      return cg::fail(ctx, CG_EINVAL, "invert: unknown precision");
    if (!a.K || !a.e1 || !a.e2 || !a.e3 || !a.invert || !a.sigma0e || !a.prior_var0)
      return cg::fail(ctx, CG_EINVAL, "invert: missing table");
    const bool mixed = a.precision == CG_PREC_MIXED;
    std::vector<cg::PlaneTab<float>> t32;
    std::vector<cg::PlaneTab<double>> t64;
    build_plane_tables<float>(&a, nzp, t32);
    if (mixed) build_plane_tables<double>(&a, nzp, t64);
    std::vector<int> inv(a.invert, a.invert + nzp);
    std::vector<std::pair<const void *, size_t>> parts = {
        {t32.data(), sizeof(cg::PlaneTab<float>) * nzp},
        {mixed ? (const void *)t64.data() : nullptr, mixed ? sizeof(cg::PlaneTab<double>) * nzp : 0},
        {inv.data(), sizeof(int) * nzp},
        {a.rxy, sizeof(float) * plane},
        {a.rt, sizeof(float) * nzp},
        {a.lam2d, sizeof(float) * plane},
    };
    std::vector<void *> dp;
    rc = stage_tables(ctx, parts, dp);
    if (rc) return rc;
    const unsigned ctiles = (unsigned)((plane + 63) / 64);
    const size_t ntiles = (size_t)ctiles * n2;
    unsigned *list = nullptr;
    if (mixed) {  // the deferred-tile list: capacity all tiles
      const size_t bytes = align_up(ntiles * sizeof(unsigned), 256);
      if (ctx->list_bytes < bytes) {
        (void)hipStreamSynchronize(ctx->stream);
        if (ctx->list) {
          (void)hipFree(ctx->list);
          ctx->in_use -= ctx->list_bytes;
        }
        ctx->list = nullptr;
        ctx->list_bytes = 0;
        hipError_t e = hipMalloc(&ctx->list, bytes);
        if (e != hipSuccess) return cg::hip_fail(ctx, e, "hipMalloc(deferred tiles)");
        ctx->list_bytes = bytes;
        ctx->in_use += bytes;
        ctx->peak = std::max(ctx->peak, ctx->in_use);
      }
      list = ctx->list;
    }
    // the scales of grid_fft_forward / grid_fft_inverse
    const long N = (long)nxp * nyp * nzp;
    const float s_fwd = 1.0f / sqrtf((float)N);
    const float s_inv = float(1.0 / sqrtf(float((size_t)N)));
    for (int i = 0; i < 3 + nt && !rc; ++i) {
      const cg_grid *in = input(i);
      cg_grid *w = work[i];
      const float s = in->cubetype == CG_CT_COVARIANCE ? 1.0f : s_fwd;
      rc = cg::fft_xz(ctx, 1, in->d, w->d, nullptr, nxp, nyp, nzp, s);
      if (!rc && nyp > 1)
        rc = cg::fft_strided(ctx, reinterpret_cast<float2 *>(w->d), nyp, cnxp, cnxp, nzp, (long)plane, -1, 1.0f,
                             CG_K_FFT_Y);
      w->cubetype = in->cubetype;
      w->transformed = 1;
    }
    if (rc) return rc;
    cg::PostParams p = post_params(&a);
    p.tab32 = static_cast<const cg::PlaneTab<float> *>(dp[0]);
    p.tab64 = static_cast<const cg::PlaneTab<double> *>(dp[1]);
    p.invert = static_cast<const int *>(dp[2]);
    p.rxy = static_cast<const float *>(dp[3]);
    p.rt = static_cast<const float *>(dp[4]);
    p.lam2d = static_cast<const float *>(dp[5]);
    // counters[0]: FP64 re-solves; [1]: deferred tiles; [8, 136): statistics slots
    unsigned long long *cnt = reinterpret_cast<unsigned long long *>(ctx->counters);
    p.list_len = mixed ? cnt : nullptr;
    p.stats = a.stats ? cnt + 8 : nullptr;
    if (mixed || a.stats) {
      rc = cg::zero_counters(ctx, cnt, 136);
      if (rc) return rc;
    }
    cg::PostZArgs z;
    z.p = p;
    z.n2 = n2;
    z.ctiles = ctiles;
    z.redo = list;
    z.redo_len = mixed ? cnt + 1 : nullptr;
    rc = cg::launch_posterior_z(ctx, z, n1, mixed);
    if (rc) return rc;
    for (int g = 0; g < 6; ++g) {
      a.cov[g]->transformed = 1;
      a.cov[g]->cubetype = CG_CT_COVARIANCE;
    }
    // the realisations while the means are still in the split layout: the colouring adds them
    // there, and the realisations' inverse passes read nothing else
    if (v->nreal > 0) {
      cg::ColourParams cp;
      std::memset(&cp, 0, sizeof(cp));
      cp.csize = (long)ref->csize;
      cp.cnxp = cnxp;
      cp.nyp = nyp;
      cp.nzp = nzp;
      cg::colour_dims(cp);
      for (int g = 0; g < 6; ++g) cp.cov[g] = reinterpret_cast<const float2 *>(a.cov[g]->d);
      cp.seed = v->seed;
      const float2 *ms[3];
      for (int g = 0; g < 3; ++g) ms[g] = reinterpret_cast<const float2 *>(a.mean[g]->d);
      for (int r = 0; r < v->nreal && !rc; r += 2) {
        const bool pair = r + 1 < v->nreal;
        for (int g = 0; g < 3; ++g) {
          cp.out[g] = reinterpret_cast<float2 *>(v->out[3 * r + g]->d);
          cp.out2[g] = pair ? reinterpret_cast<float2 *>(v->out[3 * (r + 1) + g]->d) : nullptr;
        }
        cp.stream0 = 3u * (uint32_t)r;
        cp.stream1 = 3u * (uint32_t)(r + 1);
        rc = cg::colour_split(ctx, cp, n1, n2, pair, ms);
        for (int q = r; q < (pair ? r + 2 : r + 1) && !rc; ++q) {
          for (int g = 0; g < 3 && !rc; ++g) {
            cg_grid *o = v->out[3 * q + g];
            o->cubetype = CG_CT_PARAMETER;
            if (nyp > 1)
              rc = cg::fft_strided(ctx, reinterpret_cast<float2 *>(o->d), nyp, cnxp, cnxp, nzp, (long)plane, +1, 1.0f,
                                   CG_K_FFT_Y);
            if (!rc) rc = cg::fft_xz(ctx, 0, nullptr, o->d, nullptr, nxp, nyp, nzp, s_inv);
            o->transformed = 0;
          }
        }
      }
      if (rc) return rc;
    }
    for (int g = 0; g < 3 && !rc; ++g) {
      float2 *c = reinterpret_cast<float2 *>(a.mean[g]->d);
      if (nyp > 1) rc = cg::fft_strided(ctx, c, nyp, cnxp, cnxp, nzp, (long)plane, +1, 1.0f, CG_K_FFT_Y);
      const float s = a.mean[g]->cubetype == CG_CT_COVARIANCE ? float(1.0 / (double)N) : s_inv;
      if (!rc) rc = cg::fft_xz(ctx, 0, nullptr, a.mean[g]->d, nullptr, nxp, nyp, nzp, s);
      a.mean[g]->transformed = 0;
    }
    if (!rc && a.stats) rc = post_stats(ctx, cnt, mixed, a.stats);
  }
  if (rc || split) return rc;
  // realisations: pairs (2m, 2m + 1) as cg_simulate_pair, a last odd one alone; the posterior
  // means are added in the real domain
  for (int r = 0; r < v->nreal && !rc; r += 2) {
    cg_sim_args s;
    std::memset(&s, 0, sizeof(s));
    for (int g = 0; g < 6; ++g) s.cov[g] = a.cov[g];
    for (int g = 0; g < 3; ++g) {
      s.mean[g] = a.mean[g];
      s.out[g] = v->out[3 * r + g];
    }
    s.noise_source = CG_NOISE_COUNTER;
    s.seed = v->seed;
    s.realisation = (uint32_t)r;
    if (r + 1 < v->nreal) rc = simulate_impl(ctx, &s, v->out + 3 * (r + 1));
    else rc = simulate_impl(ctx, &s, nullptr);
  }
  return rc;
}

// ---------------------------------------------------------------------------- forward modelling
int cg_forward_model(cg_ctx *ctx, int nt, const float *K, cg_grid *m[3], cg_grid *d[])
{
  if (!ctx || !K || !m || !d) return CG_EINVAL;
  if (nt < 1 || nt > cg::kMaxTheta) return cg::fail(ctx, CG_EINVAL, "forward_model: ntheta must be 1..8");
  const cg_grid *ref = m[0];
  if (!ref) return cg::fail(ctx, CG_EINVAL, "forward_model: missing grid");
  for (int g = 0; g < 3 + nt; ++g) {
    const cg_grid *x = g < 3 ? m[g] : d[g - 3];
    if (!x || x->ctx != ctx || !same_shape(x, ref)) return cg::fail(ctx, CG_EINVAL, "forward_model: bad grid");
    if (g < 3 && !x->transformed) return cg::fail(ctx, CG_ESTATE, "forward_model: elastic grids must be transformed");
  }
  DeviceGuard dg(ctx->device);
  std::vector<void *> dp;
  const size_t kb = sizeof(float) * 2 * 3 * (size_t)nt * ref->nzp;
  int rc = stage_tables(ctx, {{K, kb}}, dp);
  if (rc) return rc;
  cg::ForwardParams p;
  std::memset(&p, 0, sizeof(p));
  p.plane = (long)ref->cnxp * ref->nyp;
  p.nzp = ref->nzp;
  p.nt = nt;
  p.K = static_cast<const float2 *>(dp[0]);
  for (int g = 0; g < 3; ++g) p.m[g] = reinterpret_cast<const float2 *>(m[g]->d);
  for (int l = 0; l < nt; ++l) {
    p.d[l] = reinterpret_cast<float2 *>(d[l]->d);
    d[l]->transformed = 1;
  }
  return cg::forward_model(ctx, p);
}

int cg_grid_add_white_noise(cg_grid *g, float sigma, uint64_t seed, uint32_t stream)
{
  if (!g) return CG_EINVAL;
  if (g->transformed) return cg::fail(g->ctx, CG_ESTATE, "add_white_noise: grid is in the Fourier domain");
  DeviceGuard dg(g->ctx->device);
  return cg::add_white_noise(g->ctx, g->d, g->nxp, g->rnxp, (long)g->nyp * g->nzp, sigma, seed, stream);
}

int cg_grid_sum_squares(cg_grid *g, int ext, double *sum)
{
  if (!g || !sum) return CG_EINVAL;
  if (g->transformed) return cg::fail(g->ctx, CG_ESTATE, "sum_squares: grid is in the Fourier domain");
  cg_ctx *ctx = g->ctx;
  DeviceGuard dg(ctx->device);
  const int ny = ext ? g->nyp : g->ny, nz = ext ? g->nzp : g->nz;
  char *work = static_cast<char *>(cg::table_space(ctx, cg::box_sum_squares_work(ny, nz) + 256));
  if (!work) return CG_ENOMEM;
  double *d = reinterpret_cast<double *>(work);
  int rc = cg::box_sum_squares(ctx, g->d, ext ? g->nxp : g->nx, ny, nz, g->rnxp, g->nyp, false, work + 256, d,
                               nullptr);
  if (rc) return rc;
  hipError_t e = hipMemcpyAsync(sum, d, sizeof(double), hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  return cg::hip_fail(ctx, e, "sum_squares");
}

int cg_grid_data_variance(cg_grid *g, float *variance, uint64_t *count)
{
  if (!g || !variance) return CG_EINVAL;
  if (g->transformed) return cg::fail(g->ctx, CG_ESTATE, "data_variance: grid is in the Fourier domain");
  cg_ctx *ctx = g->ctx;
  DeviceGuard dg(ctx->device);
  char *work = static_cast<char *>(cg::table_space(ctx, cg::box_sum_squares_work(g->ny, g->nz) + 256));
  if (!work) return CG_ENOMEM;
  double *d = reinterpret_cast<double *>(work);
  unsigned long long *n = reinterpret_cast<unsigned long long *>(work + 64);
  int rc = cg::box_sum_squares(ctx, g->d, g->nx, g->ny, g->nz, g->rnxp, g->nyp, true, work + 256, d, n);
  if (rc) return rc;
  struct {
    double sum;
    unsigned long long n;
  } h;
  hipError_t e = hipMemcpyAsync(&h.sum, d, sizeof(double), hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(&h.n, n, sizeof(unsigned long long), hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) return cg::hip_fail(ctx, e, "data_variance");
  // modelavodynamic.cpp:612-617: zero when every sample is zero
  *variance = h.n == 0 ? 0.0f : static_cast<float>(h.sum / static_cast<double>(h.n));
  if (count) *count = h.n;
  return CG_OK;
}

// ---------------------------------------------------------------------------- host setup
}  // extern "C"

namespace {

// |M| = V |diag(w)| V^T of a real symmetric n x n matrix (row-major, overwritten), by cyclic
// Jacobi rotations.
template <int N>
void abs_sym(double (&a)[N][N])
{
  double v[N][N] = {};
  for (int i = 0; i < N; ++i) v[i][i] = 1.0;
  for (int sweep = 0; sweep < 64; ++sweep) {
    double off = 0.0, all = 0.0;
    for (int p = 0; p < N; ++p)
      for (int q = 0; q < N; ++q) {
        all += a[p][q] * a[p][q];
        if (p != q) off += a[p][q] * a[p][q];
      }
    if (off <= 1e-30 * all) break;
    for (int p = 0; p < N - 1; ++p)
      for (int q = p + 1; q < N; ++q) {
        if (a[p][q] == 0.0) continue;
        // rotation in the (p, q) plane that zeroes a[p][q]: t = tan of the angle
        const double th = (a[q][q] - a[p][p]) / (2.0 * a[p][q]);
        const double t = (th >= 0 ? 1.0 : -1.0) / (std::fabs(th) + std::sqrt(th * th + 1.0));
        const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
        for (int r = 0; r < N; ++r) {  // columns p, q
          const double rp = a[r][p], rq = a[r][q];
          a[r][p] = c * rp - s * rq;
          a[r][q] = s * rp + c * rq;
        }
        for (int r = 0; r < N; ++r) {  // rows p, q
          const double pr = a[p][r], qr = a[q][r];
          a[p][r] = c * pr - s * qr;
          a[q][r] = s * pr + c * qr;
        }
        for (int r = 0; r < N; ++r) {
          const double rp = v[r][p], rq = v[r][q];
          v[r][p] = c * rp - s * rq;
          v[r][q] = s * rp + c * rq;
        }
      }
  }
  double w[N];
  for (int i = 0; i < N; ++i) w[i] = std::fabs(a[i][i]);
  for (int i = 0; i < N; ++i)
    for (int j = 0; j < N; ++j) {
      double x = 0.0;
      for (int k = 0; k < N; ++k) x += v[i][k] * w[k] * v[j][k];
      a[i][j] = x;
    }
}

}  // namespace

extern "C" {

int cg_make_circ_auto_cov_pos_def(int nzp, float *cac)
{
  if (nzp < 1 || !cac) return CG_EINVAL;
  const int nf = nzp / 2 + 1;
  const double two_pi = 6.283185307179586476925286766559;
  auto series = [&](int j, int k) { return cac + (size_t)(3 * j + k) * nzp; };
  // lag-0 standard deviations, as the reference rounds them (square root of a float)
  double d[3], dinv[3];
  for (int i = 0; i < 3; ++i) {
    const float v0 = series(i, i)[0];
    if (!(v0 > 0.0f)) return CG_EINVAL;
    d[i] = (double)std::sqrt(v0);
    dinv[i] = 1.0 / (double)std::sqrt(v0);
  }
  // spectra of the nine series (sign -1, unnormalised, bins 0..nzp/2), rounded to float as
  // FFTGrid::fft1DzInPlace returns them
  std::vector<std::complex<float>> spec((size_t)9 * nf);
  for (int s = 0; s < 9; ++s) {
    const float *x = cac + (size_t)s * nzp;
    for (int m = 0; m < nf; ++m) {
      double re = 0.0, im = 0.0;
      for (int t = 0; t < nzp; ++t) {
        const double ang = two_pi * (double)((long)m * t % nzp) / nzp;
        re += x[t] * std::cos(ang);
        im -= x[t] * std::sin(ang);
      }
      spec[(size_t)s * nf + m] = std::complex<float>((float)re, (float)im);
    }
  }
  // per frequency: M = D^-1 C D^-1 (Hermitian for Hermitian pairs of cross-covariances),
  // C' = D |M| D with |M| from its real symmetric 6 x 6 form [[Re, -Im], [Im, Re]]
  for (int m = 0; m < nf; ++m) {
    double r[6][6];
    for (int j = 0; j < 3; ++j)
      for (int k = 0; k < 3; ++k) {
        const std::complex<float> c = spec[(size_t)(3 * j + k) * nf + m];
        const double re = dinv[j] * (double)c.real() * dinv[k], im = dinv[j] * (double)c.imag() * dinv[k];
        r[j][k] = re;
        r[j + 3][k + 3] = re;
        r[j][k + 3] = -im;
        r[j + 3][k] = im;
      }
    abs_sym<6>(r);
    for (int j = 0; j < 3; ++j)
      for (int k = 0; k < 3; ++k)
        spec[(size_t)(3 * j + k) * nf + m] =
            std::complex<float>((float)(d[j] * r[j][k] * d[k]), (float)(d[j] * r[j + 3][k] * d[k]));
  }
  // back to lags: half-complex to real (imaginary parts of bins 0 and nzp/2 for even nzp ignored,
  // as FFTW's c2r does), unnormalised, then times 1/nzp in float
  const float factor = (float)(1.0 / (double)nzp);
  for (int s = 0; s < 9; ++s) {
    const std::complex<float> *X = &spec[(size_t)s * nf];
    float *x = cac + (size_t)s * nzp;
    for (int t = 0; t < nzp; ++t) {
      double acc = X[0].real();
      for (int m = 1; m < nf; ++m) {
        const double ang = two_pi * (double)((long)m * t % nzp) / nzp;
        const double term = X[m].real() * std::cos(ang) - X[m].imag() * std::sin(ang);
        acc += (2 * m == nzp) ? X[m].real() * std::cos(ang) : 2.0 * term;
      }
      x[t] = (float)acc * factor;
    }
  }
  return CG_OK;
}

// CommonData::FillInData (commondata.cpp:6451-6775), the SEG-Y DATA branch, with the columns
// resolved by the caller: the traces tapered and zero-padded, their length-nt R2C spectra (the x
// pass), the bins spread into length-mt spectra, the length-mt C2R (the x pass inverse, scale
// 1 / nt), every column interpolated from its trace's fine samples.
int cg_grid_fill_from_traces(cg_grid *g, const cg_trace_fill_args *a)
{
  if (!g || !a || !a->samples || !a->length || !a->z0_data || !a->column_trace || !a->z0_grid || !a->dz_grid)
    return CG_EINVAL;
  cg_ctx *ctx = g->ctx;
  if (a->ntraces < 1 || a->stride < 1) return cg::fail(ctx, CG_EINVAL, "fill_from_traces: no traces");
  if (!(a->dz_data > 0.0f) || !(a->smooth_length >= 0.0f))
    return cg::fail(ctx, CG_EINVAL, "fill_from_traces: bad sample interval or smoothing length");
  if (a->grid_type != 0 && a->grid_type != CG_CT_DATA && a->grid_type != CG_CT_PARAMETER)
    return cg::fail(ctx, CG_EINVAL, "fill_from_traces: grid_type must be CG_CT_DATA or CG_CT_PARAMETER");
  const bool detrend = a->grid_type == CG_CT_PARAMETER;
  const int ns = (int)std::floor(a->smooth_length / a->dz_data);
  int maxlen = 0;
  for (int t = 0; t < a->ntraces; ++t) {
    const int len = a->length[t];
    if (len < 1 || len > a->stride) return cg::fail(ctx, CG_EINVAL, "fill_from_traces: trace length outside 1..stride");
    if (!detrend && len < ns)
      return cg::fail(ctx, CG_EINVAL, "fill_from_traces: a trace is shorter than its guard-zone taper");
    maxlen = std::max(maxlen, len);
  }
  const long ncol = (long)g->nyp * g->rnxp;
  for (long c = 0; c < ncol; ++c)
    if (a->column_trace[c] < -2 || a->column_trace[c] >= a->ntraces)
      return cg::fail(ctx, CG_EINVAL, "fill_from_traces: column trace index out of range");
  if (detrend)  // FillInData stops at the first box column outside the data (commondata.cpp:6767-6770)
    for (int j = 0; j < g->ny; ++j)
      for (int i = 0; i < g->nx; ++i)
        if (a->column_trace[(long)j * g->rnxp + i] == -1)
          return cg::fail(ctx, CG_EINVAL, "fill_from_traces: the traces do not cover the box");
  const int nt = cg_factorable(maxlen);
  if (nt > 1024) return cg::fail(ctx, CG_EINVAL, "fill_from_traces: traces longer than 1024 samples");
  DeviceGuard dg(ctx->device);
  cg::TraceFillParams p;
  std::memset(&p, 0, sizeof(p));
  p.ntraces = a->ntraces;
  p.stride = a->stride;
  p.dz_data = a->dz_data;
  p.ns = ns;
  p.nt = nt;
  p.rnt = 2 * (nt / 2 + 1);
  p.mt = 10 * nt;
  p.rmt = 2 * (p.mt / 2 + 1);
  p.dz_fine = a->dz_data / 10.0f;
  p.nz = g->nz;
  p.nyp = g->nyp;
  p.nzp = g->nzp;
  p.rnxp = g->rnxp;
  p.grid = g->d;
  p.detrend = detrend ? 1 : 0;
  p.trim_shift = a->trim_keeps_z0 ? 0 : 1;
  // one device block: samples, lengths, z0_data, columns (trace, z0, dz), the two trace buffers,
  // the detrended traces' TraceInfo
  Layout L;
  const size_t o_s = L.add(sizeof(float) * (size_t)a->ntraces * a->stride);
  const size_t o_len = L.add(sizeof(int) * a->ntraces);
  const size_t o_z0d = L.add(sizeof(float) * a->ntraces);
  const size_t o_ct = L.add(sizeof(int) * ncol);
  const size_t o_z0g = L.add(sizeof(float) * ncol);
  const size_t o_dzg = L.add(sizeof(float) * ncol);
  const size_t o_a = L.add(sizeof(float) * (size_t)a->ntraces * p.rnt);
  const size_t o_b = L.add(sizeof(float) * (size_t)a->ntraces * p.rmt);
  const size_t o_info = L.add(detrend ? sizeof(cg::TraceInfo) * a->ntraces : 0);
  char *dev = nullptr;
  hipError_t e = cg::dev_alloc(ctx, reinterpret_cast<void **>(&dev), L.size);
  if (e != hipSuccess) return cg::hip_fail(ctx, e, "fill_from_traces: device buffers");
  const struct {
    size_t off;
    const void *src;
    size_t bytes;
  } up[] = {{o_s, a->samples, sizeof(float) * (size_t)a->ntraces * a->stride},
            {o_len, a->length, sizeof(int) * a->ntraces},
            {o_z0d, a->z0_data, sizeof(float) * a->ntraces},
            {o_ct, a->column_trace, sizeof(int) * ncol},
            {o_z0g, a->z0_grid, sizeof(float) * ncol},
            {o_dzg, a->dz_grid, sizeof(float) * ncol}};
  for (const auto &u : up) {
    if (e == hipSuccess) e = hipMemcpyAsync(dev + u.off, u.src, u.bytes, hipMemcpyHostToDevice, ctx->stream);
  }
  int rc = cg::hip_fail(ctx, e, "fill_from_traces: copying the traces");
  p.samples = reinterpret_cast<const float *>(dev + o_s);
  p.length = reinterpret_cast<const int *>(dev + o_len);
  p.z0_data = reinterpret_cast<const float *>(dev + o_z0d);
  p.column_trace = reinterpret_cast<const int *>(dev + o_ct);
  p.z0_grid = reinterpret_cast<const float *>(dev + o_z0g);
  p.dz_grid = reinterpret_cast<const float *>(dev + o_dzg);
  p.a = reinterpret_cast<float *>(dev + o_a);
  p.b = reinterpret_cast<float *>(dev + o_b);
  p.info = detrend ? reinterpret_cast<cg::TraceInfo *>(dev + o_info) : nullptr;
  if (!rc) rc = detrend ? cg::pack_detrend(ctx, p) : cg::pack_traces(ctx, p);
  if (!rc) rc = cg::fft_x_r2c(ctx, p.a, p.a, nt, a->ntraces, 1.0f);
  if (!rc) rc = cg::spread_spectra(ctx, p);
  if (!rc) rc = cg::fft_x_c2r(ctx, p.b, p.mt, a->ntraces, 1 / static_cast<float>(nt), nullptr);
  if (!rc) rc = cg::fill_columns(ctx, p);
  cg::dev_free(ctx, dev);
  // the caller may release its host arrays on return
  e = hipStreamSynchronize(ctx->stream);
  if (!rc) rc = cg::hip_fail(ctx, e, "fill_from_traces");
  if (!rc) {
    g->cubetype = detrend ? CG_CT_PARAMETER : CG_CT_DATA;
    g->transformed = 0;
  }
  return rc;
}

int cg_factorable(int leastint)
{
  if (leastint <= 1) return 1;
  long best = 1;
  while (best < leastint) best *= 2;
  for (long a = 1; a <= 2 * best; a *= 2)
    for (long b = a; b <= 2 * best; b *= 3)
      for (long c = b; c <= 2 * best; c *= 5)
        for (long d = c; d <= 2 * best; d *= 7)
          if (d >= leastint && d < best) best = d;
  return (int)best;
}

int cg_build_kz_tables(int nz, int nzp, double dt_ms, int nt, const float *wav, const float *A, float low_cut,
                       float high_cut, float min_rel_thick, float *K, float *e1, float *e2, float *e3, int *invert)
{
  if (nz < 1 || nzp < nz || nt < 1 || nt > cg::kMaxTheta || !wav || !A || !K || !e1 || !e2 || !e3 || !invert)
    return CG_EINVAL;
  std::vector<double> ct(nzp), st(nzp);
  for (int m = 0; m < nzp; ++m) {
    ct[m] = std::cos(2.0 * M_PI * m / nzp);
    st[m] = std::sin(2.0 * M_PI * m / nzp);
  }
  // forward difference operator [-1, 0, ..., 0, +1] (wavelet1D.cpp:411-415)
  std::vector<double> diff(nzp, 0.0);
  diff[0] = -1.0;
  diff[nzp - 1] += 1.0;
  std::vector<std::vector<double>> W(nt, std::vector<double>(nzp)), S(nt, std::vector<double>(nzp));
  std::vector<float> wn(nt), sn(nt);
  for (int l = 0; l < nt; ++l) {
    const float *wl = wav + (size_t)l * nzp;
    std::vector<float> w(wl, wl + nzp), f(nzp), b(nzp), c(nzp);
    wn[l] = find_norm(w);
    // errorSmooth3 = central(backward(forward(w))), circular (wavelet1D.cpp:441-500)
    for (int i = 0; i < nzp; ++i) f[i] = (i == nzp - 1) ? w[0] - w[i] : w[i + 1] - w[i];
    for (int i = 0; i < nzp; ++i) b[i] = (i == 0) ? f[0] - f[nzp - 1] : f[i] - f[i - 1];
    for (int i = 0; i < nzp; ++i) {
      const float hi = (i == nzp - 1) ? b[0] : b[i + 1];
      const float lo = (i == 0) ? b[nzp - 1] : b[i - 1];
      c[i] = (float)(0.5 * (hi - lo));
    }
    sn[l] = find_norm(c);
    for (int i = 0; i < nzp; ++i) {
      W[l][i] = w[i];
      S[l][i] = c[i];
    }
  }
  const double lz = nz * dt_ms;
  for (int k = 0; k < nzp; ++k) {
    const float rf = (float)((nz * 1000.0f) / (lz * nzp) * std::min(k, nzp - k));
    invert[k] = (rf > low_cut * (double)min_rel_thick) && (rf < high_cut);
    double kdr, kdi;
    dft_bin(diff, ct, st, k, kdr, kdi);
    const float kDr = (float)kdr, kDi = (float)kdi;
    for (int l = 0; l < nt; ++l) {
      double wr, wi, sr, si;
      dft_bin(W[l], ct, st, k, wr, wi);
      dft_bin(S[l], ct, st, k, sr, si);
      const float kwr = (float)wr, kwi = -(float)wi;  // fillkW: conj(W)
      const float pr = kDr * kwr - kDi * kwi, pi = kDr * kwi + kDi * kwr;
      for (int j = 0; j < 3; ++j) {
        K[2 * ((k * nt + l) * 3 + j)] = pr * A[l * 3 + j];
        K[2 * ((k * nt + l) * 3 + j) + 1] = pi * A[l * 3 + j];
      }
      e1[2 * (k * nt + l)] = (float)wr / wn[l];
      e1[2 * (k * nt + l) + 1] = -((float)wi / wn[l]);
      e2[2 * (k * nt + l)] = (float)sr / sn[l];
      e2[2 * (k * nt + l) + 1] = -((float)si / sn[l]);
      e3[2 * (k * nt + l)] = 1.0f;
      e3[2 * (k * nt + l) + 1] = 0.0f;
    }
  }
  return CG_OK;
}

// Wavelet::getCAmp(k, scale) (wavelet.cpp:337-394) on the half spectrum c[0, nzp/2]: the spectrum
// of the wavelet stretched by `scale`, c(k / scale) / scale, linearly interpolated between bins
// (zero beyond nzp/2), the Hermitian mirror for k > nzp/2; float arithmetic as the reference.
static std::complex<float> stretched_bin(const std::vector<std::complex<float>> &c, int nzp, int k, float scale)
{
  const int cnzp = nzp / 2 + 1;
  const bool mirror = k >= cnzp;
  const float omega = float(mirror ? nzp - k : k) / scale;
  if (omega >= cnzp) return std::complex<float>(0.0f, 0.0f);
  const int omL = int(std::floor(omega));
  int omU = omL + 1;
  if (omU >= cnzp) omU -= 1;
  const float d = omega - float(omL);
  float re = c[omL].real() * (1.0f - d) + c[omU].real() * d;
  float im = c[omL].imag() * (1.0f - d) + c[omU].imag() * d;
  if (mirror) im = -im;
  return std::complex<float>(re / scale, im / scale);
}

int cg_grid_scale_wavelet(cg_grid *g, const cg_scale_wavelet_args *a)
{
  if (!g || !a) return CG_EINVAL;
  cg_ctx *ctx = g->ctx;
  if (a->mode != CG_SCALE_WAVELET_DIVIDE && a->mode != CG_SCALE_WAVELET_MULTIPLY)
    return cg::fail(ctx, CG_EINVAL, "scale_wavelet: unknown mode");
  if (!a->wavelet || !a->rel_thick || !(a->dt_ms > 0))
    return cg::fail(ctx, CG_EINVAL, "scale_wavelet: missing wavelet, relative thickness or sample interval");
  if (a->mode == CG_SCALE_WAVELET_DIVIDE && (!a->corr_t || !a->A || !a->prior_var0))
    return cg::fail(ctx, CG_EINVAL, "scale_wavelet: divide needs corr_t, A and prior_var0");
  if (g->transformed) return cg::fail(ctx, CG_ESTATE, "scale_wavelet: grid is in the Fourier domain");
  const int nzp = g->nzp, cnzp = nzp / 2 + 1, nz = g->nz;
  for (long t = 0; t < (long)g->nxp * g->nyp; ++t)
    if (!(a->rel_thick[t] > 0)) return cg::fail(ctx, CG_EINVAL, "scale_wavelet: relative thickness must be positive");
  DeviceGuard dg(ctx->device);
  std::vector<double> ct(nzp), st(nzp);
  for (int m = 0; m < nzp; ++m) {
    ct[m] = std::cos(2.0 * M_PI * m / nzp);
    st[m] = std::sin(2.0 * M_PI * m / nzp);
  }
  std::vector<float> wv(a->wavelet, a->wavelet + nzp);
  std::vector<double> wd(wv.begin(), wv.end());
  std::vector<float2> W(cnzp), D3(cnzp);
  std::vector<float> tau2(cnzp, 0.0f);
  for (int m = 0; m < cnzp; ++m) {
    double wr, wi;
    dft_bin(wd, ct, st, m, wr, wi);
    W[m] = make_float2((float)wr, (float)wi);
    const double om = 2.0 * M_PI * m / nzp;
    D3[m] = make_float2(0.0f, (float)(std::sin(om) * (2.0 * std::cos(om) - 2.0)));  // i sin(w) (2 cos(w) - 2)
  }
  const float wnorm = find_norm(wv);
  // band norm of errorSmooth3's spectrum (Wavelet::findNormWithinFrequencyBand, wavelet.cpp:903-925)
  const float dOmega = (float)(1.0 / (nzp * a->dt_ms * 0.001));
  const int loI = std::max<int>(0, (int)std::floor(a->low_cut / dOmega));
  const int hiI = std::min<int>(cnzp, (int)std::ceil(a->high_cut / dOmega));
  double n2 = 0.0;
  const double fac = 1.0 / (double)nzp;
  for (int m = loI; m < hiI; ++m) {
    const float sr = -D3[m].y * W[m].y, si = D3[m].y * W[m].x;  // (i d) (wr + i wi)
    n2 += 2.0 * fac * (sr * sr + si * si);
  }
  if (loI == 0) {
    const float sr = -D3[0].y * W[0].y, si = D3[0].y * W[0].x;
    n2 -= fac * (sr * sr + si * si);
  }
  const float norm_f3 = (float)std::sqrt(n2);
  const double lz = nz * a->dt_ms;
  const float maxfrequency = (float)((nzp / 2) * 1000.0 * nz) / (float)(lz * nzp);
  const float modw = wnorm * wnorm * (maxfrequency / a->high_cut);
  if (a->mode == CG_SCALE_WAVELET_DIVIDE) {
    // the reflection-coefficient time covariance (avoinversion.cpp:368-401) and its operator FFT
    double eiVar = 0.0;
    for (int l = 0; l < 3; ++l)
      for (int m = 0; m < 3; ++m) eiVar += a->A[l] * (float)a->prior_var0[3 * l + m] * a->A[m];
    std::vector<float> ei(nzp), rc(nzp);
    for (int k = 0; k < nzp; ++k) ei[k] = (float)(eiVar * a->corr_t[k]);
    for (int i = 0; i < nzp; ++i) {
      const float next = ei[(i + 1) % nzp], prev = ei[(i + nzp - 1) % nzp];
      rc[i] = 2 * ei[i] - next - prev;
    }
    std::vector<double> rcd(rc.begin(), rc.end());
    for (int m = 0; m < cnzp; ++m) {
      double r, im;
      dft_bin(rcd, ct, st, m, r, im);
      tau2[m] = std::fabs((float)r);
    }
  }
  std::vector<float> taper;
  for (int k = nz + 1; k < nzp; ++k) {
    const float d = dist_to_boundary(k, nz, nzp);
    taper.push_back(std::max<float>(1 - d * d, 0));
  }
  const size_t box = (size_t)g->nx * g->ny;
  std::vector<void *> dp;
  int rc = stage_tables(ctx,
                        {{W.data(), sizeof(float2) * cnzp},
                         {D3.data(), sizeof(float2) * cnzp},
                         {tau2.data(), sizeof(float) * cnzp},
                         {taper.empty() ? nullptr : taper.data(), sizeof(float) * taper.size()},
                         {a->rel_thick, sizeof(double) * g->nxp * g->nyp},
                         {a->shift_ms, a->shift_ms ? sizeof(float) * box : 0},
                         {a->gain, a->gain ? sizeof(float) * box : 0},
                         {a->stretch, a->stretch ? sizeof(float) * box : 0}},
                        dp);
  if (rc) return rc;
  cg::ScaleWaveletParams p;
  std::memset(&p, 0, sizeof(p));
  p.data = reinterpret_cast<float2 *>(g->d);
  p.mode = a->mode;
  p.nx = g->nx; p.ny = g->ny; p.nz = nz; p.nxp = g->nxp; p.nyp = g->nyp; p.nzp = nzp; p.cnxp = g->cnxp;
  p.lz = lz;
  p.dz = (float)a->dt_ms;
  p.low_cut = a->low_cut; p.high_cut = a->high_cut; p.wnc = a->wnc; p.error_var = a->error_var;
  p.wnorm = wnorm; p.norm_f3 = norm_f3; p.modw = modw;
  p.W = static_cast<const float2 *>(dp[0]);
  p.D3 = static_cast<const float2 *>(dp[1]);
  p.tau2 = static_cast<const float *>(dp[2]);
  p.rel_thick = static_cast<const double *>(dp[4]);
  p.shift = static_cast<const float *>(dp[5]);
  p.gain = static_cast<const float *>(dp[6]);
  p.stretch = static_cast<const float *>(dp[7]);
  return cg::scale_wavelet_filter(ctx, p, static_cast<const float *>(dp[3]));
}

int cg_build_kz_tables_variable_thickness(int nz, int nzp, double dt_ms, int nt, const float *wav, const float *A,
                                          float low_cut, float high_cut, float min_rel_thick, float *K, float *e1,
                                          float *e2, float *e3, int *invert)
{
  if (nz < 1 || nzp < nz || nt < 1 || nt > cg::kMaxTheta || !wav || !A || !K || !e1 || !e2 || !e3 || !invert)
    return CG_EINVAL;
  std::vector<double> ct(nzp), st(nzp);
  for (int m = 0; m < nzp; ++m) {
    ct[m] = std::cos(2.0 * M_PI * m / nzp);
    st[m] = std::sin(2.0 * M_PI * m / nzp);
  }
  // difference operators (wavelet1D.cpp:389-500): forward [-1, 0, ..., 0, +1]; D3 its backward then
  // central difference, as sequences
  std::vector<double> d1(nzp, 0.0), d2(nzp), d3(nzp);
  d1[0] = -1.0;
  d1[nzp - 1] += 1.0;
  for (int i = 0; i < nzp; ++i) d2[i] = (float)(i == 0 ? d1[0] - d1[nzp - 1] : d1[i] - d1[i - 1]);
  for (int i = 0; i < nzp; ++i) {
    const double hi = i == nzp - 1 ? d2[0] : d2[i + 1], lo = i == 0 ? d2[nzp - 1] : d2[i - 1];
    d3[i] = (float)(0.5 * (hi - lo));
  }
  const int cnzp = nzp / 2 + 1;
  std::vector<std::vector<std::complex<float>>> Wc(nt, std::vector<std::complex<float>>(cnzp));
  std::vector<float> wn(nt), sn(nt);
  for (int l = 0; l < nt; ++l) {
    const float *wl = wav + (size_t)l * nzp;
    std::vector<float> w(wl, wl + nzp), f(nzp), b(nzp), c(nzp);
    wn[l] = find_norm(w);
    for (int i = 0; i < nzp; ++i) f[i] = (i == nzp - 1) ? w[0] - w[i] : w[i + 1] - w[i];
    for (int i = 0; i < nzp; ++i) b[i] = (i == 0) ? f[0] - f[nzp - 1] : f[i] - f[i - 1];
    for (int i = 0; i < nzp; ++i) {
      const float hi = (i == nzp - 1) ? b[0] : b[i + 1];
      const float lo = (i == 0) ? b[nzp - 1] : b[i - 1];
      c[i] = (float)(0.5 * (hi - lo));
    }
    sn[l] = find_norm(c);  // errorSmooth3's norm
    std::vector<double> wd(w.begin(), w.end());
    for (int m = 0; m < cnzp; ++m) {  // seisWaveletForNorm's spectrum
      double wr, wi;
      dft_bin(wd, ct, st, m, wr, wi);
      Wc[l][m] = std::complex<float>((float)wr, (float)wi);
    }
  }
  const double lz = nz * dt_ms;
  for (int k = 0; k < nzp; ++k) {
    const float rf = (float)((nz * 1000.0f) / (lz * nzp) * std::min(k, nzp - k));
    invert[k] = (rf > low_cut * (double)min_rel_thick) && (rf < high_cut);
    double kdr, kdi, k3r, k3i;
    dft_bin(d1, ct, st, k, kdr, kdi);
    dft_bin(d3, ct, st, k, k3r, k3i);
    for (int l = 0; l < nt; ++l) {
      // K = D A (avoinversion.cpp:990-1000): the data were divided by the scale wavelet
      for (int j = 0; j < 3; ++j) {
        K[2 * ((k * nt + l) * 3 + j)] = (float)kdr * A[l * 3 + j];
        K[2 * ((k * nt + l) * 3 + j) + 1] = (float)kdi * A[l * 3 + j];
      }
      // errMult1 = 1 / |w_l|, errMult2 = D3 / |errorSmooth3_l| (avoinversion.cpp:1002-1013)
      e1[2 * (k * nt + l)] = 1.0f / wn[l];
      e1[2 * (k * nt + l) + 1] = 0.0f;
      const float sm = 1.0f / sn[l];
      e2[2 * (k * nt + l)] = (float)k3r * sm;
      e2[2 * (k * nt + l) + 1] = (float)k3i * sm;
      // errMult3 = fillInverseAbskWRobust (avoinversion.cpp:1296-1322): 1 / the larger modulus of
      // the wavelet spectrum at k and at k / 0.999
      std::complex<float> v = k < cnzp ? Wc[l][k] : std::conj(Wc[l][nzp - k]);
      const std::complex<float> vf = stretched_bin(Wc[l], nzp, k, 0.999f);
      const float mod = v.real() * v.real() + v.imag() * v.imag();
      const float modf = vf.real() * vf.real() + vf.imag() * vf.imag();
      const float mx = std::max(mod, modf);
      e3[2 * (k * nt + l)] = mx > 0.0f ? (float)(1.0 / std::sqrt(mx)) : wn[l] * nzp * nzp * 100.0f;
      e3[2 * (k * nt + l) + 1] = 0.0f;
    }
  }
  return CG_OK;
}

}  // extern "C"

// ============================================================================ batches
// A cg_batch holds the device descriptors of every pass of its inversions, built once; a run
// is a fixed sequence of launches (capturable into a hipGraph).
struct cg_batch {
  cg_ctx *ctx = nullptr;
  int n = 0;
  std::vector<cg_inversion> inv;
  std::vector<std::vector<cg_grid *>> outs;
  void *dev = nullptr;
  size_t dev_bytes = 0;
  // One launch of a compile-time-plan kernel over every grid of the pass that has that length.
  struct FixedGroup {
    const cg::FixedEntry *e = nullptr;
    int variant = 0;              // tile variant (the autotuned one for the first grid's geometry)
    int add = 0;                  // x inverse: the grids add a real grid
    const void *items = nullptr;  // cg::StridedArgs or cg::XArgs, device memory
    int nitems = 0, nxp = 0;
    unsigned blocks = 0;
  };
  // A pass over many grids: one launch per compiled length, and one runtime-plan launch for the
  // grids whose length has no compiled plan.
  struct Pass {
    int kind = 0, dir = -1, kclass = 0, nitems = 0;
    const cg::FftItem *items = nullptr;
    unsigned blocks = 0;
    size_t lds = 0;
    std::vector<FixedGroup> fixed;
  };
  std::vector<Pass> fwd, inv_mean;
  std::vector<std::vector<Pass>> sim;  // per realisation: z, y, x passes
  struct PostGroup {
    int nt = 0;
    bool sep = false, mixed = false;
    const cg::PostItem *items = nullptr;
    int nitems = 0;
    unsigned blocks = 0;
  };
  std::vector<PostGroup> post;
  // Colouring of one realisation number r of every inversion that has it: the colouring kernel
  // over the inversions whose realisations are not fused, and per z length one launch of the
  // colouring fused with the inverse z pass over the rest (whose z passes are then skipped); for
  // inversions whose pairs are fused, the launch at even r colours r and r + 1 together.
  struct FusedColour {
    const cg::FixedEntry *e = nullptr;
    const cg::ColourItem *items = nullptr;
    int nitems = 0;
    unsigned blocks = 0;
    bool pair = false;  // realisations r and r + 1 per item (the paired kernel)
  };
  struct ColourPass {
    const cg::ColourItem *items = nullptr;
    int nitems = 0;
    unsigned blocks = 0;
    std::vector<FusedColour> fused;
  };
  std::vector<ColourPass> colour;
  unsigned long long *counters = nullptr;
  int ncounters = 0;
  bool use_graph = false;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipEvent_t done = nullptr;
};

namespace {

// Side streams (default priority) of a context, created once: the launches of one batched pass
// (one per compiled length) and the posterior groups are independent and run concurrently;
// each pass forks from and joins back into the context stream, so the order of passes is kept
// (also inside a captured graph).
int ensure_side_streams(cg_ctx *ctx)
{
  if (ctx->fork) return CG_OK;
  hipError_t e = hipSuccess;
  for (int k = 0; k < cg_ctx::kSide && e == hipSuccess; ++k) {
    if (!ctx->side[k]) e = hipStreamCreateWithFlags(&ctx->side[k], hipStreamNonBlocking);
    if (e == hipSuccess && !ctx->join[k]) e = hipEventCreateWithFlags(&ctx->join[k], hipEventDisableTiming);
  }
  if (e == hipSuccess) e = hipEventCreateWithFlags(&ctx->fork, hipEventDisableTiming);
  return cg::hip_fail(ctx, e, "side streams");
}

// Runs launch(i) for i < n: on the context stream when n < 2, else spread over the side streams
// between a fork from and a join into the context stream.
template <typename F>
int fork_join(cg_batch *b, int n, F launch)
{
  cg_ctx *ctx = b->ctx;
  if (n < 2 || !ctx->fork) {
    for (int i = 0; i < n; ++i)
      if (int rc = launch(i)) return rc;
    return CG_OK;
  }
  hipStream_t main = ctx->stream;
  const int ns = std::min(n, cg_ctx::kSide);
  hipError_t e = hipEventRecord(ctx->fork, main);
  for (int k = 0; k < ns && e == hipSuccess; ++k) e = hipStreamWaitEvent(ctx->side[k], ctx->fork, 0);
  if (e != hipSuccess) return cg::hip_fail(ctx, e, "batch fork");
  int rc = CG_OK;
  for (int i = 0; i < n && !rc; ++i) {
    ctx->stream = ctx->side[i % ns];
    rc = launch(i);
  }
  ctx->stream = main;
  for (int k = 0; k < ns && e == hipSuccess; ++k) {
    e = hipEventRecord(ctx->join[k], ctx->side[k]);
    if (e == hipSuccess) e = hipStreamWaitEvent(main, ctx->join[k], 0);
  }
  if (rc) return rc;
  return cg::hip_fail(ctx, e, "batch join");
}

int run_pass(cg_batch *b, const cg_batch::Pass &p)
{
  cg_ctx *ctx = b->ctx;
  int rc = fork_join(b, (int)p.fixed.size(), [&](int i) {
    const auto &g = p.fixed[i];
    hipEvent_t ev;
    cg::launch_begin(ctx, p.kclass, &ev);
    int r;
    if (p.kind == cg::FFT_ITEM_STRIDED)
      r = cg::fixed_strided_list(ctx, g.e, g.variant, p.dir, static_cast<const cg::StridedArgs *>(g.items),
                                 g.nitems, g.blocks);
    else
      r = cg::fixed_x_list(ctx, g.e, g.variant, p.kind == cg::FFT_ITEM_X_R2C, g.add,
                           static_cast<const cg::XArgs *>(g.items), g.nitems, g.blocks, g.nxp);
    cg::launch_end(ctx, p.kclass, ev);
    return r;
  });
  if (rc) return rc;
  return cg::launch_fft_batch(ctx, p.kind, p.dir, p.items, p.nitems, p.blocks, p.lds, p.kclass);
}

// Waits for the batch's work and frees everything it holds.
void release_batch(cg_batch *b)
{
  cg_ctx *ctx = b->ctx;
  (void)hipStreamSynchronize(ctx->stream);  // the side streams join into it
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  if (b->done) (void)hipEventDestroy(b->done);
  if (b->dev) {
    cg::dev_free(ctx, b->dev);
    ctx->in_use -= b->dev_bytes;
  }
  delete b;
}

int enqueue_batch(cg_batch *b)
{
  cg_ctx *ctx = b->ctx;
  int rc = CG_OK;
  if (b->ncounters) rc = cg::zero_counters(ctx, b->counters, b->ncounters);
  for (auto &p : b->fwd)
    if (!rc) rc = run_pass(b, p);
  if (!rc)
    rc = fork_join(b, (int)b->post.size(), [&](int i) {
      const auto &g = b->post[i];
      return cg::launch_posterior_batch(ctx, g.nt, g.sep, g.mixed, g.items, g.nitems, g.blocks);
    });
  for (auto &p : b->inv_mean)
    if (!rc) rc = run_pass(b, p);
  for (size_t r = 0; r < b->sim.size(); ++r) {
    const auto &cp = b->colour[r];
    if (!rc)
      rc = fork_join(b, 1 + (int)cp.fused.size(), [&](int k) {
        if (k == 0) return cg::colour_batch(ctx, cp.items, cp.nitems, cp.blocks);
        const auto &f = cp.fused[k - 1];
        return f.pair ? cg::fixed_colour_zinv2_list(ctx, f.e, f.items, f.nitems, f.blocks)
                      : cg::fixed_colour_zinv_list(ctx, f.e, f.items, f.nitems, f.blocks);
      });
    for (auto &p : b->sim[r])
      if (!rc) rc = run_pass(b, p);
  }
  return rc;
}

}  // namespace

extern "C" {

int cg_batch_create(cg_ctx *ctx, int n, const cg_inversion *inv, cg_batch **out)
{
  if (!ctx || !inv || !out || n < 1) return CG_EINVAL;
  *out = nullptr;
  DeviceGuard dg(ctx->device);
  // ---- validation
  for (int i = 0; i < n; ++i) {
    const cg_inversion &v = inv[i];
    const cg_post_args &a = v.post;
    const int nt = a.ntheta;
    if (nt < 1 || nt > cg::kMaxTheta) return cg::fail(ctx, CG_EINVAL, "batch: ntheta must be 1..8");
    if (!a.K || !a.e1 || !a.e2 || !a.e3 || !a.invert || !a.sigma0e || !a.prior_var0)
      return cg::fail(ctx, CG_EINVAL, "batch: missing table");
    if (a.prior_mode != CG_PRIOR_SEPARABLE || !a.rxy || !a.rt || !a.lam2d || a.err_corr)
      return cg::fail(ctx, CG_EINVAL, "batch: inversions use the separable prior with host rxy, rt, lam2d");
    if (a.precision != CG_PREC_REFERENCE && a.precision != CG_PREC_MIXED)
      return cg::fail(ctx, CG_EINVAL, "batch: unknown precision");
    const cg_grid *ref = v.mean_in[0];
    if (!ref) return cg::fail(ctx, CG_EINVAL, "batch: missing input grid");
    std::vector<const cg_grid *> all;
    for (int g = 0; g < 3; ++g) all.insert(all.end(), {v.mean_in[g], a.mean[g]});
    for (int l = 0; l < nt; ++l) all.insert(all.end(), {v.seis_in[l], a.data[l]});
    for (int g = 0; g < 6; ++g) all.push_back(a.cov[g]);
    if (v.nreal < 0 || (v.nreal > 0 && !v.out)) return cg::fail(ctx, CG_EINVAL, "batch: bad realisations");
    for (int r = 0; r < 3 * v.nreal; ++r) all.push_back(v.out[r]);
    for (const cg_grid *g : all) {
      if (!g || g->ctx->device != ctx->device || !same_shape(g, ref))
        return cg::fail(ctx, CG_EINVAL, "batch: bad grid");
    }
    for (int g = 0; g < 3; ++g)
      if (v.mean_in[g]->transformed) return cg::fail(ctx, CG_ESTATE, "batch: inputs must be in the real domain");
    for (int l = 0; l < nt; ++l)
      if (v.seis_in[l]->transformed) return cg::fail(ctx, CG_ESTATE, "batch: inputs must be in the real domain");
  }
  cg_batch *b = new (std::nothrow) cg_batch();
  if (!b) return CG_ENOMEM;
  b->ctx = ctx;
  b->n = n;
  b->inv.assign(inv, inv + n);
  b->outs.resize(n);
  int maxR = 0;
  for (int i = 0; i < n; ++i) {
    b->outs[i].assign(inv[i].out, inv[i].out + 3 * inv[i].nreal);
    b->inv[i].out = nullptr;
    maxR = std::max(maxR, inv[i].nreal);
  }
  // ---- host-side items (device pointers filled once the layout is allocated)
  std::vector<cg::FftItem> fx, fy, fz, iz, iy, ix;
  std::vector<std::vector<cg::FftItem>> sz(maxR), sy(maxR), sx(maxR);
  std::vector<cg::ColourItem> col_items;
  std::vector<int> col_count(maxR, 0);
  std::vector<const cg::FixedEntry *> fused_e(n, nullptr);  // per inversion: fused colouring entry
  std::vector<char> paired(n, 0);  // realisations 2m and 2m + 1 coloured by one paired launch
  size_t lds_fx = 0, lds_fy = 0, lds_fz = 0, lds_ix = 0, lds_iy = 0, lds_iz = 0;
  std::vector<size_t> lds_sx(maxR, 0), lds_sy(maxR, 0), lds_sz(maxR, 0);
  int rc = CG_OK;
  auto add_x = [&](std::vector<cg::FftItem> &v, size_t &lds, int kind, const float *src, float *grid,
                   const cg_grid *g, float scale, const float *add) {
    cg::FftItem it;
    std::memset(&it, 0, sizeof(it));
    it.src = src;
    it.grid = grid;
    it.add = add;
    it.nxp = g->nxp;
    it.rows = (long)g->nyp * g->nzp;
    it.scale = scale;
    size_t l = 0;
    if (!rc) rc = cg::plan_fft_item(ctx, kind, &it, &l);
    lds = std::max(lds, l);
    v.push_back(it);
  };
  auto add_s = [&](std::vector<cg::FftItem> &v, size_t &lds, cg_grid *g, int axis) {
    cg::FftItem it;
    std::memset(&it, 0, sizeof(it));
    const long plane = (long)g->cnxp * g->nyp;
    it.data = reinterpret_cast<float2 *>(g->d);
    if (axis == 1) {
      it.n = g->nyp; it.S = g->cnxp; it.inner = g->cnxp; it.outer = g->nzp; it.OS = plane;
    } else {
      it.n = g->nzp; it.S = plane; it.inner = plane; it.outer = 1; it.OS = 0;
    }
    if (it.n < 2) return;
    it.scale = 1.0f;
    size_t l = 0;
    if (!rc) rc = cg::plan_fft_item(ctx, cg::FFT_ITEM_STRIDED, &it, &l);
    lds = std::max(lds, l);
    v.push_back(it);
  };
  for (int i = 0; i < n && !rc; ++i) {
    const cg_inversion &v = b->inv[i];
    const cg_post_args &a = v.post;
    const cg_grid *ref = v.mean_in[0];
    const size_t N = (size_t)ref->nxp * ref->nyp * ref->nzp;
    const float sf = 1.0f / sqrtf((float)N);
    const float si = float(1.0 / sqrtf(float(N)));
    for (int g = 0; g < 3; ++g) add_x(fx, lds_fx, cg::FFT_ITEM_X_R2C, v.mean_in[g]->d, a.mean[g]->d, ref, sf, nullptr);
    for (int l = 0; l < a.ntheta; ++l)
      add_x(fx, lds_fx, cg::FFT_ITEM_X_R2C, v.seis_in[l]->d, a.data[l]->d, ref, sf, nullptr);
    for (int g = 0; g < 3; ++g) add_s(fy, lds_fy, a.mean[g], 1), add_s(fz, lds_fz, a.mean[g], 2);
    for (int l = 0; l < a.ntheta; ++l) add_s(fy, lds_fy, a.data[l], 1), add_s(fz, lds_fz, a.data[l], 2);
    for (int g = 0; g < 3; ++g) {
      add_s(iz, lds_iz, a.mean[g], 2);
      add_s(iy, lds_iy, a.mean[g], 1);
      add_x(ix, lds_ix, cg::FFT_ITEM_X_C2R, nullptr, a.mean[g]->d, ref, si, nullptr);
    }
    const cg::FixedEntry *fe = cg::fixed_colour_entry(ctx, ref->nzp, (long)ref->cnxp * ref->nyp);
    fused_e[i] = fe;
    paired[i] = fe && cg::fixed_colour2_entry(ctx, v.mean_in[0]->nzp, (long)v.mean_in[0]->cnxp * v.mean_in[0]->nyp);
    for (int r = 0; r < v.nreal; ++r)
      for (int g = 0; g < 3; ++g) {
        cg_grid *o = b->outs[i][3 * r + g];
        if (!fe) add_s(sz[r], lds_sz[r], o, 2);  // fused: the colouring kernel runs the z pass
        add_s(sy[r], lds_sy[r], o, 1);
        add_x(sx[r], lds_sx[r], cg::FFT_ITEM_X_C2R, nullptr, o->d, ref, si, a.mean[g]->d);
      }
  }
  if (rc) {
    delete b;
    return rc;
  }
  // ---- layout of the device area
  Layout L;
  // Each pass: grids grouped by compiled length (one launch per group), the rest left to the
  // runtime-plan batch kernel.
  struct PassBuild {
    int kind = 0;
    std::vector<cg::FftItem> generic;
    size_t off_generic = 0;
    struct G {
      const cg::FixedEntry *e;
      int variant;
      std::vector<cg::FftItem> items;
      size_t off;
    };
    std::vector<G> groups;
  };
  // Grids are grouped by compiled length and by the tile variant their own geometry uses alone
  // (the autotuned choice), so that every grid is transformed by the same kernel as in a plain call.
  auto plan_pass = [&](const std::vector<cg::FftItem> &v, int kind) {
    PassBuild pb;
    pb.kind = kind;
    std::map<std::pair<const cg::FixedEntry *, int>, size_t> index;
    for (const auto &it : v) {
      const cg::FixedEntry *e = nullptr;
      if (ctx->fft_engine == CG_FFT_AUTO) {
        if (kind == cg::FFT_ITEM_STRIDED) e = cg::fixed_find(cg::FIX_STRIDED, it.n);
        else if (it.nxp % 2 == 0) e = cg::fixed_find(cg::FIX_XE, it.nxp / 2);
        else e = cg::fixed_find(cg::FIX_XO, it.nxp);
      }
      if (!e) {
        pb.generic.push_back(it);
        continue;
      }
      const int var = kind == cg::FFT_ITEM_STRIDED
                          ? cg::fixed_variant(ctx, e, it.S, it.inner, it.inner * it.outer)
                          : cg::fixed_variant(ctx, e, it.rows, kind == cg::FFT_ITEM_X_R2C,
                                              kind != cg::FFT_ITEM_X_R2C && it.add != nullptr);
      auto f = index.find({e, var});
      if (f == index.end()) {
        f = index.emplace(std::make_pair(e, var), pb.groups.size()).first;
        pb.groups.push_back({e, var, {}, 0});
      }
      pb.groups[f->second].items.push_back(it);
    }
    pb.off_generic = L.add(sizeof(cg::FftItem) * pb.generic.size());
    const size_t esz = kind == cg::FFT_ITEM_STRIDED ? sizeof(cg::StridedArgs) : sizeof(cg::XArgs);
    for (auto &g : pb.groups) g.off = L.add(esz * g.items.size());
    return pb;
  };
  PassBuild o_fx = plan_pass(fx, cg::FFT_ITEM_X_R2C), o_fy = plan_pass(fy, cg::FFT_ITEM_STRIDED),
            o_fz = plan_pass(fz, cg::FFT_ITEM_STRIDED);
  PassBuild o_iz = plan_pass(iz, cg::FFT_ITEM_STRIDED), o_iy = plan_pass(iy, cg::FFT_ITEM_STRIDED),
            o_ix = plan_pass(ix, cg::FFT_ITEM_X_C2R);
  std::vector<PassBuild> o_sz(maxR), o_sy(maxR), o_sx(maxR);
  std::vector<size_t> o_col(maxR);
  for (int r = 0; r < maxR; ++r) {
    o_sz[r] = plan_pass(sz[r], cg::FFT_ITEM_STRIDED);
    o_sy[r] = plan_pass(sy[r], cg::FFT_ITEM_STRIDED);
    o_sx[r] = plan_pass(sx[r], cg::FFT_ITEM_X_C2R);
    int cnt = 0;
    for (int i = 0; i < n; ++i) cnt += b->inv[i].nreal > r && !fused_e[i];
    col_count[r] = cnt;
    o_col[r] = L.add(sizeof(cg::ColourItem) * cnt);
  }
  // fused colouring: per realisation number, one list per z length (inversions in order), and
  // at even r one list of pairs (r, r + 1) per z length for the paired inversions that have both
  std::map<const cg::FixedEntry *, std::vector<int>> fused_groups;
  for (int i = 0; i < n; ++i)
    if (fused_e[i]) fused_groups[fused_e[i]].push_back(i);
  // 0: not coloured at r by a fused launch; 1: single; 2: the first of a pair; 3: the second
  auto fused_kind = [&](int i, int r) {
    const int nr = b->inv[i].nreal;
    if (!fused_e[i] || nr <= r) return 0;
    if (!paired[i]) return 1;
    if (r % 2 == 1) return 3;
    return r + 1 < nr ? 2 : 1;
  };
  std::vector<std::vector<size_t>> o_fused(maxR), o_fused2(maxR);
  for (int r = 0; r < maxR; ++r)
    for (auto &kv : fused_groups) {
      int c = 0, c2 = 0;
      for (int i : kv.second) {
        c += fused_kind(i, r) == 1;
        c2 += fused_kind(i, r) == 2;
      }
      o_fused[r].push_back(L.add(sizeof(cg::ColourItem) * std::max(c, 1)));
      o_fused2[r].push_back(L.add(sizeof(cg::ColourItem) * std::max(c2, 1)));
    }
  // per-inversion posterior tables
  struct PT {
    std::vector<cg::PlaneTab<float>> t32;
    std::vector<cg::PlaneTab<double>> t64;
    size_t o32, o64, oinv, orxy, ort, olam, olist, ninv;
  };
  std::vector<PT> pt(n);
  const size_t o_counters = L.add(sizeof(unsigned long long) * n);
  for (int i = 0; i < n; ++i) {
    const cg_post_args &a = b->inv[i].post;
    const cg_grid *ref = b->inv[i].mean_in[0];
    const size_t plane = (size_t)ref->cnxp * ref->nyp;
    build_plane_tables<float>(&a, ref->nzp, pt[i].t32);
    const bool mixed = a.precision == CG_PREC_MIXED;
    if (mixed) build_plane_tables<double>(&a, ref->nzp, pt[i].t64);
    pt[i].ninv = 0;
    for (int k = 0; k < ref->nzp; ++k) pt[i].ninv += a.invert[k] != 0;
    pt[i].o32 = L.add(sizeof(cg::PlaneTab<float>) * ref->nzp);
    pt[i].o64 = mixed ? L.add(sizeof(cg::PlaneTab<double>) * ref->nzp) : 0;
    pt[i].oinv = L.add(sizeof(int) * ref->nzp);
    pt[i].orxy = L.add(sizeof(float) * plane);
    pt[i].ort = L.add(sizeof(float) * ref->nzp);
    pt[i].olam = L.add(sizeof(float) * plane);
    pt[i].olist = mixed ? L.add(sizeof(unsigned) * std::max<size_t>(pt[i].ninv * plane, 1)) : 0;
  }
  // posterior groups by (nt, precision)
  std::map<std::pair<int, int>, std::vector<int>> groups;
  for (int i = 0; i < n; ++i) groups[{b->inv[i].post.ntheta, b->inv[i].post.precision}].push_back(i);
  std::vector<size_t> o_groups;
  for (auto &kv : groups) o_groups.push_back(L.add(sizeof(cg::PostItem) * kv.second.size()));
  // ---- allocate
  hipError_t e = cg::dev_alloc(ctx, &b->dev, L.size);
  if (e != hipSuccess) {
    delete b;
    return cg::fail(ctx, CG_ENOMEM, std::string("batch: hipMalloc: ") + hipGetErrorString(e));
  }
  b->dev_bytes = L.size;
  ctx->in_use += L.size;
  ctx->peak = std::max(ctx->peak, ctx->in_use);
  char *dev = static_cast<char *>(b->dev);
  std::vector<char> host(L.size, 0);
  auto put = [&](size_t off, const void *src, size_t bytes) {
    if (bytes) std::memcpy(host.data() + off, src, bytes);
  };
  // ---- passes: block offsets, copy items
  auto finish = [&](PassBuild &pb, int dir, int kclass, size_t lds) {
    cg_batch::Pass p;
    unsigned blk = 0;
    for (auto &it : pb.generic) {
      it.block0 = blk;
      blk += it.nblocks;
    }
    put(pb.off_generic, pb.generic.data(), sizeof(cg::FftItem) * pb.generic.size());
    p.kind = pb.kind; p.dir = dir; p.kclass = kclass; p.nitems = (int)pb.generic.size();
    p.items = reinterpret_cast<const cg::FftItem *>(dev + pb.off_generic);
    p.blocks = blk;
    p.lds = lds;
    for (auto &g : pb.groups) {
      cg_batch::FixedGroup fg;
      fg.e = g.e;
      fg.nitems = (int)g.items.size();
      fg.items = dev + g.off;
      unsigned gb = 0;
      const auto &f0 = g.items.front();
      if (pb.kind == cg::FFT_ITEM_STRIDED) {
        fg.variant = g.variant;
        const int tile = g.e->tile[fg.variant];
        std::vector<cg::StridedArgs> a;
        for (const auto &it : g.items) {
          const unsigned long ncols = (unsigned long)it.inner * (unsigned long)it.outer;
          a.push_back({it.data, it.S, it.OS, (unsigned)it.inner, (unsigned)ncols, gb, it.scale});
          gb += (unsigned)((ncols + tile - 1) / tile);
        }
        put(g.off, a.data(), sizeof(cg::StridedArgs) * a.size());
      } else {
        const bool fwd = pb.kind == cg::FFT_ITEM_X_R2C;
        fg.add = !fwd && f0.add != nullptr;
        fg.variant = g.variant;
        const int tile = g.e->tile[fg.variant];
        std::vector<cg::XArgs> a;
        for (const auto &it : g.items) {
          a.push_back({fwd ? it.src : nullptr, it.grid, fwd ? nullptr : it.add, it.rows, gb, it.scale});
          gb += (unsigned)((it.rows + tile - 1) / tile);
          fg.nxp = it.nxp;
        }
        put(g.off, a.data(), sizeof(cg::XArgs) * a.size());
      }
      fg.blocks = gb;
      p.fixed.push_back(fg);
    }
    // largest launches first, so that the side streams finish together
    std::stable_sort(p.fixed.begin(), p.fixed.end(),
                     [](const cg_batch::FixedGroup &x, const cg_batch::FixedGroup &y) { return x.blocks > y.blocks; });
    return p;
  };
  b->fwd.push_back(finish(o_fx, -1, CG_K_FFT_X, lds_fx));
  b->fwd.push_back(finish(o_fy, -1, CG_K_FFT_Y, lds_fy));
  b->fwd.push_back(finish(o_fz, -1, CG_K_FFT_Z, lds_fz));
  b->inv_mean.push_back(finish(o_iz, +1, CG_K_FFT_Z, lds_iz));
  b->inv_mean.push_back(finish(o_iy, +1, CG_K_FFT_Y, lds_iy));
  b->inv_mean.push_back(finish(o_ix, +1, CG_K_FFT_X, lds_ix));
  b->sim.resize(maxR);
  b->colour.resize(maxR);
  for (int r = 0; r < maxR; ++r) {
    b->sim[r].push_back(finish(o_sz[r], +1, CG_K_FFT_Z, lds_sz[r]));
    b->sim[r].push_back(finish(o_sy[r], +1, CG_K_FFT_Y, lds_sy[r]));
    b->sim[r].push_back(finish(o_sx[r], +1, CG_K_FFT_X, lds_sx[r]));
    std::vector<cg::ColourItem> ci;
    unsigned blk = 0;
    auto colour_item = [&](int i) {
      cg::ColourItem it;
      std::memset(&it, 0, sizeof(it));
      const cg_grid *ref = b->inv[i].mean_in[0];
      it.p.csize = (long)ref->csize;
      it.p.cnxp = ref->cnxp;
      it.p.nyp = ref->nyp;
      it.p.nzp = ref->nzp;
      cg::colour_dims(it.p);
      for (int g = 0; g < 6; ++g) it.p.cov[g] = reinterpret_cast<const float2 *>(b->inv[i].post.cov[g]->d);
      for (int g = 0; g < 3; ++g) it.p.out[g] = reinterpret_cast<float2 *>(b->outs[i][3 * r + g]->d);
      it.p.seed = b->inv[i].seed;
      it.p.stream0 = 3u * (unsigned)r;
      return it;
    };
    size_t gi_f = 0;
    for (auto &kv : fused_groups) {
      for (int kind = 1; kind <= 2; ++kind) {
        std::vector<cg::ColourItem> fi;
        unsigned fb = 0;
        const int tile = kind == 2 ? kv.first->cz2_tile : kv.first->cz_tile;
        for (int i : kv.second) {
          if (fused_kind(i, r) != kind) continue;
          cg::ColourItem it = colour_item(i);
          if (kind == 2) {  // streams 3r .. 3r + 5 with r even: three whole Philox blocks
            for (int g = 0; g < 3; ++g) it.p.out2[g] = reinterpret_cast<float2 *>(b->outs[i][3 * (r + 1) + g]->d);
            it.p.stream1 = 3u * (unsigned)(r + 1);
          }
          it.block0 = fb;
          fb += (unsigned)(((long)it.p.cnxp * it.p.nyp + tile - 1) / tile);
          fi.push_back(it);
        }
        const size_t off = kind == 2 ? o_fused2[r][gi_f] : o_fused[r][gi_f];
        if (fi.empty()) continue;
        put(off, fi.data(), sizeof(cg::ColourItem) * fi.size());
        cg_batch::FusedColour fc;
        fc.e = kv.first;
        fc.items = reinterpret_cast<const cg::ColourItem *>(dev + off);
        fc.nitems = (int)fi.size();
        fc.blocks = fb;
        fc.pair = kind == 2;
        b->colour[r].fused.push_back(fc);
      }
      ++gi_f;
    }
    for (int i = 0; i < n; ++i) {
      if (b->inv[i].nreal <= r || fused_e[i]) continue;
      cg::ColourItem it;
      std::memset(&it, 0, sizeof(it));
      const cg_grid *ref = b->inv[i].mean_in[0];
      it.p.csize = (long)ref->csize;
      it.p.cnxp = ref->cnxp;
      it.p.nyp = ref->nyp;
      it.p.nzp = ref->nzp;
      cg::colour_dims(it.p);
      for (int g = 0; g < 6; ++g) it.p.cov[g] = reinterpret_cast<const float2 *>(b->inv[i].post.cov[g]->d);
      for (int g = 0; g < 3; ++g) it.p.out[g] = reinterpret_cast<float2 *>(b->outs[i][3 * r + g]->d);
      it.p.seed = b->inv[i].seed;
      it.p.stream0 = 3u * (unsigned)r;
      it.block0 = blk;
      blk += (unsigned)((ref->csize + cg::kThreads - 1) / cg::kThreads);
      ci.push_back(it);
    }
    put(o_col[r], ci.data(), sizeof(cg::ColourItem) * ci.size());
    b->colour[r].items = reinterpret_cast<const cg::ColourItem *>(dev + o_col[r]);
    b->colour[r].nitems = (int)ci.size();
    b->colour[r].blocks = blk;
  }
  // ---- posterior items
  b->counters = reinterpret_cast<unsigned long long *>(dev + o_counters);
  b->ncounters = n;
  size_t gi = 0;
  for (auto &kv : groups) {
    std::vector<cg::PostItem> items;
    unsigned blk = 0;
    for (int i : kv.second) {
      const cg_post_args &a = b->inv[i].post;
      const cg_grid *ref = b->inv[i].mean_in[0];
      const size_t plane = (size_t)ref->cnxp * ref->nyp;
      put(pt[i].o32, pt[i].t32.data(), sizeof(cg::PlaneTab<float>) * ref->nzp);
      if (!pt[i].t64.empty()) put(pt[i].o64, pt[i].t64.data(), sizeof(cg::PlaneTab<double>) * ref->nzp);
      put(pt[i].oinv, a.invert, sizeof(int) * ref->nzp);
      put(pt[i].orxy, a.rxy, sizeof(float) * plane);
      put(pt[i].ort, a.rt, sizeof(float) * ref->nzp);
      put(pt[i].olam, a.lam2d, sizeof(float) * plane);
      cg::PostItem it;
      std::memset(&it, 0, sizeof(it));
      cg::PostParams &p = it.p;
      p.nxp = ref->nxp; p.nyp = ref->nyp; p.nzp = ref->nzp; p.cnxp = ref->cnxp; p.nt = a.ntheta;
      p.csize = (long)ref->csize;
      for (int g = 0; g < 3; ++g) p.mean[g] = reinterpret_cast<float2 *>(a.mean[g]->d);
      for (int l = 0; l < a.ntheta; ++l) p.data[l] = reinterpret_cast<float2 *>(a.data[l]->d);
      for (int g = 0; g < 6; ++g) p.cov[g] = reinterpret_cast<float2 *>(a.cov[g]->d);
      p.tab32 = reinterpret_cast<const cg::PlaneTab<float> *>(dev + pt[i].o32);
      p.tab64 = pt[i].t64.empty() ? nullptr : reinterpret_cast<const cg::PlaneTab<double> *>(dev + pt[i].o64);
      p.invert = reinterpret_cast<const int *>(dev + pt[i].oinv);
      p.rxy = reinterpret_cast<const float *>(dev + pt[i].orxy);
      p.rt = reinterpret_cast<const float *>(dev + pt[i].ort);
      p.lam2d = reinterpret_cast<const float *>(dev + pt[i].olam);
      const double *pv = a.prior_var0;
      const double s6[6] = {pv[0], pv[4], pv[8], pv[1], pv[2], pv[5]};
      for (int g = 0; g < 6; ++g) {
        p.s6f[g] = (float)s6[g];
        p.s6d[g] = s6[g];
      }
      p.prior_mode = CG_PRIOR_SEPARABLE;
      p.want_residual = a.want_residual;
      p.precision = a.precision;
      p.cond_limit = (float)(a.cond_limit > 0 ? a.cond_limit : 1e6);
      p.list_len = b->counters + i;
      p.stats = nullptr;
      p.list = pt[i].t64.empty() ? nullptr : reinterpret_cast<unsigned *>(dev + pt[i].olist);
      it.bpp = (unsigned)((plane + cg::kThreads - 1) / cg::kThreads);
      it.block0 = blk;
      blk += it.bpp * (unsigned)ref->nzp;
      items.push_back(it);
    }
    put(o_groups[gi], items.data(), sizeof(cg::PostItem) * items.size());
    cg_batch::PostGroup pg;
    pg.nt = kv.first.first;
    pg.sep = true;
    pg.mixed = kv.first.second == CG_PREC_MIXED;
    pg.items = reinterpret_cast<const cg::PostItem *>(dev + o_groups[gi]);
    pg.nitems = (int)items.size();
    pg.blocks = blk;
    b->post.push_back(pg);
    ++gi;
  }
  e = hipMemcpyAsync(b->dev, host.data(), L.size, hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&b->done, hipEventDisableTiming);
  if (e == hipSuccess && ensure_side_streams(ctx) != CG_OK) e = hipErrorOutOfMemory;
  if (e != hipSuccess) {
    release_batch(b);
    return cg::hip_fail(ctx, e, "batch: upload of descriptors");
  }
  *out = b;
  return CG_OK;
}

int cg_batch_use_graph(cg_batch *b, int enable)
{
  if (!b) return CG_EINVAL;
  b->use_graph = enable != 0;
  return CG_OK;
}

int cg_batch_run(cg_batch *b)
{
  if (!b) return CG_EINVAL;
  cg_ctx *ctx = b->ctx;
  DeviceGuard dg(ctx->device);
  int rc = CG_OK;
  if (b->use_graph) {
    if (!b->exec) {
      const bool prof = ctx->profiling;
      ctx->profiling = false;  // no event records inside the captured sequence
      hipError_t e = hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal);
      if (e != hipSuccess) {
        ctx->profiling = prof;
        return cg::hip_fail(ctx, e, "hipStreamBeginCapture");
      }
      rc = enqueue_batch(b);
      e = hipStreamEndCapture(ctx->stream, &b->graph);
      ctx->profiling = prof;
      if (rc) return rc;
      if (e == hipSuccess) e = hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0);
      if (e != hipSuccess) return cg::hip_fail(ctx, e, "batch graph capture");
    }
    rc = cg::hip_fail(ctx, hipGraphLaunch(b->exec, ctx->stream), "hipGraphLaunch");
  } else {
    rc = enqueue_batch(b);
  }
  if (rc) return rc;
  for (int i = 0; i < b->n; ++i) {  // domains after a run
    cg_inversion &v = b->inv[i];
    for (int g = 0; g < 3; ++g) {
      v.post.mean[g]->transformed = 0;
      v.post.mean[g]->cubetype = CG_CT_PARAMETER;
    }
    for (int l = 0; l < v.post.ntheta; ++l) {
      v.post.data[l]->transformed = 1;
      v.post.data[l]->cubetype = CG_CT_DATA;
    }
    for (int g = 0; g < 6; ++g) {
      v.post.cov[g]->transformed = 1;
      v.post.cov[g]->cubetype = CG_CT_COVARIANCE;
    }
    for (cg_grid *o : b->outs[i]) {
      o->transformed = 0;
      o->cubetype = CG_CT_PARAMETER;
    }
  }
  return cg::hip_fail(ctx, hipEventRecord(b->done, ctx->stream), "hipEventRecord");
}

int cg_batch_query(cg_batch *b, int *done)
{
  if (!b || !done) return CG_EINVAL;
  DeviceGuard dg(b->ctx->device);
  const hipError_t e = hipEventQuery(b->done);
  if (e == hipSuccess) {
    *done = 1;
    return CG_OK;
  }
  if (e == hipErrorNotReady) {
    *done = 0;
    return CG_OK;
  }
  return cg::hip_fail(b->ctx, e, "hipEventQuery");
}

int cg_batch_wait(cg_batch *b)
{
  if (!b) return CG_EINVAL;
  DeviceGuard dg(b->ctx->device);
  return cg::hip_fail(b->ctx, hipEventSynchronize(b->done), "hipEventSynchronize");
}

int cg_batch_destroy(cg_batch *b)
{
  if (!b) return CG_EINVAL;
  DeviceGuard dg(b->ctx->device);
  release_batch(b);
  return CG_OK;
}

}  // extern "C"
