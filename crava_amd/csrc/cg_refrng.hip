// cg_refrng.hip -- the reference's own noise stream, generated in parallel on the GPU.
//
// FFTGrid::fillInComplexNoise (fftgrid.cpp:997-1045) draws, in linear cell order, one or two
// normals per cell (none where it copies the conjugate of an earlier partner) from
// RandomGen::rnorm01 (random.cpp:38-111: Marsaglia-Bray), which consumes a variable number of
// uniforms u = seed * 2^-32 of the LCG seed <- 69069 seed + 1 (random.h:11-14,
// random.cpp:113-118).  Sequential as written; parallel here because the number of uniforms a
// normal consumes is a function of its starting offset in the uniform stream alone:
//   1. the stream is cut into chunks of kChunk offsets; for each chunk and each of kEntries
//      possible entry offsets (where the first normal starting in the chunk may begin), one
//      thread follows the normals through the chunk: exit offset (into the next chunk) and
//      normals started inside it;
//   2. groups of kGroup chunks compose those tables (one thread per group and entry);
//   3. the host chains the groups from offset 0 (a few hundred steps), walking exactly where an
//      entry falls outside the tables (a normal that consumed more than kEntries uniforms
//      across a boundary);
//   4. one thread per group resolves its chunks' entries, one thread per chunk regenerates its
//      normals into an array indexed by normal number, and one thread per cell picks its
//      normals (the number of normals drawn before a cell depends on indices only).
// The state word advances as the reference's process-wide seed does (by the uniforms consumed).
// Arithmetic is the reference's: double uniforms, double exp/log/sqrt, float rounding of
// std * normal.
#include <algorithm>
#include <cmath>
#include <vector>

#include "cg_internal.h"

// named (not anonymous) so that profilers show the kernels' names
namespace cgrng {

constexpr uint32_t kMul = 69069u;    // random.h:11
constexpr uint32_t kAdd = 1u;        // random.h:12
constexpr int kChunk = 4096;         // uniforms per chunk
constexpr int kEntries = 8;          // tabulated entry offsets per chunk
constexpr int kGroup = 256;          // chunks per group
constexpr int kMaxDraw = 1 << 20;    // uniforms one normal may consume before the stream is declared bad

struct Affine {
  uint32_t a, c;  // x -> a x + c (mod 2^32)
};

__host__ __device__ inline Affine compose(Affine f, Affine g) { return Affine{f.a * g.a, f.a * g.c + f.c}; }

// LCG state after k steps from s.
__host__ __device__ inline uint32_t lcg_skip(uint32_t s, unsigned long long k)
{
  Affine r{1u, 0u}, p{kMul, kAdd};
  while (k) {
    if (k & 1ull) r = compose(p, r);
    p = compose(p, p);
    k >>= 1;
  }
  return r.a * s + r.c;
}

__host__ __device__ inline double unif01(uint32_t &st)
{
  st = kMul * st + kAdd;
  return (double)st * (1.0 / 4294967296.0);  // random.h:14: 2^-30 / 4
}

__host__ __device__ inline double mb_g(double x)  // random.cpp:96-111
{
  const double a = 17.49731196, b = 2.36785163, c = 2.15787544;
  const double absx = fabs(x);
  double result = a * exp(-x * x / 2.0);
  if (absx < 1.0)
    result -= 2.0 * b * (3.0 - x * x) + c * (1.5 - absx);
  else if (absx < 1.5)
    result -= b * (3.0 - absx) * (3.0 - absx) + c * (1.5 - absx);
  else
    result -= b * (3.0 - absx) * (3.0 - absx);
  return result;
}

// One Marsaglia-Bray normal from state st (advanced); returns the uniforms consumed, 0 if a
// rejection loop ran past kMaxDraw.
__host__ __device__ inline int rnorm01(uint32_t &st, double &x)
{
  int n = 1;
  const double u = unif01(st);
  if (u < 0.8638) {
    const double u1 = unif01(st), u2 = unif01(st), u3 = unif01(st);
    x = 2.0 * (u1 + u2 + u3) - 3.0;
    return n + 3;
  }
  if (u < 0.9745) {
    const double u1 = unif01(st), u2 = unif01(st);
    x = 1.5 * (u1 + u2 - 1.0);
    return n + 2;
  }
  if (u < 0.9973002039) {
    double u2;
    do {
      const double u1 = unif01(st);
      u2 = unif01(st);
      x = 6.0 * u1 - 3.0;
      n += 2;
      if (n > kMaxDraw) return 0;
    } while (0.358 * u2 > mb_g(x));
    return n;
  }
  double s, t;
  do {
    double v1, v2, w;
    do {
      const double u1 = unif01(st), u2 = unif01(st);
      v1 = 2.0 * u1 - 1.0;
      v2 = 2.0 * u2 - 1.0;
      w = v1 * v1 + v2 * v2;
      n += 2;
      if (n > kMaxDraw) return 0;
    } while (w >= 1.0);
    const double c = sqrt((9.0 - 2.0 * log(w)) / w);
    s = c * v1;
    t = c * v2;
  } while ((fabs(s) <= 3.0) && (fabs(t) <= 3.0));
  x = (fabs(s) > 3.0) ? s : t;
  return n;
}

struct Hop {
  int exit;   // entry offset into the next chunk (or group); < 0: bad stream
  int count;  // normals started inside
};

// Normals starting at offsets in [pos, end), the first at pos: exit offset past end and count.
__host__ __device__ inline Hop walk(uint32_t seed, unsigned long long pos, unsigned long long end)
{
  uint32_t st = lcg_skip(seed, pos);
  int cnt = 0;
  double x;
  while (pos < end) {
    const int n = rnorm01(st, x);
    if (!n) return Hop{-1, cnt};
    pos += (unsigned)n;
    ++cnt;
  }
  return Hop{(int)(pos - end), cnt};
}

__global__ void k_chunk_tables(uint32_t seed, long nchunks, Hop *__restrict__ tab)
{
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nchunks * kEntries) return;
  const long c = t / kEntries;
  const int e = (int)(t % kEntries);
  const unsigned long long start = (unsigned long long)c * kChunk;
  tab[t] = walk(seed, start + (unsigned)e, start + kChunk);
}

// Follows entry `cur` through chunks [c0, c1): table lookups, exact walks for entries outside.
__device__ inline Hop follow(uint32_t seed, const Hop *tab, long c0, long c1, int cur)
{
  long total = 0;
  for (long c = c0; c < c1; ++c) {
    Hop h;
    if (cur < kEntries) {
      h = tab[c * kEntries + cur];
    } else {
      const unsigned long long start = (unsigned long long)c * kChunk;
      h = walk(seed, start + (unsigned)cur, start + kChunk);
    }
    if (h.exit < 0) return Hop{-1, 0};
    total += h.count;
    cur = h.exit;
  }
  return Hop{cur, (int)total};
}

__global__ void k_group_tables(uint32_t seed, long nchunks, const Hop *__restrict__ tab, Hop *__restrict__ gtab,
                               long ngroups)
{
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ngroups * kEntries) return;
  const long g = t / kEntries;
  const int e = (int)(t % kEntries);
  const long c0 = g * kGroup, c1 = c0 + kGroup < nchunks ? c0 + kGroup : nchunks;
  gtab[t] = follow(seed, tab, c0, c1, e);
}

// One thread per group: entry offset and first normal number of every chunk of the group.
__global__ void k_chunk_entries(uint32_t seed, long nchunks, const Hop *__restrict__ tab, const int *__restrict__ gentry,
                                const long long *__restrict__ gfirst, long ngroups, int *__restrict__ centry,
                                long long *__restrict__ cfirst)
{
  const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= ngroups) return;
  const long c0 = g * kGroup, c1 = c0 + kGroup < nchunks ? c0 + kGroup : nchunks;
  int cur = gentry[g];
  long long first = gfirst[g];
  for (long c = c0; c < c1; ++c) {
    centry[c] = cur;
    cfirst[c] = first;
    if (cur < 0) continue;  // past the end of the requested normals
    const Hop h = follow(seed, tab, c, c + 1, cur);
    if (h.exit < 0) {
      cur = -1;
      continue;
    }
    first += h.count;
    cur = h.exit;
  }
}

// One thread per chunk: its normals, numbered from cfirst, while below nnorm; the offset after
// the last requested normal goes to *end_pos.
__global__ void k_emit(uint32_t seed, long nchunks, const int *__restrict__ centry, const long long *__restrict__ cfirst,
                       long long nnorm, double *__restrict__ X, unsigned long long *__restrict__ end_pos)
{
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nchunks || centry[c] < 0 || cfirst[c] >= nnorm) return;
  const unsigned long long start = (unsigned long long)c * kChunk, end = start + kChunk;
  unsigned long long pos = start + (unsigned)centry[c];
  uint32_t st = lcg_skip(seed, pos);
  long long n = cfirst[c];
  double x;
  while (pos < end && n < nnorm) {
    const int used = rnorm01(st, x);
    if (!used) return;
    X[n++] = x;
    pos += (unsigned)used;
    if (n == nnorm) *end_pos = pos;
  }
}

// Draws of a special cell (x = 0 column, or the x = cnxp-1 column at an odd linear index):
// 1 if it is its own conjugate, 2 if its partner comes later, 0 if it copies an earlier one.
__host__ __device__ inline int special_draws(long jk, int nyp, int nzp, long *jkcc_out)
{
  const long jind = jk % nyp, kind = jk / nyp;
  const long jcc = jind == 0 ? 0 : nyp - jind;
  const long kcc = kind == 0 ? 0 : nzp - kind;
  const long jkcc = jcc + kcc * nyp;
  *jkcc_out = jkcc;
  return jkcc == jk ? 1 : (jkcc > jk ? 2 : 0);
}

__host__ __device__ inline bool is_special(long i, int cnxp)
{
  const long xs = i % cnxp;
  return xs == 0 || (xs == cnxp - 1 && (i % 2) == 1);
}

// Normals drawn before cell i: rowbase[jk] + those of the cells before it in its row (the
// x = 0 cell, then ordinary cells with two each; the x = cnxp-1 cell is the row's last).
__device__ inline long long draws_before(long i, int cnxp, int nyp, int nzp, const long long *rowbase)
{
  const long jk = i / cnxp, xs = i % cnxp;
  if (xs == 0) return rowbase[jk];
  long jkcc;
  const int d0 = special_draws(jk, nyp, nzp, &jkcc);
  return rowbase[jk] + d0 + 2ll * (xs - 1);
}

__global__ void k_noise_cells(float2 *__restrict__ out, long csize, int cnxp, int nyp, int nzp,
                              const long long *__restrict__ rowbase, const double *__restrict__ X)
{
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= csize) return;
  const double sd = (double)(float)(1.0 / sqrt(2.0));  // fftgrid.cpp:1003, a float
  long src = i;
  bool conj = false, self = false;
  if (is_special(i, cnxp)) {
    const long jk = i / cnxp, xs = i % cnxp;
    long jkcc;
    const int d = special_draws(jk, nyp, nzp, &jkcc);
    if (d == 1) {
      self = true;
    } else if (d == 0) {
      src = jkcc * cnxp + xs;  // an earlier cell that drew two normals
      conj = true;
    }
  }
  const long long b = draws_before(src, cnxp, nyp, nzp, rowbase);
  float2 v;
  if (self) {
    v = make_float2((float)X[b], 0.f);
  } else {
    v = make_float2((float)(sd * X[b]), (float)(sd * X[b + 1]));
    if (conj) v.y = -v.y;
  }
  out[i] = v;
}

}  // namespace cgrng

namespace cg {

using namespace cgrng;

// Fills `out` (cnxp x nyp x nzp complex) as FFTGrid::fillInComplexNoise does from the state
// *seed and advances *seed past the uniforms consumed.  Synchronises the context stream.
int fill_noise_reference(cg_ctx *ctx, float2 *out, int nxp, int nyp, int nzp, uint32_t *seed)
{
  const int cnxp = nxp / 2 + 1;
  const long rows = (long)nyp * nzp, csize = rows * cnxp;
  // normals drawn before each row (index arithmetic only)
  std::vector<long long> rowbase(rows + 1);
  long long acc = 0;
  for (long jk = 0; jk < rows; ++jk) {
    rowbase[jk] = acc;
    for (long xs = 0; xs < cnxp; xs += (cnxp > 1 ? cnxp - 1 : 1)) {
      const long i = jk * cnxp + xs;
      long jkcc;
      acc += is_special(i, cnxp) ? special_draws(jk, nyp, nzp, &jkcc) : 2;
    }
    acc += 2ll * (cnxp > 2 ? cnxp - 2 : 0);
  }
  rowbase[rows] = acc;
  const long long nnorm = acc;
  // uniforms: about 3.93 per normal on average; the margin covers the spread many times over
  const unsigned long long ulen = (unsigned long long)nnorm * 4ull + (1ull << 20);
  const long nchunks = (long)((ulen + kChunk - 1) / kChunk);
  const long ngroups = (nchunks + kGroup - 1) / kGroup;
  Hop *tab = nullptr, *gtab = nullptr;
  int *gentry = nullptr, *centry = nullptr;
  long long *gfirst = nullptr, *cfirst = nullptr, *rowbase_d = nullptr;
  double *X = nullptr;
  unsigned long long *end_pos = nullptr;
  const uint32_t s0 = *seed;
  auto release = [&]() {
    (void)hipStreamSynchronize(ctx->stream);
    for (void *p : {(void *)tab, (void *)gtab, (void *)gentry, (void *)centry, (void *)gfirst, (void *)cfirst,
                    (void *)rowbase_d, (void *)X, (void *)end_pos})
      dev_free(ctx, p);
  };
  hipError_t e = dev_alloc(ctx, (void **)&tab, sizeof(Hop) * nchunks * kEntries);
  if (e == hipSuccess) e = dev_alloc(ctx, (void **)&gtab, sizeof(Hop) * ngroups * kEntries);
  if (e == hipSuccess) e = dev_alloc(ctx, (void **)&gentry, sizeof(int) * ngroups);
  if (e == hipSuccess) e = dev_alloc(ctx, (void **)&gfirst, sizeof(long long) * ngroups);
  if (e == hipSuccess) e = dev_alloc(ctx, (void **)&centry, sizeof(int) * nchunks);
  if (e == hipSuccess) e = dev_alloc(ctx, (void **)&cfirst, sizeof(long long) * nchunks);
  if (e == hipSuccess) e = dev_alloc(ctx, (void **)&rowbase_d, sizeof(long long) * (rows + 1));
  if (e == hipSuccess) e = dev_alloc(ctx, (void **)&X, sizeof(double) * (nnorm + 2));
  if (e == hipSuccess) e = dev_alloc(ctx, (void **)&end_pos, sizeof(unsigned long long));
  if (e != hipSuccess) {
    release();
    return fail(ctx, CG_ENOMEM, "reference noise: device memory");
  }
  auto blocks = [](long n) { return (unsigned)((n + kThreads - 1) / kThreads); };
  hipEvent_t ev;
  launch_begin(ctx, CG_K_NOISE, &ev);
  k_chunk_tables<<<blocks(nchunks * kEntries), kThreads, 0, ctx->stream>>>(s0, nchunks, tab);
  k_group_tables<<<blocks(ngroups * kEntries), kThreads, 0, ctx->stream>>>(s0, nchunks, tab, gtab, ngroups);
  launch_end(ctx, CG_K_NOISE, ev);
  e = hipGetLastError();
  std::vector<Hop> gh(ngroups * kEntries);
  if (e == hipSuccess)
    e = hipMemcpyAsync(gh.data(), gtab, sizeof(Hop) * gh.size(), hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) {
    release();
    return hip_fail(ctx, e, "reference noise: chunk tables");
  }
  // chain the groups from offset 0 until every requested normal is placed
  std::vector<int> ge(ngroups, -1);
  std::vector<long long> gf(ngroups, nnorm);
  int cur = 0;
  long long first = 0;
  for (long g = 0; g < ngroups && first < nnorm; ++g) {
    ge[g] = cur;
    gf[g] = first;
    Hop h;
    if (cur < kEntries) {
      h = gh[g * kEntries + cur];
    } else {  // rare: walk the group's chunks exactly on the host
      const long c0 = g * kGroup, c1 = std::min<long>(c0 + kGroup, nchunks);
      long total = 0;
      int c_cur = cur;
      h = Hop{0, 0};
      for (long c = c0; c < c1; ++c) {
        const unsigned long long start = (unsigned long long)c * kChunk;
        const Hop w = walk(s0, start + (unsigned)c_cur, start + kChunk);
        if (w.exit < 0) {
          h.exit = -1;
          break;
        }
        total += w.count;
        c_cur = w.exit;
      }
      if (h.exit == 0) h = Hop{c_cur, (int)total};
    }
    if (h.exit < 0) {
      release();
      return fail(ctx, CG_EINVAL, "reference noise: a rejection loop exceeded its bound");
    }
    first += h.count;
    cur = h.exit;
  }
  if (first < nnorm) {
    release();
    return fail(ctx, CG_EINVAL, "reference noise: uniform stream shorter than the normals need");
  }
  e = hipMemcpyAsync(gentry, ge.data(), sizeof(int) * ngroups, hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(gfirst, gf.data(), sizeof(long long) * ngroups, hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(rowbase_d, rowbase.data(), sizeof(long long) * (rows + 1), hipMemcpyHostToDevice,
                       ctx->stream);
  if (e != hipSuccess) {
    release();
    return hip_fail(ctx, e, "reference noise: upload");
  }
  launch_begin(ctx, CG_K_NOISE, &ev);
  k_chunk_entries<<<blocks(ngroups), kThreads, 0, ctx->stream>>>(s0, nchunks, tab, gentry, gfirst, ngroups, centry,
                                                                 cfirst);
  k_emit<<<blocks(nchunks), kThreads, 0, ctx->stream>>>(s0, nchunks, centry, cfirst, nnorm, X, end_pos);
  k_noise_cells<<<blocks(csize), kThreads, 0, ctx->stream>>>(out, csize, cnxp, nyp, nzp, rowbase_d, X);
  launch_end(ctx, CG_K_NOISE, ev);
  unsigned long long endp = 0;
  e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(&endp, end_pos, sizeof(endp), hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  release();
  if (e != hipSuccess) return hip_fail(ctx, e, "reference noise");
  *seed = lcg_skip(s0, endp);
  return CG_OK;
}

}  // namespace cg
