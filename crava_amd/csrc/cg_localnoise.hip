// cg_localnoise.hip -- local noise: the posterior and the realisations rescaled trace by trace
// where the data noise varies laterally (AVOInversion::correctVpVsRho, avoinversion.cpp:2146-2302,
// its per-trace matrices newPosteriorCovPointwise, :2028-2124, and the realisation scaling of
// simulate, :1485-1504).
//
// Per trace (i, j), with angle scales s_l = scale_l(i, j) / min scale_l, D = diag(sqrt s),
//   M          = G Sigma_m G^T + D Sigma_e D                      (n_theta x n_theta)
//   Delta      = Sigma_m G^T M^+ G Sigma_m      (M^+: eigenvalues <= 1e-8 dropped)
//   sigmanew   = Delta^1/2 (Sigma_m - postVar0)^-1/2          (eigenvalues <= 1e-7 dropped)
//   kept       = (Sigma_m - Delta)^1/2 postVar0^-1/2
// (square roots over the positive eigenvalues).  Every one of these is V f(Lambda) V^T of a
// symmetric matrix, which does not depend on the eigenvectors' signs or order, so a cyclic
// Jacobi solver in FP64 reproduces the reference's LAPACK results to rounding.  One thread per
// trace builds both 3 x 3 matrices; then one pass over the grids applies them sample by sample
// with the reference's arithmetic (float differences, double sums left to right, no FMA).
#include <cmath>
#include <cstring>
#include <vector>

#include "cg_internal.h"

struct cg_local_noise {
  cg_ctx *ctx = nullptr;
  int nx = 0, ny = 0, nt = 0;
  double *sig = nullptr;  // 18 planes of nx * ny: sigmanew (9, row-major entries), then kept (9)
};

namespace cgln {

// Cyclic Jacobi eigen decomposition of the symmetric n x n matrix a (upper triangle read; a is
// destroyed): w the eigenvalues, v the eigenvectors as columns (Golub & Van Loan, Matrix
// Computations, sec. 8.5: the symmetric Schur rotation, row-by-row sweeps).  Sweeps end when
// the off-diagonal mass is below 1e-32 of the diagonal's.
template <int N>
__host__ __device__ inline void jacobi(double (&a)[N][N], double (&w)[N], double (&v)[N][N])
{
  for (int i = 0; i < N; ++i)
    for (int j = 0; j < N; ++j) {
      v[i][j] = i == j ? 1.0 : 0.0;
      if (j < i) a[i][j] = a[j][i];
    }
  for (int sweep = 0; sweep < 32; ++sweep) {
    double off = 0.0, diag = 0.0;
    for (int p = 0; p < N; ++p) {
      diag += a[p][p] * a[p][p];
      for (int q = p + 1; q < N; ++q) off += a[p][q] * a[p][q];
    }
    if (!(off > 1e-32 * diag)) break;
    for (int p = 0; p < N - 1; ++p)
      for (int q = p + 1; q < N; ++q) {
        const double apq = a[p][q];
        if (apq == 0.0) continue;
        const double tau = (a[q][q] - a[p][p]) / (2.0 * apq);
        const double t = (tau >= 0.0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
        const double c = 1.0 / sqrt(1.0 + t * t), s = t * c;
        for (int k = 0; k < N; ++k) {
          const double akp = a[k][p], akq = a[k][q];
          a[k][p] = c * akp - s * akq;
          a[k][q] = s * akp + c * akq;
        }
        for (int k = 0; k < N; ++k) {
          const double apk = a[p][k], aqk = a[q][k];
          a[p][k] = c * apk - s * aqk;
          a[q][k] = s * apk + c * aqk;
        }
        a[p][q] = a[q][p] = 0.0;
        for (int k = 0; k < N; ++k) {
          const double vkp = v[k][p], vkq = v[k][q];
          v[k][p] = c * vkp - s * vkq;
          v[k][q] = s * vkp + c * vkq;
        }
      }
  }
  for (int i = 0; i < N; ++i) w[i] = a[i][i];
}

// out = V diag(f(w)) V^T of the symmetric m, f chosen by MODE:
//   0: 1/w where w > thr, else 0     1: sqrt(w) where w > 0, else 0     2: 1/sqrt(w) where w > thr, else 0
template <int N, int MODE>
__host__ __device__ inline void sym_fn(const double (&m)[N][N], double thr, double (&out)[N][N])
{
  double a[N][N], w[N], v[N][N];
  for (int i = 0; i < N; ++i)
    for (int j = 0; j < N; ++j) a[i][j] = m[i][j];
  jacobi<N>(a, w, v);
  double f[N];
  for (int i = 0; i < N; ++i) {
    if (MODE == 0) f[i] = w[i] > thr ? 1.0 / w[i] : 0.0;
    if (MODE == 1) f[i] = w[i] > 0.0 ? sqrt(w[i]) : 0.0;
    if (MODE == 2) f[i] = w[i] > thr ? 1.0 / sqrt(w[i]) : 0.0;
  }
  for (int i = 0; i < N; ++i)
    for (int j = 0; j < N; ++j) {
      double s = 0.0;
      for (int k = 0; k < N; ++k) s += v[i][k] * f[k] * v[j][k];
      out[i][j] = s;
    }
}

// Trace-independent parts, computed once on the host.
struct Consts {
  int nt;
  long ntr;                                           // traces, nx * ny
  double gsg[cg::kMaxTheta][cg::kMaxTheta];           // G Sigma_m G^T
  double e0[cg::kMaxTheta][cg::kMaxTheta];            // Sigma_e (errThetaCov)
  double smgt[3][cg::kMaxTheta];                      // Sigma_m G^T
  double g[cg::kMaxTheta][3];                         // G
  double sm[3][3];                                    // Sigma_m (priorVar0)
  double sd_isqrt[3][3];                              // (Sigma_m - postVar0)^-1/2
  double old_isqrt[3][3];                             // postVar0^-1/2
  double min_scale[cg::kMaxTheta];
};

template <int NT>
__global__ void __launch_bounds__(64) k_trace_matrices(Consts c, const double *__restrict__ scales,
                                                       double *__restrict__ sig)
{
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= c.ntr) return;
  double sq[NT];
  for (int l = 0; l < NT; ++l) sq[l] = sqrt(scales[l * c.ntr + t] / c.min_scale[l]);
  double m[NT][NT], mi[NT][NT];
  for (int l = 0; l < NT; ++l)
    for (int k = 0; k < NT; ++k) m[l][k] = c.gsg[l][k] + (sq[l] * c.e0[l][k]) * sq[k];
  sym_fn<NT, 0>(m, 1e-8, mi);
  // Delta = ((Sigma_m G^T) M^+ G) Sigma_m
  double h2[3][NT], h3[3][3], delta[3][3];
  for (int a = 0; a < 3; ++a)
    for (int k = 0; k < NT; ++k) {
      double s = 0.0;
      for (int l = 0; l < NT; ++l) s += c.smgt[a][l] * mi[l][k];
      h2[a][k] = s;
    }
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b) {
      double s = 0.0;
      for (int k = 0; k < NT; ++k) s += h2[a][k] * c.g[k][b];
      h3[a][b] = s;
    }
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b) {
      double s = 0.0;
      for (int k = 0; k < 3; ++k) s += h3[a][k] * c.sm[k][b];
      delta[a][b] = s;
    }
  double local[3][3], dsq[3][3], lsq[3][3];
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b) local[a][b] = c.sm[a][b] - delta[a][b];
  sym_fn<3, 1>(delta, 0.0, dsq);
  sym_fn<3, 1>(local, 0.0, lsq);
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b) {
      double s = 0.0, u = 0.0;
      for (int k = 0; k < 3; ++k) {
        s += dsq[a][k] * c.sd_isqrt[k][b];
        u += lsq[a][k] * c.old_isqrt[k][b];
      }
      sig[(long)(3 * a + b) * c.ntr + t] = s;
      sig[(long)(9 + 3 * a + b) * c.ntr + t] = u;
    }
}

constexpr int kZChunk = 16;  // samples of one trace per thread

// The posterior update of correctVpVsRho (avoinversion.cpp:2243-2263): inside the box,
// p <- float(m + s_a0 d0 + s_a1 d1 + s_a2 d2) with d = p - m in float and the sum in double.
__global__ void __launch_bounds__(cg::kThreads)
k_correct_posterior(int nx, int ny, int nz, int rnxp, int nyp, const double *__restrict__ sig, float *p0, float *p1,
                    float *p2, const float *__restrict__ m0, const float *__restrict__ m1,
                    const float *__restrict__ m2)
{
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
  if (i >= nx) return;
  const long ntr = (long)nx * ny, t = i + (long)nx * j;
  double s[9];
  for (int e = 0; e < 9; ++e) s[e] = sig[e * ntr + t];
  const int k0 = blockIdx.z * kZChunk, k1 = min(nz, k0 + kZChunk);
  for (int k = k0; k < k1; ++k) {
    const long idx = i + (long)rnxp * (j + (long)nyp * k);
    const float a = p0[idx], b = p1[idx], c = p2[idx];
    const float ma = m0[idx], mb = m1[idx], mc = m2[idx];
    const double d0 = (double)(a - ma), d1 = (double)(b - mb), d2 = (double)(c - mc);
    p0[idx] = (float)((double)ma + s[0] * d0 + s[1] * d1 + s[2] * d2);
    p1[idx] = (float)((double)mb + s[3] * d0 + s[4] * d1 + s[5] * d2);
    p2[idx] = (float)((double)mc + s[6] * d0 + s[7] * d1 + s[8] * d2);
  }
}

// The realisation scaling of simulate (avoinversion.cpp:1485-1504) and the posterior added
// after it (FFTGrid::add over the whole padded grid): inside the box
// r <- float(k_a0 r0 + k_a1 r1 + k_a2 r2) (double, left to right), then r += p in float everywhere.
__global__ void __launch_bounds__(cg::kThreads)
k_correct_realisation(int nx, int ny, int nz, int rnxp, int nyp, int nzp, const double *__restrict__ sig, float *r0,
                      float *r1, float *r2, const float *__restrict__ q0, const float *__restrict__ q1,
                      const float *__restrict__ q2)
{
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
  if (i >= rnxp) return;
  const bool col = i < nx && j < ny;
  const long ntr = (long)nx * ny, t = i + (long)nx * j;
  double s[9];
  for (int e = 0; e < 9; ++e) s[e] = col ? sig[(9 + e) * ntr + t] : 0.0;
  const int k0 = blockIdx.z * kZChunk, k1 = min(nzp, k0 + kZChunk);
  for (int k = k0; k < k1; ++k) {
    const long idx = i + (long)rnxp * (j + (long)nyp * k);
    float a = r0[idx], b = r1[idx], c = r2[idx];
    if (col && k < nz) {
      const double x = a, y = b, z = c;
      a = (float)(s[0] * x + s[1] * y + s[2] * z);
      b = (float)(s[3] * x + s[4] * y + s[5] * z);
      c = (float)(s[6] * x + s[7] * y + s[8] * z);
    }
    r0[idx] = a + q0[idx];
    r1[idx] = b + q1[idx];
    r2[idx] = c + q2[idx];
  }
}

// Grid2D::FindMin (libs/nrlib/grid/grid2d.hpp:182-191) in storage order.
double find_min(const double *v, long n, double missing)
{
  double m = v[0];
  for (long i = 0; i < n; ++i)
    if ((m == missing || v[i] < m) && v[i] != missing) m = v[i];
  return m;
}

template <int NT>
int launch_trace(cg_ctx *ctx, const Consts &c, const double *scales, double *sig)
{
  const unsigned blocks = (unsigned)((c.ntr + 63) / 64);
  k_trace_matrices<NT><<<blocks, 64, 0, ctx->stream>>>(c, scales, sig);
  return cg::hip_fail(ctx, hipGetLastError(), "k_trace_matrices");
}

int trace_matrices(cg_ctx *ctx, const Consts &c, const double *scales, double *sig)
{
  switch (c.nt) {
    case 1: return launch_trace<1>(ctx, c, scales, sig);
    case 2: return launch_trace<2>(ctx, c, scales, sig);
    case 3: return launch_trace<3>(ctx, c, scales, sig);
    case 4: return launch_trace<4>(ctx, c, scales, sig);
    case 5: return launch_trace<5>(ctx, c, scales, sig);
    case 6: return launch_trace<6>(ctx, c, scales, sig);
    case 7: return launch_trace<7>(ctx, c, scales, sig);
    case 8: return launch_trace<8>(ctx, c, scales, sig);
  }
  return cg::fail(ctx, CG_EINVAL, "local noise: 1 to 8 angles");
}

bool grids_ok(const cg_local_noise *ln, cg_grid *const *a, cg_grid *const *b)
{
  for (int g = 0; g < 3; ++g) {
    if (!a[g] || !b[g] || a[g]->ctx != ln->ctx || b[g]->ctx != ln->ctx) return false;
    if (a[g]->nx != ln->nx || a[g]->ny != ln->ny || a[g]->transformed || b[g]->transformed) return false;
    if (a[g]->rnxp != b[g]->rnxp || a[g]->nyp != b[g]->nyp || a[g]->nzp != b[g]->nzp || a[g]->nz != b[g]->nz ||
        a[g]->nx != b[g]->nx || a[g]->ny != b[g]->ny || a[g]->rnxp != a[0]->rnxp || a[g]->nyp != a[0]->nyp ||
        a[g]->nzp != a[0]->nzp || a[g]->nz != a[0]->nz)
      return false;
  }
  return true;
}

}  // namespace cgln

extern "C" {

int cg_local_noise_create(cg_ctx *ctx, const cg_local_noise_args *a, cg_local_noise **out)
{
  using namespace cgln;
  if (!ctx || !a || !out) return CG_EINVAL;
  *out = nullptr;
  const int nt = a->ntheta;
  if (a->nx < 1 || a->ny < 1 || nt < 1 || nt > cg::kMaxTheta || !a->G || !a->prior_var0 || !a->post_var0 ||
      !a->err_theta_cov || !a->noise_scale)
    return cg::fail(ctx, CG_EINVAL, "local noise: bad arguments (1 to 8 angles, all tables given)");
  for (int l = 0; l < nt; ++l)
    if (!a->noise_scale[l]) return cg::fail(ctx, CG_EINVAL, "local noise: missing scale grid");
  (void)hipSetDevice(ctx->device);
  Consts c;
  std::memset(&c, 0, sizeof(c));
  c.nt = nt;
  c.ntr = (long)a->nx * a->ny;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) c.sm[i][j] = a->prior_var0[3 * i + j];
  for (int l = 0; l < nt; ++l)
    for (int b = 0; b < 3; ++b) c.g[l][b] = a->G[3 * l + b];
  for (int l = 0; l < nt; ++l)
    for (int k = 0; k < nt; ++k) c.e0[l][k] = a->err_theta_cov[nt * l + k];
  // H1 = G Sigma_m, G Sigma_m G^T, Sigma_m G^T (avoinversion.cpp:2045-2049, 2068)
  double h1[cg::kMaxTheta][3];
  for (int l = 0; l < nt; ++l)
    for (int b = 0; b < 3; ++b) {
      double s = 0.0;
      for (int k = 0; k < 3; ++k) s += c.g[l][k] * c.sm[k][b];
      h1[l][b] = s;
    }
  for (int l = 0; l < nt; ++l)
    for (int m = 0; m < nt; ++m) {
      double s = 0.0;
      for (int k = 0; k < 3; ++k) s += h1[l][k] * c.g[m][k];
      c.gsg[l][m] = s;
    }
  for (int b = 0; b < 3; ++b)
    for (int l = 0; l < nt; ++l) {
      double s = 0.0;
      for (int k = 0; k < 3; ++k) s += c.sm[b][k] * c.g[l][k];
      c.smgt[b][l] = s;
    }
  double sdel[3][3], pv[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      pv[i][j] = a->post_var0[3 * i + j];
      sdel[i][j] = c.sm[i][j] - pv[i][j];
    }
  sym_fn<3, 2>(sdel, 1e-7, c.sd_isqrt);  // avoinversion.cpp:2106-2119
  sym_fn<3, 2>(pv, 0.0, c.old_isqrt);    // avoinversion.cpp:2182-2191 (lib_matr_eigen)
  for (int l = 0; l < nt; ++l) c.min_scale[l] = find_min(a->noise_scale[l], c.ntr, -99999.0);
  cg_local_noise *ln = new cg_local_noise;
  ln->ctx = ctx;
  ln->nx = a->nx;
  ln->ny = a->ny;
  ln->nt = nt;
  double *scales = nullptr;
  hipError_t e = cg::dev_alloc(ctx, (void **)&ln->sig, sizeof(double) * 18 * (size_t)c.ntr);
  if (e == hipSuccess) e = cg::dev_alloc(ctx, (void **)&scales, sizeof(double) * nt * (size_t)c.ntr);
  for (int l = 0; l < nt && e == hipSuccess; ++l)
    e = hipMemcpyAsync(scales + (size_t)l * c.ntr, a->noise_scale[l], sizeof(double) * c.ntr,
                       hipMemcpyHostToDevice, ctx->stream);
  int rc = e == hipSuccess ? CG_OK : cg::hip_fail(ctx, e, "local noise: tables");
  if (rc == CG_OK) {
    hipEvent_t ev;
    cg::launch_begin(ctx, CG_K_ELEMENTWISE, &ev);
    rc = trace_matrices(ctx, c, scales, ln->sig);
    cg::launch_end(ctx, CG_K_ELEMENTWISE, ev);
  }
  // the caller's scale arrays may go once this returns
  if (rc == CG_OK) rc = cg::hip_fail(ctx, hipStreamSynchronize(ctx->stream), "local noise: synchronise");
  if (scales) cg::dev_free(ctx, scales);
  if (rc != CG_OK) {
    if (ln->sig) cg::dev_free(ctx, ln->sig);
    delete ln;
    return rc;
  }
  *out = ln;
  return CG_OK;
}

int cg_local_noise_matrices(const cg_local_noise *ln, double *sigmanew, double *kept)
{
  if (!ln) return CG_EINVAL;
  cg_ctx *ctx = ln->ctx;
  (void)hipSetDevice(ctx->device);
  const size_t ntr = (size_t)ln->nx * ln->ny;
  std::vector<double> h(18 * ntr);
  hipError_t e = hipMemcpyAsync(h.data(), ln->sig, sizeof(double) * h.size(), hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) return cg::hip_fail(ctx, e, "local noise: matrices");
  for (size_t t = 0; t < ntr; ++t)
    for (int q = 0; q < 9; ++q) {
      if (sigmanew) sigmanew[9 * t + q] = h[q * ntr + t];
      if (kept) kept[9 * t + q] = h[(9 + q) * ntr + t];
    }
  return CG_OK;
}

int cg_local_noise_correct(cg_local_noise *ln, cg_grid *const post[3], cg_grid *const mean[3])
{
  if (!ln || !post || !mean) return CG_EINVAL;
  cg_ctx *ctx = ln->ctx;
  if (!cgln::grids_ok(ln, post, mean))
    return cg::fail(ctx, CG_EINVAL, "local noise: grids must be real-domain, of the scales' nx x ny and alike");
  (void)hipSetDevice(ctx->device);
  const cg_grid *g = post[0];
  const dim3 grid((unsigned)((g->nx + cg::kThreads - 1) / cg::kThreads), (unsigned)g->ny,
                  (unsigned)((g->nz + cgln::kZChunk - 1) / cgln::kZChunk));
  if (g->nz == 0) return CG_OK;
  hipEvent_t ev;
  cg::launch_begin(ctx, CG_K_ELEMENTWISE, &ev);
  cgln::k_correct_posterior<<<grid, cg::kThreads, 0, ctx->stream>>>(g->nx, g->ny, g->nz, g->rnxp, g->nyp, ln->sig,
                                                                     post[0]->d, post[1]->d, post[2]->d, mean[0]->d,
                                                                     mean[1]->d, mean[2]->d);
  cg::launch_end(ctx, CG_K_ELEMENTWISE, ev);
  return cg::hip_fail(ctx, hipGetLastError(), "k_correct_posterior");
}

int cg_local_noise_correct_realisation(cg_local_noise *ln, cg_grid *const real[3], cg_grid *const post[3])
{
  if (!ln || !real || !post) return CG_EINVAL;
  cg_ctx *ctx = ln->ctx;
  if (!cgln::grids_ok(ln, real, post))
    return cg::fail(ctx, CG_EINVAL, "local noise: grids must be real-domain, of the scales' nx x ny and alike");
  (void)hipSetDevice(ctx->device);
  const cg_grid *g = real[0];
  const dim3 grid((unsigned)((g->rnxp + cg::kThreads - 1) / cg::kThreads), (unsigned)g->nyp,
                  (unsigned)((g->nzp + cgln::kZChunk - 1) / cgln::kZChunk));
  hipEvent_t ev;
  cg::launch_begin(ctx, CG_K_ELEMENTWISE, &ev);
  cgln::k_correct_realisation<<<grid, cg::kThreads, 0, ctx->stream>>>(g->nx, g->ny, g->nz, g->rnxp, g->nyp, g->nzp,
                                                                       ln->sig, real[0]->d, real[1]->d, real[2]->d,
                                                                       post[0]->d, post[1]->d, post[2]->d);
  cg::launch_end(ctx, CG_K_ELEMENTWISE, ev);
  return cg::hip_fail(ctx, hipGetLastError(), "k_correct_realisation");
}

void cg_local_noise_destroy(cg_local_noise *ln)
{
  if (!ln) return;
  (void)hipSetDevice(ln->ctx->device);
  (void)hipStreamSynchronize(ln->ctx->stream);
  cg::dev_free(ln->ctx, ln->sig);
  delete ln;
}

}  // extern "C"
