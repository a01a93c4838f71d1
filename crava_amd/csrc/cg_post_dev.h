// cg_post_dev.h -- device side of the per-wavenumber posterior shared by the posterior passes
// (cg_post.hip: one wavenumber per lane over the natural spectra; cg_post_z.hip: the same
// update inside the pass that also runs the second z-transform stage): complex helpers, the
// register Cholesky factorisation with lib_matr's accept/reject rules, forward substitution,
// the update of one wavenumber, and the wave-aggregated statistics counter.
#pragma once

#include <hip/hip_runtime.h>

#include "cg_internal.h"
#include "cg_mem.h"
#include "cg_post.h"

// named (not anonymous) so that profilers show the kernels' names
namespace cgpost {

using cg::gld;
using cg::gst;
using cg::kat;

template <typename R>
struct cpx {
  R re, im;
};

template <typename R>
__device__ __forceinline__ cpx<R> mk(R a, R b)
{
  return cpx<R>{a, b};
}

constexpr double kTol = 1e-20;  // lib_matr.c:19

__device__ __forceinline__ int tri(int i, int j) { return i * (i + 1) / 2 + j; }  // j <= i

// Fused multiply-add in the working precision: the complex products below are written as chains
// of fma(-a, b, acc) so that each real product costs one instruction.
__device__ __forceinline__ float fmr(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fmr(double a, double b, double c) { return __builtin_fma(a, b, c); }

// Reciprocal: one v_rcp (about 1 ulp) in FP32, an exact division in FP64.
__device__ __forceinline__ float rcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ double rcp(double x) { return 1.0 / x; }


// Cholesky of the Hermitian M (lower triangle) with lib_matrCholCpx's accept/reject rules
// (lib_matr.c:78-167): fail if M00 <= 0, or if a diagonal entry or a pivot of M / M00 is
// <= TOL.  The reference divides M by M00, factors, and rescales L by sqrt(M00); factoring M
// directly and testing against TOL * M00 is the same decision without the two rescaling
// passes.  On success m holds L and rd[i] = 1 / L_ii.  cond_lb = max pivot / min pivot, a
// lower bound of cond(M).
template <int NT, typename R>
__device__ __forceinline__ int chol(cpx<R> (&m)[NT * (NT + 1) / 2], R (&rd)[NT], R &cond_lb)
{
  const R factor = m[0].re;
  if (!(factor > R(0))) return 1;
  const double tol = kTol * (double)factor;
  R pmin = R(1e30), pmax = R(0);
#pragma unroll
  for (int i = 0; i < NT; ++i) {
    if (!((double)m[tri(i, i)].re > tol)) return 1;
#pragma unroll
    for (int j = 0; j < i; ++j) {
      R lr = m[tri(i, j)].re, li = m[tri(i, j)].im;
#pragma unroll
      for (int k = 0; k < j; ++k) {  // a_ij - sum_k L_ik conj(L_jk)
        const cpx<R> a = m[tri(i, k)], b = m[tri(j, k)];
        lr = fmr(-a.re, b.re, fmr(-a.im, b.im, lr));
        li = fmr(-a.im, b.re, fmr(a.re, b.im, li));
      }
      m[tri(i, j)].re = lr * rd[j];
      m[tri(i, j)].im = li * rd[j];
    }
    R s = 0;
#pragma unroll
    for (int k = 0; k < i; ++k) s = fmr(m[tri(i, k)].re, m[tri(i, k)].re, fmr(m[tri(i, k)].im, m[tri(i, k)].im, s));
    const R piv = m[tri(i, i)].re - s;
    if (!((double)piv > tol)) return 1;
    pmin = piv < pmin ? piv : pmin;
    pmax = piv > pmax ? piv : pmax;
    const R l = sqrt(piv);
    m[tri(i, i)].re = l;
    m[tri(i, i)].im = 0;
    rd[i] = rcp(l);
  }
  cond_lb = pmax / pmin;
  return 0;
}

// x <- L^-1 x (forward substitution, as lib_matrLXeqBMatCpx, lib_matr.c:390-410).
template <int NT, typename R>
__device__ __forceinline__ void lsolve(const cpx<R> (&L)[NT * (NT + 1) / 2], const R (&rd)[NT], cpx<R> (&x)[NT])
{
#pragma unroll
  for (int i = 0; i < NT; ++i) {
    R xr = x[i].re, xi = x[i].im;
#pragma unroll
    for (int j = 0; j < i; ++j) {
      const cpx<R> l = L[tri(i, j)];
      xr = fmr(-x[j].re, l.re, fmr(x[j].im, l.im, xr));
      xi = fmr(-x[j].im, l.re, fmr(-x[j].re, l.im, xi));
    }
    x[i].re = xr * rd[i];
    x[i].im = xi * rd[i];
  }
}

// The update for one wavenumber in precision R.  With M = L L^H and Y = L^-1 (K Sigma_m):
//   (K Sigma_m)^H M^-1 (K Sigma_m) = Y^H Y,   X^H r = Y^H (L^-1 r)   (X = M^-1 K Sigma_m),
// which is the reference's M^-1 solve (lib_matrAXeqBMatCpx, forward then backward
// substitution) rearranged so that only forward substitutions remain.
// Returns 0 on success (outputs written), 1 on Cholesky failure (outputs untouched).
template <int NT, typename R, bool SEP, class Tab>
__device__ __forceinline__ int update(const Tab &T, R r, R lam, const R (&P6)[6], cpx<R> (&m3)[3],
                                      cpx<R> (&d)[NT], cpx<R> (&cov6)[6], cpx<R> (&res)[NT], bool want_res,
                                      R &cond_lb)
{
  // Separable prior: Y holds L^-1 S (S = K Sigma0), and K Sigma_m = r S enters the outputs as
  // Y^H Y r^2 and Y^H z r; the grid prior: Y holds L^-1 K Sigma_m itself (scale 1).
  cpx<R> Y[3][NT];
  cpx<R> M[NT * (NT + 1) / 2];
  if (SEP) {
#pragma unroll
    for (int l = 0; l < NT; ++l)
#pragma unroll
      for (int a = 0; a < 3; ++a) Y[a][l] = mk<R>(T.S[l][a].re, T.S[l][a].im);
#pragma unroll
    for (int l = 0; l < NT; ++l)
#pragma unroll
      for (int m = 0; m <= l; ++m) {
        const int t = tri(l, m);
        M[t].re = fmr(r, T.G[t].re, fmr(lam, T.Q[t].re, l == m ? T.D[l] : R(0)));
        M[t].im = fmr(r, T.G[t].im, lam * T.Q[t].im);
      }
  } else {
    const R Pm[3][3] = {{P6[0], P6[3], P6[4]}, {P6[3], P6[1], P6[5]}, {P6[4], P6[5], P6[2]}};
#pragma unroll
    for (int l = 0; l < NT; ++l)
#pragma unroll
      for (int b = 0; b < 3; ++b) {
        R xr = 0, xi = 0;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
          xr = fmr(T.K[l][a].re, Pm[a][b], xr);
          xi = fmr(T.K[l][a].im, Pm[a][b], xi);
        }
        Y[b][l] = mk<R>(xr, xi);
      }
#pragma unroll
    for (int l = 0; l < NT; ++l)
#pragma unroll
      for (int m = 0; m <= l; ++m) {
        const int t = tri(l, m);
        R xr = fmr(lam, T.Q[t].re, l == m ? T.D[l] : R(0)), xi = lam * T.Q[t].im;
#pragma unroll
        for (int a = 0; a < 3; ++a) {  // (K Sigma_m K^H)_lm = sum_a KS_la conj(K_ma)
          xr = fmr(Y[a][l].re, T.K[m][a].re, fmr(Y[a][l].im, T.K[m][a].im, xr));
          xi = fmr(Y[a][l].im, T.K[m][a].re, fmr(-Y[a][l].re, T.K[m][a].im, xi));
        }
        M[t].re = xr;
        M[t].im = xi;
      }
  }
  R rd[NT];
  if (chol<NT, R>(M, rd, cond_lb)) return 1;
  // z = L^-1 (d - K m)
  cpx<R> z[NT];
#pragma unroll
  for (int l = 0; l < NT; ++l) {
    R xr = d[l].re, xi = d[l].im;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      xr = fmr(-T.K[l][a].re, m3[a].re, fmr(T.K[l][a].im, m3[a].im, xr));
      xi = fmr(-T.K[l][a].im, m3[a].re, fmr(-T.K[l][a].re, m3[a].im, xi));
    }
    z[l] = mk<R>(xr, xi);
  }
  lsolve<NT, R>(M, rd, z);
#pragma unroll
  for (int a = 0; a < 3; ++a) lsolve<NT, R>(M, rd, Y[a]);
  // Sigma' = Sigma_m - (r^2) Y^H Y (entries 00 11 22 01 02 12), m' = m + (r) Y^H z
  const R s2 = SEP ? r * r : R(1), s1 = SEP ? r : R(1);
  const int ia[6] = {0, 1, 2, 0, 0, 1}, ib[6] = {0, 1, 2, 1, 2, 2};
#pragma unroll
  for (int g = 0; g < 6; ++g) {
    R xr = 0, xi = 0;
#pragma unroll
    for (int l = 0; l < NT; ++l) {
      const cpx<R> u = Y[ia[g]][l], v = Y[ib[g]][l];
      xr = fmr(u.re, v.re, fmr(u.im, v.im, xr));
      if (g >= 3) xi = fmr(u.re, v.im, fmr(-u.im, v.re, xi));
    }
    cov6[g] = mk<R>(fmr(-s2, xr, P6[g]), -s2 * xi);
  }
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    R xr = 0, xi = 0;
#pragma unroll
    for (int l = 0; l < NT; ++l) {
      xr = fmr(Y[a][l].re, z[l].re, fmr(Y[a][l].im, z[l].im, xr));
      xi = fmr(Y[a][l].re, z[l].im, fmr(-Y[a][l].im, z[l].re, xi));
    }
    m3[a].re = fmr(s1, xr, m3[a].re);
    m3[a].im = fmr(s1, xi, m3[a].im);
  }
  if (want_res) {
#pragma unroll
    for (int l = 0; l < NT; ++l) {
      R xr = d[l].re, xi = d[l].im;
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        xr = fmr(-T.K[l][a].re, m3[a].re, fmr(T.K[l][a].im, m3[a].im, xr));
        xi = fmr(-T.K[l][a].im, m3[a].re, fmr(-T.K[l][a].re, m3[a].im, xi));
      }
      res[l] = mk<R>(xr, xi);
    }
  }
  return 0;
}

// Wave-aggregated counter increment: one atomic per wavefront, spread over 64 slots by block
// so that waves do not serialise on one address.  Every lane of the wave must reach this
// point (the kernels below have no early returns).
__device__ __forceinline__ void wave_count(unsigned long long *slots, bool pred)
{
  const unsigned long long m = __ballot(pred);
  if (__lane_id() == 0 && m) atomicAdd(slots + (blockIdx.x & 63u), (unsigned long long)__popcll(m));
}

}  // namespace cgpost
