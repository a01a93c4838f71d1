/*
 * orc_post_impl.h -- per-wavenumber posterior and colouring, instantiated twice by
 * orc_post.c: REAL=float (the reference's FP32 arithmetic, operation for operation) and
 * REAL=double (the FP64 accuracy reference).  Test infrastructure only; see oracle.h.
 *
 * Restated reference code:
 *   avoinversion.cpp:1016-1083          per-wavenumber update, constant-thickness branch
 *   avoinversion.cpp:1236-1272          getNextErrorVariance
 *   seismicparametersholder.cpp:485-556 getNextParameterCovariance
 *   avoinversion.cpp:1421-1460          simulate: colouring of the noise
 *   libs/lib/lib_matr.c                 lib_matrCholCpx 78-167, lib_matrAXeqBMatCpx 330-365,
 *                                       lib_matrProdCholVec 453-469, products 524-736,
 *                                       add/subtract 744-783, lib_matrAdjoint 790-799
 */

#define CPX  CAT(cpx_, SUF)
#define CAT(a, b) CAT2(a, b)
#define CAT2(a, b) a##b

typedef struct { REAL re, im; } CPX;

/* lib_matrCholCpx: normalise by M[0][0], Cholesky with TOL, rescale L by sqrt(factor). */
static int CAT(chol_, SUF)(int n, CPX m[ORC_MAXT][ORC_MAXT])
{
  REAL factor = m[0][0].re;
  if (factor <= 0) return 1;
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++) {
      m[i][j].re = m[i][j].re / factor;
      m[i][j].im = m[i][j].im / factor;
    }
  for (int i = 0; i < n; i++) {
    if (m[i][i].re <= ORC_TOL) return 1;
    for (int j = 0; j < i; j++) {
      CPX lr = {0, 0};
      for (int k = 0; k < j; k++) {
        lr.re += (m[i][k].re * m[j][k].re) + (m[i][k].im * m[j][k].im);
        lr.im += -(m[i][k].re * m[j][k].im) + (m[i][k].im * m[j][k].re);
      }
      REAL help = m[j][j].re * m[j][j].re + m[j][j].im * m[j][j].im;
      m[i][j].re = ((m[i][j].re - lr.re) * m[j][j].re + (m[i][j].im - lr.im) * m[j][j].im) / help;
      /* the reference evaluates this with the freshly updated real part (m[j][j].im is 0) */
      m[i][j].im = ((m[i][j].im - lr.im) * m[j][j].re - (m[i][j].re - lr.re) * m[j][j].im) / help;
    }
    CPX lr = {0, 0};
    for (int k = 0; k < i; k++)
      lr.re += (m[i][k].re * m[i][k].re) + (m[i][k].im * m[i][k].im);
    lr.re = m[i][i].re - lr.re;
    if (lr.re <= ORC_TOL) return 1;
    m[i][i].re = (REAL)sqrt(lr.re);
    m[i][i].im = 0.0;
  }
  REAL sf = (REAL)sqrt(factor);
  for (int i = 0; i < n; i++)
    for (int j = 0; j <= i; j++) {
      m[i][j].re *= sf;
      m[i][j].im *= sf;
    }
  return 0;
}

/* lib_matrAXeqBMatCpx: forward then backward substitution with L and L^H, ncol columns. */
static void CAT(axeqb_, SUF)(int n, CPX L[ORC_MAXT][ORC_MAXT], CPX x[ORC_MAXT][3], int ncol)
{
  for (int c = 0; c < ncol; c++) {
    for (int i = 0; i < n; i++) {
      CPX lx = x[i][c];
      for (int j = 0; j < i; j++) {
        lx.re -= (x[j][c].re * L[i][j].re - x[j][c].im * L[i][j].im);
        lx.im -= (x[j][c].im * L[i][j].re + x[j][c].re * L[i][j].im);
      }
      REAL help = (L[i][i].re * L[i][i].re + L[i][i].im * L[i][i].im);
      x[i][c].re = (lx.re * L[i][i].re + lx.im * L[i][i].im) / help;
      x[i][c].im = (lx.im * L[i][i].re - lx.re * L[i][i].im) / help;
    }
    for (int i = n - 1; i >= 0; i--) {
      CPX lx = x[i][c];
      for (int j = n - 1; j > i; j--) {
        lx.re = lx.re - (x[j][c].re * L[j][i].re + x[j][c].im * L[j][i].im);
        lx.im = lx.im - (-x[j][c].re * L[j][i].im + x[j][c].im * L[j][i].re);
      }
      REAL help = (L[i][i].re * L[i][i].re + L[i][i].im * L[i][i].im);
      x[i][c].re = (lx.re * L[i][i].re - lx.im * L[i][i].im) / help;
      x[i][c].im = (lx.im * L[i][i].re + lx.re * L[i][i].im) / help;
    }
  }
}

long CAT(orc_posterior_, SUF)(const orc_post_spec *s, REAL *mean, REAL *data, REAL *cov, uint8_t *flags)
{
  const int nxp = s->nxp, nyp = s->nyp, nzp = s->nzp, nt = s->ntheta;
  const int cnxp = nxp / 2 + 1;
  const long csize = (long)cnxp * nyp * nzp;
  const double wnc = s->wnc;
  long nfail = 0;
  const double *pv = s->prior_var0;
  /* priorVar0_ entries in the order the covariance grids are read: 00 11 22 01 02 12 */
  const double s6[6] = {pv[0], pv[4], pv[8], pv[1], pv[2], pv[5]};

  for (int k = 0; k < nzp; k++) {
    const int inv = s->invert[k];
    CPX K[ORC_MAXT][3], e1[ORC_MAXT], e2[ORC_MAXT], e3[ORC_MAXT];
    for (int l = 0; l < nt; l++) {
      for (int j = 0; j < 3; j++) {
        K[l][j].re = (REAL)s->K[2 * ((k * nt + l) * 3 + j)];
        K[l][j].im = (REAL)s->K[2 * ((k * nt + l) * 3 + j) + 1];
      }
      e1[l].re = (REAL)s->e1[2 * (k * nt + l)];
      e1[l].im = (REAL)s->e1[2 * (k * nt + l) + 1];
      e2[l].re = (REAL)s->e2[2 * (k * nt + l)];
      e2[l].im = (REAL)s->e2[2 * (k * nt + l) + 1];
      e3[l].re = (REAL)s->e3[2 * (k * nt + l)];
      e3[l].im = (REAL)s->e3[2 * (k * nt + l) + 1];
    }
    for (int j = 0; j < nyp; j++) {
      for (int i = 0; i < cnxp; i++) {
        const long idx = i + (long)cnxp * (j + (long)nyp * k);
        CPX m3[3], dat[ORC_MAXT], res[ORC_MAXT];
        for (int g = 0; g < 3; g++) {
          m3[g].re = mean[2 * (g * csize + idx)];
          m3[g].im = mean[2 * (g * csize + idx) + 1];
        }
        for (int l = 0; l < nt; l++) {
          dat[l].re = data[2 * (l * csize + idx)];
          dat[l].im = data[2 * (l * csize + idx) + 1];
          res[l] = dat[l];
        }
        /* getNextParameterCovariance: |Re(c_ab)/s_ab| * s_ab, imaginary part zero */
        REAL pr[6];
        for (int g = 0; g < 6; g++) {
          REAL t;
          if (s->prior_mode == 0) {
            t = (REAL)cov[2 * (g * csize + idx)];
            if (s6[g] != 0) t = t / (REAL)s6[g];
          } else {
            t = (REAL)s->rxy[(long)j * cnxp + i] * (REAL)s->rt[k];
          }
          REAL a = (REAL)sqrt(t * t);
          pr[g] = a * (REAL)s6[g];
        }
        CPX P[3][3];
        P[0][0].re = pr[0]; P[1][1].re = pr[1]; P[2][2].re = pr[2];
        P[0][1].re = P[1][0].re = pr[3];
        P[0][2].re = P[2][0].re = pr[4];
        P[1][2].re = P[2][1].re = pr[5];
        for (int a = 0; a < 3; a++)
          for (int b = 0; b < 3; b++) P[a][b].im = 0;

        /* getNextErrorVariance */
        REAL lt = s->lam_grid ? (REAL)s->lam_grid[2 * idx] : (REAL)s->lam2d[(long)j * cnxp + i];
        REAL lam = (REAL)sqrt(lt * lt);
        CPX E[ORC_MAXT][ORC_MAXT];
        if (inv) {
          for (int l = 0; l < nt; l++)
            for (int m = 0; m < nt; m++) {
              const double sl = s->sigma0e[l * nt + m];
              E[l][m].re = (REAL)(0.5 * (1.0 - wnc) * sl * lam * (REAL)(e1[l].re * e1[m].re + e1[l].im * e1[m].im));
              E[l][m].re += (REAL)(0.5 * (1.0 - wnc) * sl * lam * (REAL)(e2[l].re * e2[m].re + e2[l].im * e2[m].im));
              if (l == m) {
                E[l][m].re += (REAL)(wnc * sl * e3[l].re * e3[l].re);
                E[l][m].im = 0.0;
              } else {
                E[l][m].im = (REAL)(0.5 * (1.0 - wnc) * sl * lam * (REAL)(-e1[l].re * e1[m].im + e1[l].im * e1[m].re));
                E[l][m].im += (REAL)(0.5 * (1.0 - wnc) * sl * lam * (REAL)(-e2[l].re * e2[m].im + e2[l].im * e2[m].re));
              }
            }
        }
        uint8_t flag = 0;
        if (inv) {
          flag = 1;
          CPX KS[ORC_MAXT][3], M[ORC_MAXT][ORC_MAXT];
          for (int l = 0; l < nt; l++) /* lib_matrProdCpx(K, parVar, n, 3, 3) */
            for (int b = 0; b < 3; b++) {
              CPX x = {0, 0};
              for (int a = 0; a < 3; a++) {
                x.re += K[l][a].re * P[a][b].re - K[l][a].im * P[a][b].im;
                x.im += K[l][a].im * P[a][b].re + K[l][a].re * P[a][b].im;
              }
              KS[l][b] = x;
            }
          for (int l = 0; l < nt; l++) /* lib_matrProdAdjointCpx(KS, K, n, 3, n) */
            for (int m = 0; m < nt; m++) {
              CPX x = {0, 0};
              for (int a = 0; a < 3; a++) {
                x.re += KS[l][a].re * K[m][a].re + KS[l][a].im * K[m][a].im;
                x.im += KS[l][a].im * K[m][a].re - KS[l][a].re * K[m][a].im;
              }
              M[l][m] = x;
            }
          for (int l = 0; l < nt; l++) /* lib_matrAddMatCpx(errVar, n, n, margVar) */
            for (int m = 0; m < nt; m++) {
              M[l][m].re += E[l][m].re;
              M[l][m].im += E[l][m].im;
            }
          if (CAT(chol_, SUF)(nt, M) == 0) {
            CPX KScc[3][ORC_MAXT];
            for (int l = 0; l < nt; l++) /* lib_matrAdjoint */
              for (int a = 0; a < 3; a++) {
                KScc[a][l].re = KS[l][a].re;
                KScc[a][l].im = -KS[l][a].im;
              }
            CAT(axeqb_, SUF)(nt, M, KS, 3);
            for (int a = 0; a < 3; a++) /* reduceVar = KScc * KS; parVar -= reduceVar */
              for (int b = 0; b < 3; b++) {
                CPX x = {0, 0};
                for (int l = 0; l < nt; l++) {
                  x.re += KScc[a][l].re * KS[l][b].re - KScc[a][l].im * KS[l][b].im;
                  x.im += KScc[a][l].im * KS[l][b].re + KScc[a][l].re * KS[l][b].im;
                }
                P[a][b].re -= x.re;
                P[a][b].im -= x.im;
              }
            for (int l = 0; l < nt; l++) { /* ijkData -= K * mean */
              CPX x = {0, 0};
              for (int a = 0; a < 3; a++) {
                x.re += K[l][a].re * m3[a].re - K[l][a].im * m3[a].im;
                x.im += K[l][a].im * m3[a].re + K[l][a].re * m3[a].im;
              }
              dat[l].re -= x.re;
              dat[l].im -= x.im;
            }
            CPX ans[3];
            for (int a = 0; a < 3; a++) { /* lib_matrProdAdjointMatVecCpx(KS, ijkData, 3, n) */
              CPX x = {0, 0};
              for (int l = 0; l < nt; l++) {
                x.re += KS[l][a].re * dat[l].re + KS[l][a].im * dat[l].im;
                x.im += -KS[l][a].im * dat[l].re + KS[l][a].re * dat[l].im;
              }
              ans[a] = x;
            }
            for (int a = 0; a < 3; a++) {
              m3[a].re += ans[a].re;
              m3[a].im += ans[a].im;
            }
            for (int l = 0; l < nt; l++) { /* ijkRes -= K * mean_post */
              CPX x = {0, 0};
              for (int a = 0; a < 3; a++) {
                x.re += K[l][a].re * m3[a].re - K[l][a].im * m3[a].im;
                x.im += K[l][a].im * m3[a].re + K[l][a].re * m3[a].im;
              }
              res[l].re -= x.re;
              res[l].im -= x.im;
            }
          } else {
            flag = 2;
            nfail++;
          }
        }
        if (flags) flags[idx] = flag;
        for (int g = 0; g < 3; g++) {
          mean[2 * (g * csize + idx)] = m3[g].re;
          mean[2 * (g * csize + idx) + 1] = m3[g].im;
        }
        const CPX out6[6] = {P[0][0], P[1][1], P[2][2], P[0][1], P[0][2], P[1][2]};
        for (int g = 0; g < 6; g++) {
          cov[2 * (g * csize + idx)] = out6[g].re;
          cov[2 * (g * csize + idx) + 1] = out6[g].im;
        }
        for (int l = 0; l < nt; l++) {
          data[2 * (l * csize + idx)] = res[l].re;
          data[2 * (l * csize + idx) + 1] = res[l].im;
        }
      }
    }
  }
  return nfail;
}

long CAT(orc_colour_, SUF)(long csize, const REAL *cov, REAL *seed)
{
  long nfail = 0;
  for (long idx = 0; idx < csize; idx++) {
    CPX P[ORC_MAXT][ORC_MAXT];
    const int map[6][2] = {{0, 0}, {1, 1}, {2, 2}, {0, 1}, {0, 2}, {1, 2}};
    for (int g = 0; g < 6; g++) {
      P[map[g][0]][map[g][1]].re = cov[2 * (g * csize + idx)];
      P[map[g][0]][map[g][1]].im = cov[2 * (g * csize + idx) + 1];
    }
    P[1][0].re = P[0][1].re; P[1][0].im = -P[0][1].im;
    P[2][0].re = P[0][2].re; P[2][0].im = -P[0][2].im;
    P[2][1].re = P[1][2].re; P[2][1].im = -P[1][2].im;
    CPX z[3];
    for (int g = 0; g < 3; g++) {
      z[g].re = seed[2 * (g * csize + idx)];
      z[g].im = seed[2 * (g * csize + idx) + 1];
    }
    CPX out[3] = {{0, 0}, {0, 0}, {0, 0}};
    if (CAT(chol_, SUF)(3, P) == 0) { /* lib_matrProdCholVec */
      for (int a = 0; a < 3; a++)
        for (int b = 0; b < a + 1; b++) {
          out[a].re += P[a][b].re * z[b].re - P[a][b].im * z[b].im;
          out[a].im += P[a][b].re * z[b].im + P[a][b].im * z[b].re;
        }
    } else {
      nfail++;
    }
    for (int g = 0; g < 3; g++) {
      seed[2 * (g * csize + idx)] = out[g].re;
      seed[2 * (g * csize + idx) + 1] = out[g].im;
    }
  }
  return nfail;
}

#undef CPX
