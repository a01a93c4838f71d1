/* A C caller of the ABI, compiled against include/cravagpu.h and linked with libcravagpu.so: the
 * host-only entry points and the argument checks that run before any device work, so that header
 * and library signatures are exercised outside any binding layer.  Exit status 0 on success; each
 * failed check prints its line. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "cravagpu.h"

static int failures = 0;
#define CHECK(cond)                                                   \
  do {                                                                \
    if (!(cond)) {                                                    \
      fprintf(stderr, "abi_check: %s:%d: %s\n", __FILE__, __LINE__, #cond); \
      ++failures;                                                     \
    }                                                                 \
  } while (0)

int main(void)
{
  /* FFTGrid::findClosestFactorableNumber: smallest 2^a 3^b 5^c 7^d >= n */
  CHECK(cg_factorable(1) == 1);
  CHECK(cg_factorable(11) == 12);
  CHECK(cg_factorable(178) == 180);
  CHECK(cg_factorable(563) == 567);
  CHECK(cg_factorable(1129) == 1134);
  /* arena sizes: a grid's in-place layout (2 (nxp/2+1) nyp nzp floats), 256-byte aligned */
  CHECK(cg_arena_grid_bytes(576, 576, 320) == (uint64_t)(2 * 289 * 576 * 320) * 4);
  CHECK(cg_arena_grid_bytes(7, 3, 5) % 256 == 0 && cg_arena_grid_bytes(7, 3, 5) >= 2 * 4 * 3 * 5 * 4);

  /* per-k_z tables: one angle, a spike wavelet w = delta_0: W(k) = 1, so K = D A with the
   * forward difference D = e^{2 pi i k / nzp} - 1, e1 = conj(W) / |w| = 1, e3 = 1 */
  enum { NZ = 20, NZP = 24 };
  float wav[NZP] = {0};
  wav[0] = 1.0f;
  const float A[3] = {0.5f, -0.25f, 0.375f};
  float K[NZP * 3 * 2], e1[NZP * 2], e2[NZP * 2], e3[NZP * 2];
  int invert[NZP];
  CHECK(cg_build_kz_tables(NZ, NZP, 4.0, 1, wav, A, 5.0f, 55.0f, 1.0f, K, e1, e2, e3, invert) == CG_OK);
  for (int k = 0; k < NZP; ++k) {
    const double w = 2.0 * 3.14159265358979323846 * k / NZP;
    const double dr = cos(w) - 1.0, di = sin(w);
    for (int j = 0; j < 3; ++j) {
      CHECK(fabs(K[2 * (3 * k + j)] - dr * A[j]) < 1e-5);
      CHECK(fabs(K[2 * (3 * k + j) + 1] - di * A[j]) < 1e-5);
    }
    CHECK(fabs(e1[2 * k] - 1.0) < 1e-6 && fabs(e1[2 * k + 1]) < 1e-6);
    CHECK(e3[2 * k] == 1.0f && e3[2 * k + 1] == 0.0f);
    /* f(k) = 1000 / (dt nzp) min(k, nzp - k) Hz inverted inside (5, 55) */
    const double f = 1000.0 / (4.0 * NZP) * (k < NZP - k ? k : NZP - k);
    CHECK(invert[k] == (f > 5.0 && f < 55.0));
  }
  CHECK(cg_build_kz_tables(NZ, NZP, 4.0, 0, wav, A, 5.0f, 55.0f, 1.0f, K, e1, e2, e3, invert) == CG_EINVAL);
  CHECK(cg_build_kz_tables(NZ, NZP, 4.0, 1, NULL, A, 5.0f, 55.0f, 1.0f, K, e1, e2, e3, invert) == CG_EINVAL);
  /* the non-constant-thickness tables: K = D A, e1 = 1 / |w| = 1 (a spike has norm 1) */
  CHECK(cg_build_kz_tables_variable_thickness(NZ, NZP, 4.0, 1, wav, A, 5.0f, 55.0f, 1.0f, K, e1, e2, e3, invert) ==
        CG_OK);
  for (int k = 0; k < NZP; ++k) {
    const double w = 2.0 * 3.14159265358979323846 * k / NZP;
    CHECK(fabs(K[2 * (3 * k)] - (cos(w) - 1.0) * A[0]) < 1e-5 && fabs(K[2 * (3 * k) + 1] - sin(w) * A[0]) < 1e-5);
    CHECK(fabs(e1[2 * k] - 1.0) < 1e-6 && e1[2 * k + 1] == 0.0f);
  }

  /* MakeCircAutoCovPosDef: a positive definite spectrum comes back unchanged */
  enum { N = 16 };
  float cac[3 * 3 * N];
  memset(cac, 0, sizeof(cac));
  for (int a = 0; a < 3; ++a) cac[(a * 3 + a) * N] = 1.0f + a;  /* white, lag 0 only */
  float keep[3 * 3 * N];
  memcpy(keep, cac, sizeof(cac));
  CHECK(cg_make_circ_auto_cov_pos_def(N, cac) == CG_OK);
  for (int i = 0; i < 9 * N; ++i) CHECK(fabs(cac[i] - keep[i]) < 1e-5);
  cac[0] = -1.0f; /* a lag-0 variance that is not positive */
  CHECK(cg_make_circ_auto_cov_pos_def(N, cac) == CG_EINVAL);

  /* argument checks that precede any device call */
  CHECK(cg_ctx_create(NULL, 0, 0) == CG_EINVAL);
  CHECK(cg_ctx_destroy(NULL) == CG_EINVAL);
  CHECK(cg_grid_destroy(NULL) == CG_EINVAL);
  CHECK(cg_arena_destroy(NULL) == CG_EINVAL);
  CHECK(cg_posterior(NULL, NULL) == CG_EINVAL);
  CHECK(cg_posterior_noalloc(NULL, NULL) == CG_EINVAL);
  CHECK(cg_simulate(NULL, NULL) == CG_EINVAL);
  CHECK(cg_fft_r2c_inplace(NULL) == CG_EINVAL);
  CHECK(cg_fft_c2r_inplace(NULL) == CG_EINVAL);
  CHECK(cg_grid_data_variance(NULL, NULL, NULL) == CG_EINVAL);
  CHECK(cg_grid_scale_wavelet(NULL, NULL) == CG_EINVAL);
  CHECK(cg_grid_write_storm(NULL, NULL, NULL, 0) == CG_EINVAL);
  CHECK(cg_grid_read_storm(NULL, NULL, NULL) == CG_EINVAL);
  CHECK(cg_grid_write_segy(NULL, NULL, NULL) == CG_EINVAL);
  CHECK(cg_grid_sum_squares(NULL, 0, NULL) == CG_EINVAL);
  CHECK(cg_batch_create(NULL, 0, NULL, NULL) == CG_EINVAL);
  CHECK(cg_batch_run(NULL) == CG_EINVAL);
  CHECK(cg_invert(NULL, NULL) == CG_EINVAL);
  CHECK(cg_grid_fill_from_traces(NULL, NULL) == CG_EINVAL);
  CHECK(cg_segy_read(NULL, NULL, NULL) == CG_EINVAL);
  {
    cg_local_noise *ln = (cg_local_noise *)1;
    CHECK(cg_local_noise_create(NULL, NULL, &ln) == CG_EINVAL && ln == (cg_local_noise *)1);
    CHECK(cg_local_noise_matrices(NULL, NULL, NULL) == CG_EINVAL);
    CHECK(cg_local_noise_correct(NULL, NULL, NULL) == CG_EINVAL);
    CHECK(cg_local_noise_correct_realisation(NULL, NULL, NULL) == CG_EINVAL);
    cg_local_noise_destroy(NULL); /* a no-op */
  }
  {
    int dims[6];
    CHECK(cg_grid_dims(NULL, dims) == CG_EINVAL);
  }
  if (failures == 0) printf("abi_check: ok\n");
  return failures == 0 ? 0 : 1;
}
