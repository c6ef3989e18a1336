/* AddressSanitizer / UndefinedBehaviorSanitizer run of the FP64 oracle (test
 * infrastructure): every entry point on small odd-sized problems, one thread.
 * Built and run by tests/test_sanitizers.py; exits 0 when every output is
 * finite (any sanitizer report aborts the run with a non-zero status). */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../oracle/avionic_oracle.h"

static unsigned long long rng = 88172645463325252ULL;
static double urand(void) {
  rng ^= rng << 13;
  rng ^= rng >> 7;
  rng ^= rng << 17;
  return (double)(rng >> 11) / 9007199254740992.0;
}
static double *cvec(long n) {
  double *v = malloc(sizeof(double) * 2 * n);
  for (long i = 0; i < 2 * n; ++i) v[i] = urand() - 0.5;
  return v;
}
static int finite_all(const double *v, long n) {
  for (long i = 0; i < n; ++i)
    if (!isfinite(v[i])) return 0;
  return 1;
}
#define CHECK(cond, what)                                 \
  do {                                                    \
    if (!(cond)) {                                        \
      fprintf(stderr, "oracle_check: %s failed\n", what); \
      return 1;                                           \
    }                                                     \
  } while (0)

static orc_params params(int method) {
  orc_params p;
  memset(&p, 0, sizeof p);
  p.method = method;
  p.maxIt = 12;
  p.sigma = p.tau = 1.0 / 3.0;
  p.sigmaTauRatio = 1.0;
  p.tsw = method == ORC_ICTGV2 ? 4.0 : 5.0;
  p.tsw2 = 0.5;
  p.lambda = 3.0;
  p.alpha0 = sqrt(2.0);
  p.alpha1 = 1.0;
  p.alpha = 0.5;
  p.gapEvery = 5;
  p.dt = 1.0;
  p.dx = p.dy = p.dz = 1.0;
  return p;
}

int main(void) {
  const int W = 9, H = 7, T = 3, C = 2;
  const long N = (long)W * H * T;
  double *b1 = cvec((long)C * H * W), *data = cvec((long)T * C * H * W);
  double *mask = malloc(sizeof(double) * N);
  for (long i = 0; i < N; ++i) mask[i] = urand() < 0.5 ? 1.0 : 0.0;
  /* FFTs of odd and composite lengths */
  for (int n = 1; n <= 30; ++n) {
    double *a = cvec(n);
    orc_fft1(a, n, -1);
    orc_fft1(a, n, +1);
    CHECK(finite_all(a, 2 * n), "orc_fft1");
    free(a);
  }
  double *img = cvec((long)H * W);
  orc_fft2(img, H, W, -1, 1);
  orc_fft2(img, H, W, +1, 0);
  CHECK(finite_all(img, 2L * H * W), "orc_fft2");
  /* the five 2D+t loops, centred and not, with gap logs */
  for (int m = 0; m <= 4; ++m)
    for (int cen = 0; cen <= 1; ++cen) {
      orc_params p = params(m);
      p.centered = cen;
      const int P = orc_primal_components(m);
      double *x = cvec(N), *all = malloc(sizeof(double) * 2 * N * P);
      double sig[12], tau[12], gap[8];
      int ng = 0;
      orc_recon(&p, W, H, C, T, data, b1, mask, x, all, sig, tau, gap, &ng, 1);
      CHECK(finite_all(x, 2 * N) && finite_all(all, 2 * N * P) && finite_all(sig, 12), "orc_recon");
      free(x);
      free(all);
    }
  /* volumetric TGV2 */
  {
    orc_params p = params(ORC_TGV2_3D);
    double *x = cvec(N), *d3 = cvec((long)C * N);
    orc_recon3(&p, W, H, T, C, d3, b1, mask, x, NULL, NULL, NULL, 1);
    CHECK(finite_all(x, 2 * N), "orc_recon3");
    free(x);
    free(d3);
  }
  /* radial: operator pair, exact NUDFT and loop */
  {
    const int M = 23, J = 3;
    const double osf = 1.5;
    double *traj = malloc(sizeof(double) * 2 * M * T), *dcf = malloc(sizeof(double) * M * T);
    for (long i = 0; i < 2L * M * T; ++i) traj[i] = urand() - 0.5;
    for (long i = 0; i < (long)M * T; ++i) dcf[i] = urand();
    double *x = cvec(N), *k = cvec((long)T * C * M), *k2 = cvec((long)T * C * M);
    orc_nufft_K(x, b1, traj, dcf, W, H, C, T, M, J, osf, k);
    orc_nufft_KH(k, b1, traj, dcf, W, H, C, T, M, J, osf, x);
    orc_nudft_K(x, b1, traj, dcf, W, H, C, T, M, k2);
    CHECK(finite_all(k, 2L * T * C * M) && finite_all(x, 2 * N) && finite_all(k2, 2L * T * C * M),
          "orc_nufft");
    orc_params p = params(ORC_TGV2);
    orc_recon_nc(&p, W, H, C, T, M, J, osf, traj, dcf, k, b1, x, NULL, NULL, NULL, 1);
    CHECK(finite_all(x, 2 * N), "orc_recon_nc");
    const double d6[6] = {1, 1, 1, 1, 1, 1};
    CHECK(isfinite(orc_opnorm_nc(ORC_TGV2, b1, W, H, C, T, M, J, osf, traj, dcf, d6, 5)),
          "orc_opnorm_nc");
    free(traj);
    free(dcf);
    free(x);
    free(k);
    free(k2);
  }
  /* operator norms and coil construction */
  {
    const double d6[6] = {1, 1, 1, 1, 1, 1};
    CHECK(isfinite(orc_opnorm(ORC_ICTGV2, b1, mask, W, H, C, T, 0, d6, 5)), "orc_opnorm");
    CHECK(isfinite(orc_opnorm3(b1, mask, W, H, T, C, d6, 5)), "orc_opnorm3");
    double prm[16] = {0.1, 1e-4, 8, 0.3, 0.3, 1.41, 1.0, 1e-4, 8, 0.3, 0.3, 1e-3, 8, 0, 0, 0};
    double *u = cvec((long)C * H * W), *b = cvec((long)C * H * W);  /* both [C][H][W] */
    int order[2];
    orc_coil_construction(data, mask, W, H, C, T, 1, prm, u, b, order);
    CHECK(finite_all(u, 2L * C * H * W) && finite_all(b, 2L * C * H * W), "orc_coil_construction");
    free(u);
    free(b);
  }
  printf("oracle_check: ok\n");
  free(b1);
  free(data);
  free(mask);
  free(img);
  return 0;
}
