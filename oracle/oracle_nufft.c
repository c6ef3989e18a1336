/*
 * FP64 oracle of the non-Cartesian (radial) operator pair.
 * TEST INFRASTRUCTURE ONLY (see avionic_oracle.h).
 *
 * Operator (matlab/backward_opt_noncart.m:2-21, forward_opt_noncart.m:2-18,
 * with the NUFFT of the Fessler IRT toolbox, which is not in the reference
 * tree: its Kaiser-Bessel gridding is restated here from the published
 * method, Jackson et al. IEEE TMI 1991 and Beatty et al. IEEE TMI 2005):
 *
 *   (K x)_{t,c,m} = sqrt(w_{t,m}) / sqrt(HW) *
 *        sum_{y,x} b1_c[y,x] x_t[y,x] exp(-2 pi i (kx_m (x - W/2) + ky_m (y - H/2)))
 *   K^H its adjoint,
 *
 * with k in cycles per pixel in [-0.5, 0.5).  orc_nudft_K evaluates this sum
 * exactly; orc_nufft_K / orc_nufft_KH approximate it by Kaiser-Bessel gridding
 * on a grid oversampled by osf (kernel width J, beta of Beatty et al.),
 * deapodised with the kernel's continuous Fourier transform.  The gridded pair
 * is exactly adjoint by construction (the adjoint test pins that), and its
 * distance to the exact sum is bounded in the tests.
 */
#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "avionic_oracle.h"

typedef double complex dc;

/* Modified Bessel function I0 (power series; converges for all x). */
static double bessel_i0(double x) {
  double s = 1.0, t = 1.0, q = 0.25 * x * x;
  for (int k = 1; k < 200; k++) {
    t *= q / ((double)k * k);
    s += t;
    if (t < 1e-17 * s) break;
  }
  return s;
}

/* beta = pi sqrt((J/osf)^2 (osf - 1/2)^2 - 0.8)  (Beatty et al. 2005, eq. 5) */
double orc_kb_beta(int J, double osf) {
  double a = ((double)J / osf) * (osf - 0.5);
  return M_PI * sqrt(a * a - 0.8);
}

/* psi(u) = I0(beta sqrt(1 - (2u/J)^2)) on |u| <= J/2. */
static double kb(double u, int J, double beta) {
  double r = 2.0 * u / J;
  if (fabs(r) > 1.0) return 0.0;
  return bessel_i0(beta * sqrt(1.0 - r * r));
}

/* Continuous FT of psi at nu (cycles per grid sample):
 * J sinh(sqrt(beta^2 - (pi J nu)^2)) / sqrt(...), or sin(...) past the root. */
static double kb_ft(double nu, int J, double beta) {
  double a = M_PI * J * nu;
  double d = beta * beta - a * a;
  if (d > 1e-12) {
    double s = sqrt(d);
    return J * sinh(s) / s;
  } else if (d < -1e-12) {
    double s = sqrt(-d);
    return J * sin(s) / s;
  }
  return (double)J;
}

static int grid_len(int n, double osf) { return (int)lround(osf * n); }

/* wrap an integer into [0, G) */
static inline int wrapi(int q, int G) {
  int r = q % G;
  return r < 0 ? r + G : r;
}

void orc_nudft_K(const double *xr, const double *b1r, const double *traj,
                 const double *dcf, int W, int H, int C, int T, int M,
                 double *kr) {
  const dc *x = (const dc *)xr, *b1 = (const dc *)b1r;
  dc *kd = (dc *)kr;
  long HW = (long)H * W;
  double sc = 1.0 / sqrt((double)HW);
#pragma omp parallel for collapse(2) schedule(static)
  for (int t = 0; t < T; t++)
    for (int m = 0; m < M; m++) {
      double kx = traj[((long)t * 2 + 0) * M + m], ky = traj[((long)t * 2 + 1) * M + m];
      double wt = dcf ? sqrt(dcf[(long)t * M + m]) : 1.0;
      for (int c = 0; c < C; c++) {
        dc acc = 0;
        for (int y = 0; y < H; y++)
          for (int xx = 0; xx < W; xx++) {
            double ph = -2.0 * M_PI * (kx * (xx - W / 2) + ky * (y - H / 2));
            acc += b1[c * HW + (long)y * W + xx] * x[t * HW + (long)y * W + xx] *
                   (cos(ph) + I * sin(ph));
          }
        kd[((long)t * C + c) * M + m] = acc * wt * sc;
      }
    }
}

void orc_nufft_K(const double *xr, const double *b1r, const double *traj,
                 const double *dcf, int W, int H, int C, int T, int M, int J,
                 double osf, double *kr) {
  const dc *x = (const dc *)xr, *b1 = (const dc *)b1r;
  dc *kd = (dc *)kr;
  const int Gx = grid_len(W, osf), Gy = grid_len(H, osf);
  const long HW = (long)H * W, G2 = (long)Gx * Gy;
  const double beta = orc_kb_beta(J, osf), sc = 1.0 / sqrt((double)HW);
  double *dx = malloc(sizeof(double) * W), *dy = malloc(sizeof(double) * H);
  for (int i = 0; i < W; i++) dx[i] = kb_ft((double)(i - W / 2) / Gx, J, beta);
  for (int i = 0; i < H; i++) dy[i] = kb_ft((double)(i - H / 2) / Gy, J, beta);
#pragma omp parallel
  {
    dc *grid = malloc(sizeof(dc) * G2);
#pragma omp for collapse(2) schedule(static)
    for (int t = 0; t < T; t++)
      for (int c = 0; c < C; c++) {
        memset(grid, 0, sizeof(dc) * G2);
        for (int y = 0; y < H; y++)
          for (int xx = 0; xx < W; xx++)
            grid[(long)wrapi(y - H / 2, Gy) * Gx + wrapi(xx - W / 2, Gx)] =
                b1[c * HW + (long)y * W + xx] * x[t * HW + (long)y * W + xx] / (dy[y] * dx[xx]);
        orc_fft2((double *)grid, Gy, Gx, -1, 0);
        for (int m = 0; m < M; m++) {
          double sx = traj[((long)t * 2 + 0) * M + m] * Gx;
          double sy = traj[((long)t * 2 + 1) * M + m] * Gy;
          dc acc = 0;
          for (int qy = (int)ceil(sy - J / 2.0); qy <= (int)floor(sy + J / 2.0); qy++) {
            double wy = kb(sy - qy, J, beta);
            for (int qx = (int)ceil(sx - J / 2.0); qx <= (int)floor(sx + J / 2.0); qx++)
              acc += wy * kb(sx - qx, J, beta) * grid[(long)wrapi(qy, Gy) * Gx + wrapi(qx, Gx)];
          }
          double wt = dcf ? sqrt(dcf[(long)t * M + m]) : 1.0;
          kd[((long)t * C + c) * M + m] = acc * wt * sc;
        }
      }
    free(grid);
  }
  free(dx);
  free(dy);
}

void orc_nufft_KH(const double *kr, const double *b1r, const double *traj,
                  const double *dcf, int W, int H, int C, int T, int M, int J,
                  double osf, double *xr) {
  const dc *kd = (const dc *)kr, *b1 = (const dc *)b1r;
  dc *x = (dc *)xr;
  const int Gx = grid_len(W, osf), Gy = grid_len(H, osf);
  const long HW = (long)H * W, G2 = (long)Gx * Gy;
  const double beta = orc_kb_beta(J, osf), sc = 1.0 / sqrt((double)HW);
  double *dx = malloc(sizeof(double) * W), *dy = malloc(sizeof(double) * H);
  for (int i = 0; i < W; i++) dx[i] = kb_ft((double)(i - W / 2) / Gx, J, beta);
  for (int i = 0; i < H; i++) dy[i] = kb_ft((double)(i - H / 2) / Gy, J, beta);
#pragma omp parallel
  {
    dc *grid = malloc(sizeof(dc) * G2);
#pragma omp for schedule(static)
    for (int t = 0; t < T; t++) {
      for (long i = 0; i < HW; i++) x[t * HW + i] = 0;
      for (int c = 0; c < C; c++) {
        memset(grid, 0, sizeof(dc) * G2);
        for (int m = 0; m < M; m++) {
          double sx = traj[((long)t * 2 + 0) * M + m] * Gx;
          double sy = traj[((long)t * 2 + 1) * M + m] * Gy;
          double wt = dcf ? sqrt(dcf[(long)t * M + m]) : 1.0;
          dc v = kd[((long)t * C + c) * M + m] * wt * sc;
          for (int qy = (int)ceil(sy - J / 2.0); qy <= (int)floor(sy + J / 2.0); qy++) {
            double wy = kb(sy - qy, J, beta);
            for (int qx = (int)ceil(sx - J / 2.0); qx <= (int)floor(sx + J / 2.0); qx++)
              grid[(long)wrapi(qy, Gy) * Gx + wrapi(qx, Gx)] += wy * kb(sx - qx, J, beta) * v;
          }
        }
        orc_fft2((double *)grid, Gy, Gx, +1, 0);
        for (int y = 0; y < H; y++)
          for (int xx = 0; xx < W; xx++)
            x[t * HW + (long)y * W + xx] +=
                conj(b1[c * HW + (long)y * W + xx]) *
                grid[(long)wrapi(y - H / 2, Gy) * Gx + wrapi(xx - W / 2, Gx)] / (dy[y] * dx[xx]);
      }
    }
    free(grid);
  }
  free(dx);
  free(dy);
}
