/* FP64 CPU oracle of the coil-sensitivity estimation (TEST INFRASTRUCTURE
 * ONLY: the checker of the HIP path, never linked into the product).
 *
 * Restates CoilConstruction (CUDA/src/coil_construction.cpp:22-557) and
 * CartesianCoilConstruction::TimeAveragedReconstruction
 * (cartesian_coil_construction.cpp:23-87):
 *   1. time-averaged coil images crec_c and their sum-of-squares u0;
 *   2. |b1_c| by H1 regularisation: (mu |u0|^2 + L) x_c = mu |u0| |crec_c|
 *      solved by conjugate gradients, L the periodic 2D Laplacian, then
 *      normalised to unit sum of squares (B1FromUH1, :363-446);
 *   3. coils taken in a greedy order (FindMaximumSum :52-73, then the coil
 *      whose |crec| overlaps most with the previous one); per coil a TGV2
 *      image estimate u from the coils so far (UTGV2Recon :75-216) and the
 *      phase of the next coil map from a quadratically regularised fit
 *      (B1Recon :217-362);
 *   4. a final B1Recon over all coils.
 * The 2D stencils follow utils::Gradient2D / Divergence2D /
 * SymmetricGradient2D / SymmetricDivergence2D (utils.cpp:156-179, 525-566,
 * 817-845, 1078-1115): forward differences zero at the last index, backward
 * differences u_0 / u_i - u_{i-1} / -u_{n-2}; the H1 operator uses periodic
 * differences (cg_forward_operation.h:46-67, diff3 / bdiff3 with wrap).
 * The conjugate gradient solver is the textbook method (the reference uses
 * AGILE's ConjugateGradient, not vendored) with the reference's tolerances
 * 1e-12 (relative and absolute, on the squared residual norm) and at most 800
 * iterations.
 */
#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "avionic_oracle.h"

typedef double complex dc;

static inline double a2(dc z) { return creal(z) * creal(z) + cimag(z) * cimag(z); }

/* ---------------------------------------------------------------- 2D stencils */
static void grad2(const dc *u, int W, int H, dc *gx, dc *gy) {
  for (int y = 0; y < H; y++)
    for (int x = 0; x < W; x++) {
      long i = (long)y * W + x;
      gx[i] = x < W - 1 ? u[i + 1] - u[i] : 0.0;
      gy[i] = y < H - 1 ? u[i + W] - u[i] : 0.0;
    }
}
static inline dc bd(const dc *a, long i, long s, int pos, int len) {
  if (len == 1) return 0.0;
  if (pos == 0) return a[i];
  if (pos == len - 1) return -a[i - s];
  return a[i] - a[i - s];
}
static inline dc fd(const dc *a, long i, long s, int pos, int len) {
  return pos < len - 1 ? a[i + s] - a[i] : 0.0;
}
/* Divergence2D: d-x p0 + d-y p1 */
static void div2(const dc *p0, const dc *p1, int W, int H, dc *out) {
  for (int y = 0; y < H; y++)
    for (int x = 0; x < W; x++) {
      long i = (long)y * W + x;
      out[i] = bd(p0, i, 1, x, W) + bd(p1, i, W, y, H);
    }
}
/* SymmetricGradient2D: (d-x v0, d-y v1, (d-y v0 + d-x v1) / 2) */
static void symgrad2(const dc *v0, const dc *v1, int W, int H, dc *e0, dc *e1, dc *e2) {
  for (int y = 0; y < H; y++)
    for (int x = 0; x < W; x++) {
      long i = (long)y * W + x;
      e0[i] = bd(v0, i, 1, x, W);
      e1[i] = bd(v1, i, W, y, H);
      e2[i] = 0.5 * (bd(v0, i, W, y, H) + bd(v1, i, 1, x, W));
    }
}
/* SymmetricDivergence2D: (d+x q0 + d+y q2, d+x q2 + d+y q1) */
static void symdiv2(const dc *q0, const dc *q1, const dc *q2, int W, int H, dc *o0, dc *o1) {
  for (int y = 0; y < H; y++)
    for (int x = 0; x < W; x++) {
      long i = (long)y * W + x;
      o0[i] = fd(q0, i, 1, x, W) + fd(q2, i, W, y, H);
      o1[i] = fd(q2, i, 1, x, W) + fd(q1, i, W, y, H);
    }
}
/* -(bdiff_x(diff_x u) + bdiff_y(diff_y u)) with periodic differences: the
 * positive semi-definite periodic Laplacian D^T D. */
static void laplace_periodic(const dc *u, int W, int H, dc *out) {
  for (int y = 0; y < H; y++)
    for (int x = 0; x < W; x++) {
      long i = (long)y * W + x;
      int xp = (x + 1) % W, xm = (x + W - 1) % W, yp = (y + 1) % H, ym = (y + H - 1) % H;
      /* g(i) = u(i+1) - u(i); b(i) = g(i) - g(i-1) = u(i+1) - 2u(i) + u(i-1) */
      dc bx = u[(long)y * W + xp] - 2.0 * u[i] + u[(long)y * W + xm];
      dc by = u[(long)yp * W + x] - 2.0 * u[i] + u[(long)ym * W + x];
      out[i] = -(bx + by);
    }
}

void orc_laplace_periodic(const double *u, int W, int H, double *out) {
  laplace_periodic((const dc *)u, W, H, (dc *)out);
}

/* ------------------------------------------------------------ time average */
/* CartesianCoilConstruction::TimeAveragedReconstruction: crec_c = F(sum_t
 * d[t][c] / max(1, sum_t M_t)) (the unitary image transform of K^H, no mask),
 * u = sqrt(sum_c |crec_c|^2), times the phase of sum_c crec_c when
 * apply_phase. */
void orc_coil_tar(const double *kdatar, const double *mask, int W, int H, int C, int T,
                  int centered, int apply_phase, double *ur, double *crecr) {
  const dc *kd = (const dc *)kdatar;
  dc *u = (dc *)ur, *crec = (dc *)crecr;
  long HW = (long)W * H;
  double *msum = malloc(sizeof(double) * HW);
  for (long p = 0; p < HW; p++) {
    double s = 0.0;
    for (int t = 0; t < T; t++) s += mask ? mask[t * HW + p] : 1.0;
    msum[p] = s > 1.0 ? s : 1.0;
  }
  dc *angle = calloc(HW, sizeof(dc));
  for (long p = 0; p < HW; p++) u[p] = 0.0;
  for (int c = 0; c < C; c++) {
    dc *img = crec + c * HW;
    for (long p = 0; p < HW; p++) {
      dc s = 0.0;
      for (int t = 0; t < T; t++) s += kd[((long)t * C + c) * HW + p];
      img[p] = s / msum[p];
    }
    orc_fft2((double *)img, H, W, -1, centered);
    for (long p = 0; p < HW; p++) {
      img[p] /= sqrt((double)HW);
      angle[p] += img[p];
      u[p] += a2(img[p]);
    }
  }
  for (long p = 0; p < HW; p++) {
    u[p] = sqrt(creal(u[p]));
    if (apply_phase) u[p] *= cexp(I * carg(angle[p]));
  }
  free(msum);
  free(angle);
}

/* ----------------------------------------------------------------- B1FromUH1 */
/* |b1_c| (coil_construction.cpp:363-446): solve (mu |u|^2 + L) x = conj(u)
 * mu |crec_c| per coil by CG from x = 0, b1_c = |x_c| / sqrt(sum_c |x_c|^2).
 * Returns the largest CG iteration count. */
int orc_coil_b1_from_uh1(const double *ur, const double *crecr, int W, int H, int C, double mu,
                         double *b1r) {
  const dc *u = (const dc *)ur, *crec = (const dc *)crecr;
  dc *b1 = (dc *)b1r;
  long N = (long)W * H;
  double *muu = malloc(sizeof(double) * N);
  for (long i = 0; i < N; i++) muu[i] = mu * a2(u[i]);
  dc *x = malloc(sizeof(dc) * N), *r = malloc(sizeof(dc) * N), *p = malloc(sizeof(dc) * N);
  dc *ap = malloc(sizeof(dc) * N), *rhs = malloc(sizeof(dc) * N);
  double *nrm = calloc(N, sizeof(double));
  int maxit_used = 0;
  for (int c = 0; c < C; c++) {
    for (long i = 0; i < N; i++) {
      rhs[i] = conj(u[i]) * (mu * cabs(crec[c * N + i]));
      x[i] = 0.0;
      r[i] = rhs[i];
      p[i] = r[i];
    }
    double rho = 0.0;
    for (long i = 0; i < N; i++) rho += a2(r[i]);
    const double rho0 = rho;
    int it = 0;
    if (rho0 > 1e-12) {
      for (; it < 800; it++) {
        laplace_periodic(p, W, H, ap);
        double pap = 0.0;
        for (long i = 0; i < N; i++) {
          ap[i] += muu[i] * p[i];
          pap += creal(conj(p[i]) * ap[i]);
        }
        const double alpha = rho / pap;
        double rn = 0.0;
        for (long i = 0; i < N; i++) {
          x[i] += alpha * p[i];
          r[i] -= alpha * ap[i];
          rn += a2(r[i]);
        }
        if (rn <= 1e-12 * rho0 || rn <= 1e-12) {
          it++;
          break;
        }
        const double beta = rn / rho;
        for (long i = 0; i < N; i++) p[i] = r[i] + beta * p[i];
        rho = rn;
      }
    }
    if (it > maxit_used) maxit_used = it;
    for (long i = 0; i < N; i++) {
      b1[c * N + i] = x[i];
      nrm[i] += a2(x[i]);
    }
  }
  for (long i = 0; i < N; i++) nrm[i] = sqrt(nrm[i]);
  for (int c = 0; c < C; c++)
    for (long i = 0; i < N; i++) b1[c * N + i] = cabs(b1[c * N + i] / nrm[i]);
  free(muu); free(x); free(r); free(p); free(ap); free(rhs); free(nrm);
  return maxit_used;
}

/* -------------------------------------------------------------- coil order */
/* FindMaximumSum (coil_construction.cpp:52-73): the coil with the largest
 * sum of |crec_c| (first of equals). */
int orc_coil_find_max_sum(const double *crecr, int W, int H, int C) {
  const dc *crec = (const dc *)crecr;
  long N = (long)W * H;
  int best = -1;
  double bs = 0.0;
  for (int c = 0; c < C; c++) {
    double s = 0.0;
    for (long i = 0; i < N; i++) s += cabs(crec[c * N + i]);
    if (best < 0 || s > bs) {
      best = c;
      bs = s;
    }
  }
  return best;
}

/* --------------------------------------------------------------- UTGV2Recon */
/* coil_construction.cpp:75-216: PD iterations of
 *   min_u nu TGV2(u) + 1/2 sum_{c in inds} |b1_c u - crec_c|^2
 * in 2D from x = x_inout (duals zero), fixed steps sigma = tau.
 * p: [nu, alpha0, alpha1, sigma, tau, iterations]. */
void orc_coil_utgv2(const double *b1r, const double *crecr, const int *inds, int n_inds, int W,
                    int H, const double *p, double *xr) {
  const dc *b1 = (const dc *)b1r, *crec = (const dc *)crecr;
  dc *x1 = (dc *)xr;
  const double nu = p[0], a0 = p[1], a1 = p[2], sig = p[3], tau = p[4];
  const int iters = (int)p[5];
  long N = (long)W * H;
  dc *lw = calloc(N, sizeof(dc)), *ib = malloc(sizeof(dc) * N);
  double *sb = calloc(N, sizeof(double));
  for (int k = 0; k < n_inds; k++) {
    const dc *bc = b1 + (long)inds[k] * N, *cc = crec + (long)inds[k] * N;
    for (long i = 0; i < N; i++) {
      lw[i] += conj(bc[i]) * cc[i];
      sb[i] += a2(bc[i]);
    }
  }
  for (long i = 0; i < N; i++) ib[i] = 1.0 / (1.0 + tau * sb[i]);
  dc *ext1 = malloc(sizeof(dc) * N), *x2 = calloc(2 * N, sizeof(dc)), *ext2 = calloc(2 * N, sizeof(dc));
  dc *y1 = calloc(2 * N, sizeof(dc)), *y2 = calloc(3 * N, sizeof(dc));
  dc *t0 = malloc(sizeof(dc) * 3 * N), *dv = malloc(sizeof(dc) * N), *sd = malloc(sizeof(dc) * 2 * N);
  memcpy(ext1, x1, sizeof(dc) * N);
  for (int it = 0; it < iters; it++) {
    grad2(ext1, W, H, t0, t0 + N);
    for (long i = 0; i < 2 * N; i++) y1[i] += sig * (t0[i] - ext2[i]);
    symgrad2(ext2, ext2 + N, W, H, t0, t0 + N, t0 + 2 * N);
    for (long i = 0; i < 3 * N; i++) y2[i] += sig * t0[i];
    /* ProximalMap2D(y1, 1/(alpha1 nu)), ProximalMap2DSym(y2, 1/(alpha0 nu)) */
    for (long i = 0; i < N; i++) {
      double n1 = sqrt(a2(y1[i]) + a2(y1[N + i])) / (a1 * nu);
      double s1 = n1 > 1.0 ? n1 : 1.0;
      y1[i] /= s1;
      y1[N + i] /= s1;
      double n2 = sqrt(a2(y2[i]) + a2(y2[N + i]) + 2.0 * a2(y2[2 * N + i])) / (a0 * nu);
      double s2 = n2 > 1.0 ? n2 : 1.0;
      y2[i] /= s2;
      y2[N + i] /= s2;
      y2[2 * N + i] /= s2;
    }
    div2(y1, y1 + N, W, H, dv);
    symdiv2(y2, y2 + N, y2 + 2 * N, W, H, sd, sd + N);
    for (long i = 0; i < N; i++) {
      dc xn = (x1[i] + tau * dv[i] + tau * lw[i]) * ib[i];
      ext1[i] = 2.0 * xn - x1[i];
      x1[i] = xn;
    }
    for (long i = 0; i < 2 * N; i++) {
      dc xn = x2[i] + tau * (y1[i] + sd[i]);
      ext2[i] = 2.0 * xn - x2[i];
      x2[i] = xn;
    }
  }
  free(lw); free(ib); free(sb); free(ext1); free(x2); free(ext2); free(y1); free(y2);
  free(t0); free(dv); free(sd);
}

/* ----------------------------------------------------------------- B1Recon */
/* coil_construction.cpp:217-362: per coil, PD iterations of
 *   min_b mu/2 |grad b|^2 + 1/2 |b u_n - crec_c|^2,  u_n = u0 / max |u0|,
 * from b = 0 (duals zero), then b normalised to unit sum of squares over the
 * `coils` maps.  p: [sigma, tau]. */
void orc_coil_b1_recon(const double *u0r, const double *crecr, int coils, int W, int H,
                       int iters, double mu, const double *p, double *xr) {
  const dc *u0 = (const dc *)u0r, *crec = (const dc *)crecr;
  dc *x1 = (dc *)xr;
  const double sig = p[0], tau = p[1];
  long N = (long)W * H;
  double mx = 0.0;
  for (long i = 0; i < N; i++) mx = cabs(u0[i]) > mx ? cabs(u0[i]) : mx;
  dc *un = malloc(sizeof(dc) * N);
  for (long i = 0; i < N; i++) un[i] = u0[i] / mx;
  for (long i = 0; i < (long)coils * N; i++) x1[i] = 0.0;
  dc *ext1 = calloc((long)coils * N, sizeof(dc)), *y1 = calloc(2L * coils * N, sizeof(dc));
  dc *t0 = malloc(sizeof(dc) * 2 * N), *dv = malloc(sizeof(dc) * N);
  for (int it = 0; it < iters; it++) {
    for (int c = 0; c < coils; c++) {
      grad2(ext1 + c * N, W, H, t0, t0 + N);
      dc *yx = y1 + (2L * c) * N, *yy = yx + N;
      for (long i = 0; i < N; i++) {
        yx[i] = (yx[i] + sig * t0[i]) / (1.0 + sig / mu);
        yy[i] = (yy[i] + sig * t0[N + i]) / (1.0 + sig / mu);
      }
    }
    for (int c = 0; c < coils; c++) {
      dc *yx = y1 + (2L * c) * N, *yy = yx + N;
      div2(yx, yy, W, H, dv);
      for (long i = 0; i < N; i++) {
        const long k = c * N + i;
        dc xn = (x1[k] + tau * dv[i] + tau * conj(un[i]) * crec[k]) / (1.0 + tau * a2(un[i]));
        ext1[k] = 2.0 * xn - x1[k];
        x1[k] = xn;
      }
    }
  }
  for (long i = 0; i < N; i++) {
    double s = 0.0;
    for (int c = 0; c < coils; c++) s += a2(x1[c * N + i]);
    s = sqrt(s);
    for (int c = 0; c < coils; c++) x1[c * N + i] /= s;
  }
  free(un); free(ext1); free(y1); free(t0); free(dv);
}

/* ------------------------------------------------ PerformCoilConstruction */
/* coil_construction.cpp:447-557 from the time-averaged images.
 * params: [uH1mu, uReg, uNrIt, uSigma, uTau, uAlpha0, uAlpha1, b1Reg,
 *          b1NrIt, b1Sigma, b1Tau, b1FinalReg, b1FinalNrIt].
 * u_out, b1_out: [C][H][W]; order_out (optional): the coil order. */
void orc_coil_construction_from(const double *crecr, const double *u0r, int W, int H, int C,
                                const double *prm, double *ur, double *b1r, int *order_out) {
  const dc *crec = (const dc *)crecr;
  dc *u = (dc *)ur, *b1 = (dc *)b1r;
  long N = (long)W * H;
  dc *u0 = malloc(sizeof(dc) * N);
  for (long i = 0; i < N; i++) u0[i] = cabs(((const dc *)u0r)[i]);
  dc *absb1 = malloc(sizeof(dc) * C * N);
  orc_coil_b1_from_uh1((const double *)u0, crecr, W, H, C, prm[0], (double *)absb1);
  for (long i = 0; i < (long)C * N; i++) {
    b1[i] = 0.0;
    u[i] = 0.0;
  }
  int *inds = malloc(sizeof(int) * C);
  int n = 0;
  inds[n++] = orc_coil_find_max_sum(crecr, W, H, C);
  memcpy(b1 + inds[0] * N, absb1 + inds[0] * N, sizeof(dc) * N);
  const double up[6] = {prm[1], prm[5], prm[6], prm[3], prm[4], prm[2]};
  dc *umax = calloc(N, sizeof(dc));
  orc_coil_utgv2(b1r, crecr, inds, n, W, H, up, (double *)umax);
  memcpy(u, umax, sizeof(dc) * N);
  double *w = calloc(N, sizeof(double));
  dc *bm = malloc(sizeof(dc) * N), *cm = malloc(sizeof(dc) * N), *x1 = malloc(sizeof(dc) * N);
  const double bp[2] = {prm[9], prm[10]};
  for (int cnt = 1; cnt < C; cnt++) {
    const int prev = inds[cnt - 1];
    for (long i = 0; i < N; i++) w[i] = sqrt(w[i] * w[i] + a2(absb1[prev * N + i]));
    int best = -1;
    double bs = 0.0;
    for (int c = 0; c < C; c++) {
      int used = 0;
      for (int k = 0; k < n; k++) used |= inds[k] == c;
      if (used) continue;
      double s = 0.0;
      for (long i = 0; i < N; i++) s += cabs(crec[prev * N + i]) * cabs(crec[c * N + i]);
      if (best < 0 || s > bs) {
        best = c;
        bs = s;
      }
    }
    inds[n++] = best;
    for (long i = 0; i < N; i++) {
      bm[i] = absb1[best * N + i] * u[(cnt - 1) * N + i] * w[i];
      cm[i] = crec[best * N + i] * w[i];
    }
    orc_coil_b1_recon((const double *)bm, (const double *)cm, 1, W, H, (int)prm[8], prm[7], bp,
                      (double *)x1);
    for (long i = 0; i < N; i++) b1[best * N + i] = absb1[best * N + i] * cexp(I * carg(x1[i]));
    memcpy(umax, u + (cnt - 1) * N, sizeof(dc) * N);
    orc_coil_utgv2(b1r, crecr, inds, n, W, H, up, (double *)umax);
    memcpy(u + cnt * N, umax, sizeof(dc) * N);
  }
  /* final correction over all coils from the last image */
  orc_coil_b1_recon((const double *)umax, crecr, C, W, H, (int)prm[12], prm[11], bp, b1r);
  if (order_out)
    for (int k = 0; k < C; k++) order_out[k] = inds[k];
  free(u0); free(absb1); free(inds); free(umax); free(w); free(bm); free(cm); free(x1);
}

void orc_coil_construction(const double *kdata, const double *mask, int W, int H, int C, int T,
                           int centered, const double *prm, double *u_out, double *b1_out,
                           int *order_out) {
  long N = (long)W * H;
  double *u0 = malloc(sizeof(double) * 2 * N), *crec = malloc(sizeof(double) * 2 * C * N);
  orc_coil_tar(kdata, mask, W, H, C, T, centered, 1, u0, crec);
  orc_coil_construction_from(crec, u0, W, H, C, prm, u_out, b1_out, order_out);
  free(u0);
  free(crec);
}
