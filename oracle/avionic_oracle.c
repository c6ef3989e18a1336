/*
 * FP64 CPU oracle for the AVIONIC primal-dual reconstruction loop.
 * TEST INFRASTRUCTURE ONLY (see avionic_oracle.h): the checker and the timed
 * CPU baseline, never part of the product path.
 *
 * Citations are reference paths relative to the AVIONIC source tree.
 */
#include "avionic_oracle.h"

#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef double complex dc;

#define IDX(x, y, t) ((long)(x) + (long)W * ((long)(y) + (long)H * (long)(t)))

/* ======================================================================== */
/* FFT: recursive mixed-radix decimation in time, written for this oracle.   */
/* Independent of the GPU's Stockham formulation so that both are checked    */
/* against each other and against a naive DFT.                               */
/* ======================================================================== */

typedef struct {
  int n;
  int nf;
  int f[40];
  dc *tw; /* tw[k] = exp(sign * 2 pi i k / n) */
} fftplan;

static void plan_init(fftplan *pl, int n, int sign) {
  pl->n = n;
  pl->nf = 0;
  int m = n;
  while (m % 4 == 0) { pl->f[pl->nf++] = 4; m /= 4; }
  while (m % 2 == 0) { pl->f[pl->nf++] = 2; m /= 2; }
  for (int p = 3; m > 1; p += 2) {
    while (m % p == 0) { pl->f[pl->nf++] = p; m /= p; }
    if (p * p > m && m > 1) { pl->f[pl->nf++] = m; m = 1; }
  }
  pl->tw = (dc *)malloc(sizeof(dc) * (size_t)(n > 0 ? n : 1));
  for (int k = 0; k < n; k++) {
    double a = 2.0 * M_PI * (double)k / (double)n;
    pl->tw[k] = cos(a) + (double)sign * sin(a) * I;
  }
}

static void plan_free(fftplan *pl) { free(pl->tw); }

/* out[0..n) = DFT of in[0, s, 2s, ...]; factors f[fi..] multiply to n. */
static void fft_rec(const dc *in, dc *out, int n, long s, const fftplan *pl,
                    int fi, dc *tmp) {
  if (n == 1) { out[0] = in[0]; return; }
  int p = pl->f[fi];
  int m = n / p;
  for (int q = 0; q < p; q++)
    fft_rec(in + q * s, out + q * m, m, s * p, pl, fi + 1, tmp);
  long tws = pl->n / n; /* twiddle stride: W_n^j = tw[j*tws] */
  long pws = pl->n / p; /* W_p^j = tw[j*pws] */
  for (int k = 0; k < m; k++) {
    for (int q = 0; q < p; q++)
      tmp[q] = out[q * m + k] * pl->tw[((long)q * k * tws) % pl->n];
    for (int r = 0; r < p; r++) {
      dc acc = 0;
      for (int q = 0; q < p; q++) acc += tmp[q] * pl->tw[((long)q * r % p) * pws];
      out[k + r * m] = acc;
    }
  }
}

static void fft_plan_exec(const fftplan *pl, dc *a, dc *work, dc *tmp) {
  memcpy(work, a, sizeof(dc) * (size_t)pl->n);
  fft_rec(work, a, pl->n, 1, pl, 0, tmp);
}

void orc_fft1(double *a, int n, int sign) {
  fftplan pl;
  plan_init(&pl, n, sign);
  dc *work = (dc *)malloc(sizeof(dc) * (size_t)n);
  dc *tmp = (dc *)malloc(sizeof(dc) * (size_t)n);
  fft_plan_exec(&pl, (dc *)a, work, tmp);
  free(work);
  free(tmp);
  plan_free(&pl);
}

/* 2D transform of one H x W image (x fastest).  centered: ifftshift before,
 * fftshift after, in both dimensions (the CenteredForward/CenteredInverse
 * variants used at cartesian_operator.cpp:90-99,142-151). */
static void fft2_one(dc *img, int H, int W, const fftplan *pw, const fftplan *ph,
                     int centered, dc *line, dc *work, dc *tmp, dc *shift) {
  if (centered) {
    for (int y = 0; y < H; y++)
      for (int x = 0; x < W; x++)
        shift[y * W + x] = img[((y + H / 2) % H) * W + (x + W / 2) % W];
    memcpy(img, shift, sizeof(dc) * (size_t)H * W);
  }
  for (int y = 0; y < H; y++) fft_plan_exec(pw, img + (long)y * W, work, tmp);
  for (int x = 0; x < W; x++) {
    for (int y = 0; y < H; y++) line[y] = img[(long)y * W + x];
    fft_plan_exec(ph, line, work, tmp);
    for (int y = 0; y < H; y++) img[(long)y * W + x] = line[y];
  }
  if (centered) {
    int cy = (H + 1) / 2, cx = (W + 1) / 2;
    for (int y = 0; y < H; y++)
      for (int x = 0; x < W; x++)
        shift[y * W + x] = img[((y + cy) % H) * W + (x + cx) % W];
    memcpy(img, shift, sizeof(dc) * (size_t)H * W);
  }
}

void orc_fft2(double *a, int H, int W, int sign, int centered) {
  fftplan pw, ph;
  plan_init(&pw, W, sign);
  plan_init(&ph, H, sign);
  int L = H > W ? H : W;
  dc *line = malloc(sizeof(dc) * L), *work = malloc(sizeof(dc) * L);
  dc *shift = malloc(sizeof(dc) * (size_t)H * W), *tmp = malloc(sizeof(dc) * L);
  fft2_one((dc *)a, H, W, &pw, &ph, centered, line, work, tmp, shift);
  free(line); free(work); free(shift); free(tmp);
  plan_free(&pw); plan_free(&ph);
}

/* ======================================================================== */
/* Cartesian operator pair                                                   */
/* ======================================================================== */

/* K: image -> k-space.  (K u)_{t,c} = M_t . F^{-1}_ortho(b1_c . u_t), inverse
 * DFT (+i exponent) with 1/sqrt(HW).  matlab/backward_opt.m:10-14;
 * CUDA/src/cartesian_operator.cpp:122-160 (BackwardOperation). */
void orc_cart_K(const double *xr, const double *b1r, const double *mask, int W,
                int H, int C, int T, int centered, double *zr) {
  const dc *x = (const dc *)xr, *b1 = (const dc *)b1r;
  dc *z = (dc *)zr;
  long HW = (long)H * W;
  double sc = 1.0 / sqrt((double)HW);
  fftplan pw, ph;
  plan_init(&pw, W, +1);
  plan_init(&ph, H, +1);
#pragma omp parallel
  {
    int L = H > W ? H : W;
    dc *line = malloc(sizeof(dc) * L), *work = malloc(sizeof(dc) * L);
    dc *shift = malloc(sizeof(dc) * HW), *tmp = malloc(sizeof(dc) * L);
#pragma omp for schedule(static) collapse(2)
    for (int t = 0; t < T; t++)
      for (int c = 0; c < C; c++) {
        dc *o = z + ((long)t * C + c) * HW;
        for (long i = 0; i < HW; i++) o[i] = b1[c * HW + i] * x[t * HW + i];
        fft2_one(o, H, W, &pw, &ph, centered, line, work, tmp, shift);
        for (long i = 0; i < HW; i++) {
          o[i] *= sc;
          if (mask) o[i] *= mask[t * HW + i];
        }
      }
    free(line); free(work); free(shift); free(tmp);
  }
  plan_free(&pw); plan_free(&ph);
}

/* K^H: k-space -> image.  (K^H z)_t = sum_c conj(b1_c) . F_ortho(M_t . z_{t,c}),
 * forward DFT (-i exponent) with 1/sqrt(HW).  matlab/forward_opt.m:9-13;
 * CUDA/src/cartesian_operator.cpp:65-111 (ForwardOperation).  Unlike the
 * reference (which masks its input in place, cartesian_operator.cpp:83-88),
 * the input is left unmodified. */
void orc_cart_KH(const double *zr, const double *b1r, const double *mask, int W,
                 int H, int C, int T, int centered, double *xr) {
  const dc *z = (const dc *)zr, *b1 = (const dc *)b1r;
  dc *x = (dc *)xr;
  long HW = (long)H * W;
  double sc = 1.0 / sqrt((double)HW);
  fftplan pw, ph;
  plan_init(&pw, W, -1);
  plan_init(&ph, H, -1);
#pragma omp parallel
  {
    int L = H > W ? H : W;
    dc *line = malloc(sizeof(dc) * L), *work = malloc(sizeof(dc) * L);
    dc *shift = malloc(sizeof(dc) * HW), *img = malloc(sizeof(dc) * HW);
    dc *tmp = malloc(sizeof(dc) * L);
#pragma omp for schedule(static)
    for (int t = 0; t < T; t++) {
      dc *o = x + (long)t * HW;
      for (long i = 0; i < HW; i++) o[i] = 0;
      for (int c = 0; c < C; c++) {
        const dc *zi = z + ((long)t * C + c) * HW;
        for (long i = 0; i < HW; i++)
          img[i] = mask ? zi[i] * mask[t * HW + i] : zi[i];
        fft2_one(img, H, W, &pw, &ph, centered, line, work, tmp, shift);
        for (long i = 0; i < HW; i++) o[i] += conj(b1[c * HW + i]) * img[i] * sc;
      }
    }
    free(line); free(work); free(shift); free(img); free(tmp);
  }
  plan_free(&pw); plan_free(&ph);
}

/* ======================================================================== */
/* Finite differences (matlab/get_sp_fdif.m, get_sp_bdif.m)                  */
/* ======================================================================== */

/* delta+ along an axis: u[i+1]-u[i], 0 at the last index (get_sp_fdif.m:8-22,
 * AGILE diff3 as used at utils.cpp:104-117). */
static inline dc fdiff(const dc *u, long i, long stride, int pos, int len) {
  return (pos < len - 1) ? u[i + stride] - u[i] : 0;
}
/* delta- = -(delta+)^T: u0 at the first index, u[i]-u[i-1] inside, -u[n-2] at
 * the last (get_sp_bdif.m:8-24, sym_bgrad_3_3.m:17-27). */
static inline dc bdiff(const dc *u, long i, long stride, int pos, int len) {
  if (len == 1) return 0;
  if (pos == 0) return u[i];
  if (pos == len - 1) return -u[i - stride];
  return u[i] - u[i - stride];
}

/* grad u = (dx+ u/dx, dy+ u/dy, dt+ u/dt).  utils.cpp:99-118; sp_grad_3_1.m. */
void orc_grad(const double *ur, int W, int H, int T, double dx, double dy,
              double dt, double *gr) {
  const dc *u = (const dc *)ur;
  dc *g = (dc *)gr;
  long N = (long)W * H * T, sx = 1, sy = W, st = (long)W * H;
#pragma omp parallel for schedule(static)
  for (int t = 0; t < T; t++)
    for (int y = 0; y < H; y++)
      for (int x = 0; x < W; x++) {
        long i = IDX(x, y, t);
        g[i] = fdiff(u, i, sx, x, W) / dx;
        g[N + i] = fdiff(u, i, sy, y, H) / dy;
        g[2 * N + i] = fdiff(u, i, st, t, T) / dt;
      }
}

/* Symmetrised gradient E(v), order (xx, yy, tt, xy, xt, yt) with backward
 * differences.  utils.cpp:397-456; sp_sym_grad_3_3.m:6-11.  The yt entry
 * divides its y part by dy (MATLAB's ds); the CUDA code uses dx there
 * (utils.cpp:453), identical because every solver sets dx == dy. */
void orc_symgrad(const double *vr, int W, int H, int T, double dx, double dy,
                 double dt, double *er) {
  const dc *v = (const dc *)vr;
  dc *e = (dc *)er;
  long N = (long)W * H * T, sx = 1, sy = W, st = (long)W * H;
  const dc *v0 = v, *v1 = v + N, *v2 = v + 2 * N;
#pragma omp parallel for schedule(static)
  for (int t = 0; t < T; t++)
    for (int y = 0; y < H; y++)
      for (int x = 0; x < W; x++) {
        long i = IDX(x, y, t);
        e[i] = bdiff(v0, i, sx, x, W) / dx;
        e[N + i] = bdiff(v1, i, sy, y, H) / dy;
        e[2 * N + i] = bdiff(v2, i, st, t, T) / dt;
        e[3 * N + i] = 0.5 * (bdiff(v0, i, sy, y, H) / dy + bdiff(v1, i, sx, x, W) / dx);
        e[4 * N + i] = 0.5 * (bdiff(v0, i, st, t, T) / dt + bdiff(v2, i, sx, x, W) / dx);
        e[5 * N + i] = 0.5 * (bdiff(v1, i, st, t, T) / dt + bdiff(v2, i, sy, y, H) / dy);
      }
}

/* div p = -grad^T p = dx- p0/dx + dy- p1/dy + dt- p2/dt.  utils.cpp:733-757
 * (minus diff3trans); sp_div_3_3.m:6-7. */
void orc_div(const double *pr, int W, int H, int T, double dx, double dy,
             double dt, double *outr) {
  const dc *p = (const dc *)pr;
  dc *o = (dc *)outr;
  long N = (long)W * H * T, sx = 1, sy = W, st = (long)W * H;
#pragma omp parallel for schedule(static)
  for (int t = 0; t < T; t++)
    for (int y = 0; y < H; y++)
      for (int x = 0; x < W; x++) {
        long i = IDX(x, y, t);
        o[i] = bdiff(p, i, sx, x, W) / dx + bdiff(p + N, i, sy, y, H) / dy +
               bdiff(p + 2 * N, i, st, t, T) / dt;
      }
}

/* div_sym q = -E^* q (off-diagonals weighted twice): forward differences.
 * utils.cpp:961-1012 (minus bdiff3trans); sp_div_3_6.m:6-16. */
void orc_symdiv(const double *qr, int W, int H, int T, double dx, double dy,
                double dt, double *outr) {
  const dc *q = (const dc *)qr;
  dc *o = (dc *)outr;
  long N = (long)W * H * T, sx = 1, sy = W, st = (long)W * H;
  const dc *q0 = q, *q1 = q + N, *q2 = q + 2 * N, *q3 = q + 3 * N,
           *q4 = q + 4 * N, *q5 = q + 5 * N;
#pragma omp parallel for schedule(static)
  for (int t = 0; t < T; t++)
    for (int y = 0; y < H; y++)
      for (int x = 0; x < W; x++) {
        long i = IDX(x, y, t);
        o[i] = fdiff(q0, i, sx, x, W) / dx + fdiff(q3, i, sy, y, H) / dy +
               fdiff(q4, i, st, t, T) / dt;
        o[N + i] = fdiff(q3, i, sx, x, W) / dx + fdiff(q1, i, sy, y, H) / dy +
                   fdiff(q5, i, st, t, T) / dt;
        o[2 * N + i] = fdiff(q4, i, sx, x, W) / dx + fdiff(q5, i, sy, y, H) / dy +
                       fdiff(q2, i, st, t, T) / dt;
      }
}

static inline double abs2(dc a) { return creal(a) * creal(a) + cimag(a) * cimag(a); }

/* y <- y / max(1, |y|_3 / radius).  utils.cpp:1341-1345 + 1297-1305;
 * norm_3.m; tv_dmri.m:203-207. */
void orc_prox3(double *yr, long N, double radius) {
  dc *y = (dc *)yr;
#pragma omp parallel for schedule(static)
  for (long i = 0; i < N; i++) {
    double n = sqrt(abs2(y[i]) + abs2(y[N + i]) + abs2(y[2 * N + i]));
    double s = n / radius;
    if (s < 1.0) s = 1.0;
    y[i] /= s; y[N + i] /= s; y[2 * N + i] /= s;
  }
}

/* q <- q / max(1, |q|_6 / radius), |q|_6^2 = diag + 2 off-diagonal.
 * utils.cpp:1347-1351, 610-630; norm_6.m:3-4. */
void orc_prox6(double *qr, long N, double radius) {
  dc *q = (dc *)qr;
#pragma omp parallel for schedule(static)
  for (long i = 0; i < N; i++) {
    double n2 = abs2(q[i]) + abs2(q[N + i]) + abs2(q[2 * N + i]) +
                2.0 * (abs2(q[3 * N + i]) + abs2(q[4 * N + i]) + abs2(q[5 * N + i]));
    double s = sqrt(n2) / radius;
    if (s < 1.0) s = 1.0;
    for (int k = 0; k < 6; k++) q[k * N + i] /= s;
  }
}

/* ======================================================================== */
/* Scalars                                                                    */
/* ======================================================================== */

/* Complete elliptic integral of the second kind E(m) by the arithmetic-
 * geometric mean (Abramowitz & Stegun 17.6), the method of utils.cpp:18-47 /
 * MATLAB ellipke, here iterated to double precision. */
double orc_ellipke(double m) {
  if (m < 0.0 || m > 1.0) return NAN;
  if (m == 1.0) return 1.0;
  double a = 1.0, b = sqrt(1.0 - m), s = m, pw = 1.0;
  for (int i = 0; i < 64; i++) {
    double c = 0.5 * (a - b);
    double an = 0.5 * (a + b);
    b = sqrt(a * b);
    a = an;
    pw *= 2.0;
    double w = pw * c * c;
    s += w;
    if (w < 1e-17) break;
  }
  return (M_PI / (2.0 * a)) * (1.0 - 0.5 * s);
}

/* (ds, dt) from the time/space weight.  pd_recon.cpp:72-98;
 * get_weights_3d.m:12-29.  Returns -1 for a non-positive weight
 * (pd_recon.cpp:94 throws invalid_argument). */
int orc_time_space_weights(double tsw, double *ds, double *dt) {
  double r = 1.0 / tsw, w1, w2;
  if (r > 0.0 && r <= 1.0) {
    w2 = M_PI / (2.0 * orc_ellipke(1.0 - r * r));
    w1 = r * w2;
  } else if (r > 1.0) {
    r = 1.0 / r;
    w1 = M_PI / (2.0 * orc_ellipke(1.0 - r * r));
    w2 = r * w1;
  } else {
    return -1;
  }
  *ds = 1.0 / w1;
  *dt = 1.0 / w2;
  return 0;
}

/* lambda = k * (W H T / ||mask||^2) + d; mask NULL = fully sampled.
 * cartesian_operator.cpp:47-62. */
double orc_adapt_lambda_cart(const double *mask, int W, int H, int T, double k,
                             double d) {
  double subfac = (double)W * H * T;
  if (mask) {
    double s = 0;
    for (long i = 0; i < (long)W * H * T; i++) s += mask[i] * mask[i];
    subfac /= s;
  } else {
    subfac /= (double)W * H * T;
  }
  return subfac * k + d;
}

/* pd_recon.cpp:40-61; steps_*_dmri.m:17-30. */
void orc_adapt_step(double nKx, double nx, double ratio, double *sigma,
                    double *tau) {
  double tmp = nx / nKx, theta = 0.95;
  if ((*sigma) * (*tau) > tmp * tmp) {
    if (theta * theta * (*sigma) * (*tau) < tmp * tmp) {
      *sigma *= theta;
      *tau *= theta;
    } else {
      *sigma = tmp * ratio;
      *tau = tmp / ratio;
    }
  }
}

/* ======================================================================== */
/* Solver building blocks                                                     */
/* ======================================================================== */

int orc_primal_components(int method) {
  return (method == ORC_TV || method == ORC_TVTEMP) ? 1
         : method == ORC_TGV2                      ? 4
         : method == ORC_ICTV                      ? 2
                                                   : 8;
}
int orc_dual_components(int method) {
  return method == ORC_TVTEMP ? 1
         : method == ORC_TV   ? 3
         : method == ORC_TGV2 ? 9
         : method == ORC_ICTV ? 6
                              : 18;
}

/* Temporal forward difference and its adjoint (TVTEMP: utils::Gradient_temp
 * and Divergence_temp, utils.cpp:134-143 and 795-805; the difference is zero
 * at the last frame, as for the t component of the full gradient).
 * tgrad: out_t = (x_{t+1} - x_t) / dt for t < T-1, 0 at t = T-1.
 * tgrad_T: the adjoint, out_t = (y_{t-1} [t > 0] - y_t [t < T-1]) / dt, which
 * is minus the temporal divergence. */
static void tgrad(const dc *x, int W, int H, int T, double dt, dc *out) {
  long F = (long)W * H;
  for (int t = 0; t < T; t++)
    for (long i = 0; i < F; i++)
      out[t * F + i] = t < T - 1 ? (x[(t + 1) * F + i] - x[t * F + i]) / dt : 0.0;
}
static void tgrad_T(const dc *y, int W, int H, int T, double dt, dc *out) {
  long F = (long)W * H;
  for (int t = 0; t < T; t++)
    for (long i = 0; i < F; i++) {
      dc v = 0.0;
      if (t > 0) v += y[(t - 1) * F + i];
      if (t < T - 1) v -= y[t * F + i];
      out[t * F + i] = v / dt;
    }
}

double orc_tvtemp_norm(const double *ur, int W, int H, int T, double dt) {
  long N = (long)W * H * T;
  dc *g = malloc(sizeof(dc) * N);
  tgrad((const dc *)ur, W, H, T, dt, g);
  double s = 0.0;
  for (long i = 0; i < N; i++) s += cabs(g[i]);
  free(g);
  return s;
}

void orc_tgrad(const double *x, int W, int H, int T, double dt, double *out) {
  tgrad((const dc *)x, W, H, T, dt, (dc *)out);
}
void orc_tgrad_T(const double *y, int W, int H, int T, double dt, double *out) {
  tgrad_T((const dc *)y, W, H, T, dt, (dc *)out);
}

/* Regulariser part of K applied to a primal state (P comps) -> D comps.
 * TV:     [grad u]                                     (tv.cpp:140-144)
 * TGV2:   [grad u - v, E v]                            (tgv2.cpp:238-253)
 * ICTGV2: [grad1(u-w) - v, E1 v, grad2 w - v2, E2 v2]  (ictgv2.cpp:338-361)
 * d6 = {dx, dy, dt, dx2, dy2, dt2}. */
static void reg_K(int method, const dc *x, int W, int H, int T, const double *d6,
                  dc *out, dc *tmp) {
  long N = (long)W * H * T;
  if (method == ORC_TVTEMP) {  /* [grad_t u] (tv_temp.cpp:130-132) */
    tgrad(x, W, H, T, d6[2], out);
    return;
  }
  if (method == ORC_ICTV) {  /* [grad1(u - w), grad2 w] (ictv.cpp:266-271) */
    const dc *u = x, *w = x + N;
    for (long i = 0; i < N; i++) tmp[i] = u[i] - w[i];
    orc_grad((const double *)tmp, W, H, T, d6[0], d6[1], d6[2], (double *)out);
    orc_grad((const double *)w, W, H, T, d6[3], d6[4], d6[5], (double *)(out + 3 * N));
    return;
  }
  if (method == ORC_TV) {
    orc_grad((const double *)x, W, H, T, d6[0], d6[1], d6[2], (double *)out);
    return;
  }
  const dc *u = x, *v = x + N;
  if (method == ORC_TGV2) {
    orc_grad((const double *)u, W, H, T, d6[0], d6[1], d6[2], (double *)out);
    for (long i = 0; i < 3 * N; i++) out[i] -= v[i];
    orc_symgrad((const double *)v, W, H, T, d6[0], d6[1], d6[2], (double *)(out + 3 * N));
    return;
  }
  const dc *w = x + 4 * N, *v2 = x + 5 * N;
  for (long i = 0; i < N; i++) tmp[i] = u[i] - w[i];
  orc_grad((const double *)tmp, W, H, T, d6[0], d6[1], d6[2], (double *)out);
  for (long i = 0; i < 3 * N; i++) out[i] -= v[i];
  orc_symgrad((const double *)v, W, H, T, d6[0], d6[1], d6[2], (double *)(out + 3 * N));
  orc_grad((const double *)w, W, H, T, d6[3], d6[4], d6[5], (double *)(out + 9 * N));
  for (long i = 0; i < 3 * N; i++) out[9 * N + i] -= v2[i];
  orc_symgrad((const double *)v2, W, H, T, d6[3], d6[4], d6[5], (double *)(out + 12 * N));
}

/* Adjoint of reg_K (as used in the primal descent), D comps -> P comps:
 * TV:     [-div p]                                       (tv.cpp:156-159)
 * TGV2:   [-div p, -p - divsym q]                        (tgv2.cpp:271-285)
 * ICTGV2: [-div1 y1, -y1 - divsym1 y2, div1 y1 - div2 y3, -y3 - divsym2 y4]
 *                                                        (ictgv2.cpp:389-416) */
static void reg_KT(int method, const dc *y, int W, int H, int T, const double *d6,
                   dc *out) {
  long N = (long)W * H * T;
  if (method == ORC_TVTEMP) {  /* [-div_t y] (tv_temp.cpp:145-147) */
    tgrad_T(y, W, H, T, d6[2], out);
    return;
  }
  if (method == ORC_ICTV) {  /* [-div1 y1, div1 y1 - div2 y3] (ictv.cpp:301-310) */
    orc_div((const double *)y, W, H, T, d6[0], d6[1], d6[2], (double *)out);
    orc_div((const double *)(y + 3 * N), W, H, T, d6[3], d6[4], d6[5], (double *)(out + N));
    for (long i = 0; i < N; i++) {
      out[N + i] = out[i] - out[N + i];
      out[i] = -out[i];
    }
    return;
  }
  orc_div((const double *)y, W, H, T, d6[0], d6[1], d6[2], (double *)out);
  if (method == ORC_TV) {
    for (long i = 0; i < N; i++) out[i] = -out[i];
    return;
  }
  orc_symdiv((const double *)(y + 3 * N), W, H, T, d6[0], d6[1], d6[2], (double *)(out + N));
  for (long i = 0; i < 3 * N; i++) out[N + i] = -y[i] - out[N + i];
  if (method == ORC_TGV2) {
    for (long i = 0; i < N; i++) out[i] = -out[i];
    return;
  }
  orc_div((const double *)(y + 9 * N), W, H, T, d6[3], d6[4], d6[5], (double *)(out + 4 * N));
  for (long i = 0; i < N; i++) {
    out[4 * N + i] = out[i] - out[4 * N + i];
    out[i] = -out[i];
  }
  orc_symdiv((const double *)(y + 12 * N), W, H, T, d6[3], d6[4], d6[5], (double *)(out + 5 * N));
  for (long i = 0; i < 3 * N; i++) out[5 * N + i] = -y[9 * N + i] - out[5 * N + i];
}

/* Weights of the squared dual norm used by the step rule (SumOfSquares3/6,
 * utils.cpp:1398-1418): 2 on the off-diagonal symmetric-gradient entries. */
static double dual_weight(int method, int comp) {
  int c = comp;
  if (method == ORC_TV || method == ORC_TVTEMP || method == ORC_ICTV) return 1.0;
  if (method == ORC_ICTGV2 && c >= 9) c -= 9;
  return (c >= 6) ? 2.0 : 1.0;
}

/* Dual projections with the radii of tgv2.cpp:258-259 and ictgv2.cpp:367-381;
 * TV uses radius 1 (tv.cpp:150). */
static void prox_all(int method, dc *y, long N, double alpha0, double alpha1,
                     double alpha) {
  if (method == ORC_TVTEMP) {  /* ProximalMap1D(y, 1) (tv_temp.cpp:138) */
    for (long i = 0; i < N; i++) {
      double a = cabs(y[i]);
      if (a > 1.0) y[i] /= a;
    }
    return;
  }
  if (method == ORC_TV) { orc_prox3((double *)y, N, 1.0); return; }
  if (method == ORC_ICTV) {  /* ictv.cpp:284-290 */
    double m = alpha < 1.0 - alpha ? alpha : 1.0 - alpha;
    orc_prox3((double *)y, N, alpha1 * (alpha / m));
    orc_prox3((double *)(y + 3 * N), N, alpha1 * ((1.0 - alpha) / m));
    return;
  }
  if (method == ORC_TGV2) {
    orc_prox3((double *)y, N, alpha1);
    orc_prox6((double *)(y + 3 * N), N, alpha0);
    return;
  }
  double m = alpha < 1.0 - alpha ? alpha : 1.0 - alpha;
  orc_prox3((double *)y, N, alpha1 * (alpha / m));
  orc_prox6((double *)(y + 3 * N), N, alpha0 * (alpha / m));
  orc_prox3((double *)(y + 9 * N), N, alpha1 * ((1.0 - alpha) / m));
  orc_prox6((double *)(y + 12 * N), N, alpha0 * ((1.0 - alpha) / m));
}

/* The data operator of a reconstruction: Cartesian or non-Cartesian. */
typedef struct {
  int W, H, C, T;
  long kn;  /* k-space length (complex entries) */
  int centered;
  const double *mask;
  /* non-Cartesian */
  int nc, M, J;
  double osf;
  const double *traj, *dcf;
  int vol;  /* volumetric operator: T is the depth */
} orc_op;

void orc_fft3(double *a, int D, int H, int W, int sign) {
  long P = (long)H * W;
  for (int d = 0; d < D; d++) orc_fft2(a + 2 * d * P, H, W, sign, 0);
  double *col = malloc(sizeof(double) * 2 * D);
  for (long p = 0; p < P; p++) {
    for (int d = 0; d < D; d++) {
      col[2 * d] = a[2 * (d * P + p)];
      col[2 * d + 1] = a[2 * (d * P + p) + 1];
    }
    orc_fft1(col, D, sign);
    for (int d = 0; d < D; d++) {
      a[2 * (d * P + p)] = col[2 * d];
      a[2 * (d * P + p) + 1] = col[2 * d + 1];
    }
  }
  free(col);
}

/* cartesian_operator3d.cpp:216-280 (K) and :64-133 (K^H) */
void orc_cart3_K(const double *xr, const double *b1r, const double *mask, int W, int H, int D,
                 int C, double *zr) {
  long N = (long)W * H * D;
  const dc *x = (const dc *)xr, *b1 = (const dc *)b1r;
  dc *z = (dc *)zr;
  double s = 1.0 / sqrt((double)N);
  for (int c = 0; c < C; c++) {
    dc *zc = z + c * N;
    for (long i = 0; i < N; i++) zc[i] = b1[c * N + i] * x[i];
    orc_fft3((double *)zc, D, H, W, +1);
    for (long i = 0; i < N; i++) zc[i] *= s * (mask ? mask[i] : 1.0);
  }
}

void orc_cart3_KH(const double *zr, const double *b1r, const double *mask, int W, int H, int D,
                  int C, double *xr) {
  long N = (long)W * H * D;
  const dc *z = (const dc *)zr, *b1 = (const dc *)b1r;
  dc *x = (dc *)xr, *t = malloc(sizeof(dc) * N);
  double s = 1.0 / sqrt((double)N);
  for (long i = 0; i < N; i++) x[i] = 0.0;
  for (int c = 0; c < C; c++) {
    for (long i = 0; i < N; i++) t[i] = z[c * N + i] * (mask ? mask[i] : 1.0);
    orc_fft3((double *)t, D, H, W, -1);
    for (long i = 0; i < N; i++) x[i] += conj(b1[c * N + i]) * t[i] * s;
  }
  free(t);
}

static void op_K(const orc_op *op, const double *x, const double *b1, double *z) {
  if (op->vol)
    orc_cart3_K(x, b1, op->mask, op->W, op->H, op->T, op->C, z);
  else if (op->nc)
    orc_nufft_K(x, b1, op->traj, op->dcf, op->W, op->H, op->C, op->T, op->M, op->J, op->osf, z);
  else
    orc_cart_K(x, b1, op->mask, op->W, op->H, op->C, op->T, op->centered, z);
}
static void op_KH(const orc_op *op, const double *z, const double *b1, double *x) {
  if (op->vol)
    orc_cart3_KH(z, b1, op->mask, op->W, op->H, op->T, op->C, x);
  else if (op->nc)
    orc_nufft_KH(z, b1, op->traj, op->dcf, op->W, op->H, op->C, op->T, op->M, op->J, op->osf, x);
  else
    orc_cart_KH(z, b1, op->mask, op->W, op->H, op->C, op->T, op->centered, x);
}

static void adapt_step_op(int method, const dc *delta, const double *b1, const orc_op *op,
                          const double *d6, double ratio, double *sigma, double *tau,
                          double *nKx_out, double *nx_out) {
  int W = op->W, H = op->H, T = op->T;
  long N = (long)W * H * T;
  int P = orc_primal_components(method), D = orc_dual_components(method);
  dc *ky = malloc(sizeof(dc) * D * N), *tmp = malloc(sizeof(dc) * N);
  dc *kz = malloc(sizeof(dc) * (size_t)op->kn);
  reg_K(method, delta, W, H, T, d6, ky, tmp);
  op_K(op, (const double *)delta, b1, (double *)kz);
  double s = 0.0, sx = 0.0;
  /* ICTV's AdaptStepSize (ictv.cpp:96-131) overwrites its regulariser sum
   * before use: its nKx is the data term alone. */
  for (int k = 0; k < D && method != ORC_ICTV; k++) {
    double wk = dual_weight(method, k), acc = 0.0;
    for (long i = 0; i < N; i++) acc += abs2(ky[k * N + i]);
    s += wk * acc;
  }
  for (long i = 0; i < op->kn; i++) s += abs2(kz[i]);
  for (long i = 0; i < (long)P * N; i++) sx += abs2(delta[i]);
  double nKx = sqrt(s), nx = sqrt(sx);
  if (nKx_out) *nKx_out = nKx;
  if (nx_out) *nx_out = nx;
  orc_adapt_step(nKx, nx, ratio, sigma, tau);
  free(ky); free(tmp); free(kz);
}

void orc_solver_adapt_step(int method, const double *deltar, const double *b1,
                           const double *mask, int W, int H, int C, int T,
                           int centered, const double *d6, double ratio,
                           double *sigma, double *tau, double *nKx_out,
                           double *nx_out) {
  orc_op op = {W, H, C, T, (long)C * W * H * T, centered, mask, 0, 0, 0, 0.0, NULL, NULL};
  adapt_step_op(method, (const dc *)deltar, b1, &op, d6, ratio, sigma, tau, nKx_out, nx_out);
}

/* ======================================================================== */
/* Operator norm by power iteration                                           */
/* ======================================================================== */

/* Deterministic start vector: entry j of the real array (real and imaginary
 * parts in turn) is splitmix64(j + seed) mapped to [-1, 1). */
static double hash_unit(unsigned long long j) {
  unsigned long long z = j + 0x9e3779b97f4a7c15ull;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  z ^= z >> 31;
  return (double)(z >> 11) * (2.0 / 9007199254740992.0) - 1.0;
}

/* M x = K_reg^* K_reg x + [K^H K u, 0, ...]: K_full^* K_full for the stacked
 * operator [K_reg; K] with K_reg^* the adjoint of the primal descent (the
 * pairing with off-diagonal weight 2). */
static void opnorm_apply(int method, const dc *x, const double *b1, const orc_op *op,
                         const double *d6, dc *out, dc *ky, dc *tmp, dc *kz, dc *img) {
  long N = (long)op->W * op->H * op->T;
  reg_K(method, x, op->W, op->H, op->T, d6, ky, tmp);
  reg_KT(method, ky, op->W, op->H, op->T, d6, out);
  op_K(op, (const double *)x, b1, (double *)kz);
  op_KH(op, (const double *)kz, b1, (double *)img);
  for (long i = 0; i < N; i++) out[i] += img[i];
}

/* ||K_full|| after `iters` power-iteration steps, following
 * matlab/getopnorm.m:1-16 (x = y / |y|; y = M x; l = <y, x>), returned as
 * sqrt(|l|) (getopnorm returns the Rayleigh quotient l itself). */
static double opnorm_op(int method, const double *b1, const orc_op *op, const double *d6,
                        int iters) {
  long N = (long)op->W * op->H * op->T;
  int P = orc_primal_components(method), D = orc_dual_components(method);
  long n = (long)P * N;
  dc *x = malloc(sizeof(dc) * n), *y = malloc(sizeof(dc) * n);
  dc *ky = malloc(sizeof(dc) * D * N), *tmp = malloc(sizeof(dc) * N);
  dc *kz = malloc(sizeof(dc) * (size_t)op->kn), *img = malloc(sizeof(dc) * N);
  for (long i = 0; i < n; i++)
    x[i] = hash_unit(2 * (unsigned long long)i) + I * hash_unit(2 * (unsigned long long)i + 1);
  opnorm_apply(method, x, b1, op, d6, y, ky, tmp, kz, img);
  double l = 0.0;
  for (int it = 0; it < iters; it++) {
    double n2 = 0.0;
    for (long i = 0; i < n; i++) n2 += abs2(y[i]);
    double inv = n2 > 0.0 ? 1.0 / sqrt(n2) : 1.0;
    for (long i = 0; i < n; i++) x[i] = y[i] * inv;
    opnorm_apply(method, x, b1, op, d6, y, ky, tmp, kz, img);
    l = 0.0;
    for (long i = 0; i < n; i++) l += creal(conj(y[i]) * x[i]);
  }
  free(x); free(y); free(ky); free(tmp); free(kz); free(img);
  return sqrt(fabs(l));
}

double orc_opnorm(int method, const double *b1, const double *mask, int W, int H, int C, int T,
                  int centered, const double *d6, int iters) {
  orc_op op = {W, H, C, T, (long)C * W * H * T, centered, mask, 0, 0, 0, 0.0, NULL, NULL};
  return opnorm_op(method, b1, &op, d6, iters);
}

double orc_opnorm3(const double *b1, const double *mask, int W, int H, int D, int C,
                   const double *d6, int iters) {
  orc_op op = {W, H, C, D, (long)C * W * H * D, 0, mask, 0, 0, 0, 0.0, NULL, NULL, 1};
  return opnorm_op(ORC_TGV2, b1, &op, d6, iters);
}

double orc_opnorm_nc(int method, const double *b1, int W, int H, int C, int T, int M, int J,
                     double osf, const double *traj, const double *dcf, const double *d6,
                     int iters) {
  orc_op op = {W, H, C, T, (long)C * M * T, 0, NULL, 1, M, J, osf, traj, dcf};
  return opnorm_op(method, b1, &op, d6, iters);
}

/* ======================================================================== */
/* Functionals and the primal-dual gap                                        */
/* ======================================================================== */

static double sum_norm3(const dc *g, long N) {
  double s = 0;
  for (long i = 0; i < N; i++)
    s += sqrt(abs2(g[i]) + abs2(g[N + i]) + abs2(g[2 * N + i]));
  return s;
}
static double sum_norm6(const dc *e, long N) {
  double s = 0;
  for (long i = 0; i < N; i++)
    s += sqrt(abs2(e[i]) + abs2(e[N + i]) + abs2(e[2 * N + i]) +
              2.0 * (abs2(e[3 * N + i]) + abs2(e[4 * N + i]) + abs2(e[5 * N + i])));
  return s;
}

/* sum |grad u|_3.  utils.cpp:1119-1132; get_tv.m. */
double orc_tv_norm(const double *u, int W, int H, int T, double dx, double dy,
                   double dt) {
  long N = (long)W * H * T;
  dc *g = malloc(sizeof(dc) * 3 * N);
  orc_grad(u, W, H, T, dx, dy, dt, (double *)g);
  double s = sum_norm3(g, N);
  free(g);
  return s;
}

double orc_ictv_norm(const double *ur, const double *wr, int W, int H, int T, double alpha1,
                     double alpha, const double *d6) {
  long N = (long)W * H * T;
  const dc *u = (const dc *)ur, *w = (const dc *)wr;
  dc *d = malloc(sizeof(dc) * N);
  for (long i = 0; i < N; i++) d[i] = u[i] - w[i];
  double n1 = orc_tv_norm((const double *)d, W, H, T, d6[0], d6[1], d6[2]);
  double n2 = orc_tv_norm(wr, W, H, T, d6[3], d6[4], d6[5]);
  free(d);
  double m = alpha < 1.0 - alpha ? alpha : 1.0 - alpha;
  return alpha1 * (alpha / m) * n1 + alpha1 * ((1.0 - alpha) / m) * n2;
}

/* alpha1 sum|grad u - v|_3 + alpha0 sum|E v|_6.  utils.cpp:1179-1219;
 * get_tgv2.m. */
double orc_tgv2_norm(const double *ur, const double *vr, int W, int H, int T,
                     double alpha0, double alpha1, double dx, double dy,
                     double dt) {
  long N = (long)W * H * T;
  const dc *v = (const dc *)vr;
  dc *g = malloc(sizeof(dc) * 3 * N), *e = malloc(sizeof(dc) * 6 * N);
  orc_grad(ur, W, H, T, dx, dy, dt, (double *)g);
  for (long i = 0; i < 3 * N; i++) g[i] -= v[i];
  orc_symgrad(vr, W, H, T, dx, dy, dt, (double *)e);
  double s = alpha1 * sum_norm3(g, N) + alpha0 * sum_norm6(e, N);
  free(g); free(e);
  return s;
}

/* (alpha/m) TGV2(u - w, v) + ((1-alpha)/m) TGV2(w, v2), m = min(alpha,1-alpha).
 * utils.cpp:1245-1287; get_ictgv2.m. */
double orc_ictgv2_norm(const double *ur, const double *vr, const double *wr,
                       const double *v2r, int W, int H, int T, double alpha0,
                       double alpha1, double alpha, double dx, double dy,
                       double dt, double dx2, double dy2, double dt2) {
  long N = (long)W * H * T;
  const dc *u = (const dc *)ur, *w = (const dc *)wr;
  dc *d = malloc(sizeof(dc) * N);
  for (long i = 0; i < N; i++) d[i] = u[i] - w[i];
  double n1 = orc_tgv2_norm((double *)d, vr, W, H, T, alpha0, alpha1, dx, dy, dt);
  double n2 = orc_tgv2_norm(wr, v2r, W, H, T, alpha0, alpha1, dx2, dy2, dt2);
  free(d);
  double m = alpha < 1.0 - alpha ? alpha : 1.0 - alpha;
  return (alpha / m) * n1 + ((1.0 - alpha) / m) * n2;
}

static double l1(const dc *a, long n) {
  double s = 0;
  for (long i = 0; i < n; i++) s += cabs(a[i]);
  return s;
}

/* G* of tv.cpp:202-225, tgv2.cpp:130-165, ictgv2.cpp:178-232:
 *  g1 = lambda/2 ||K x1 - d||^2, g2 = Re<d,z> + ||z||^2/(2 lambda),
 *  g3 = ||K^H z - div1 y1||_1, and for TGV2/ICTGV2 the l1 norms of the
 *  remaining dual residuals (norm1 of a complex vector = sum |.|). */
double orc_gstar(int method, const double *x1r, const double *yr,
                 const double *zr, const double *datar, const double *b1,
                 const double *mask, int W, int H, int C, int T, int centered,
                 double lambda, const double *d6) {
  long N = (long)W * H * T, CN = (long)C * N;
  const dc *y = (const dc *)yr, *z = (const dc *)zr, *data = (const dc *)datar;
  dc *kz = malloc(sizeof(dc) * CN);
  orc_cart_K(x1r, b1, mask, W, H, C, T, centered, (double *)kz);
  double g1 = 0, g2 = 0, zz = 0;
  for (long i = 0; i < CN; i++) {
    g1 += abs2(kz[i] - data[i]);
    g2 += creal(conj(data[i]) * z[i]);
    zz += abs2(z[i]);
  }
  g1 *= 0.5 * lambda;
  g2 += zz / (2.0 * lambda);
  free(kz);
  dc *img = malloc(sizeof(dc) * N), *dv = malloc(sizeof(dc) * N);
  dc *sd = malloc(sizeof(dc) * 3 * N);
  orc_cart_KH(zr, b1, mask, W, H, C, T, centered, (double *)img);
  if (method == ORC_TVTEMP) {  /* K^H z - div_t y (tv_temp.cpp:205-208) */
    tgrad_T(y, W, H, T, d6[2], dv);
    for (long i = 0; i < N; i++) dv[i] = -dv[i];
  } else {
    orc_div(yr, W, H, T, d6[0], d6[1], d6[2], (double *)dv);
  }
  for (long i = 0; i < N; i++) img[i] -= dv[i];
  double g = g1 + g2 + l1(img, N);
  if (method == ORC_ICTV) {  /* + |div1 y1 - div2 y3|_1 (ictv.cpp:178-183) */
    dc *dv3 = malloc(sizeof(dc) * N);
    orc_div((const double *)(y + 3 * N), W, H, T, d6[3], d6[4], d6[5], (double *)dv3);
    for (long i = 0; i < N; i++) dv3[i] = dv[i] - dv3[i];
    g += l1(dv3, N);
    free(dv3);
  } else if (method != ORC_TV && method != ORC_TVTEMP) {
    orc_symdiv((const double *)(y + 3 * N), W, H, T, d6[0], d6[1], d6[2], (double *)sd);
    for (long i = 0; i < 3 * N; i++) sd[i] += y[i];
    g += l1(sd, N) + l1(sd + N, N) + l1(sd + 2 * N, N);
  }
  if (method == ORC_ICTGV2) {
    dc *dv3 = malloc(sizeof(dc) * N);
    orc_div((const double *)(y + 9 * N), W, H, T, d6[3], d6[4], d6[5], (double *)dv3);
    for (long i = 0; i < N; i++) dv3[i] = dv[i] - dv3[i];
    g += l1(dv3, N);
    free(dv3);
    orc_symdiv((const double *)(y + 12 * N), W, H, T, d6[3], d6[4], d6[5], (double *)sd);
    for (long i = 0; i < 3 * N; i++) sd[i] += y[9 * N + i];
    g += l1(sd, N) + l1(sd + N, N) + l1(sd + 2 * N, N);
  }
  free(img); free(dv); free(sd);
  return g;
}

/* |G* + R(x)|, not yet divided by N.  tv.cpp:227-235, tgv2.cpp:167-176,
 * ictgv2.cpp:234-249. */
double orc_pdgap(int method, const double *xr, const double *yr,
                 const double *zr, const double *data, const double *b1,
                 const double *mask, int W, int H, int C, int T, int centered,
                 double lambda, double alpha0, double alpha1, double alpha,
                 const double *d6) {
  long N = (long)W * H * T;
  const dc *x = (const dc *)xr;
  double gs = orc_gstar(method, xr, yr, zr, data, b1, mask, W, H, C, T,
                        centered, lambda, d6);
  double r;
  if (method == ORC_TVTEMP)
    r = orc_tvtemp_norm(xr, W, H, T, d6[2]);
  else if (method == ORC_ICTV)
    r = orc_ictv_norm(xr, (const double *)(x + N), W, H, T, alpha1, alpha, d6);
  else if (method == ORC_TV)
    r = orc_tv_norm(xr, W, H, T, d6[0], d6[1], d6[2]);
  else if (method == ORC_TGV2)
    r = orc_tgv2_norm(xr, (const double *)(x + N), W, H, T, alpha0, alpha1,
                      d6[0], d6[1], d6[2]);
  else
    r = orc_ictgv2_norm(xr, (const double *)(x + N), (const double *)(x + 4 * N),
                        (const double *)(x + 5 * N), W, H, T, alpha0, alpha1,
                        alpha, d6[0], d6[1], d6[2], d6[3], d6[4], d6[5]);
  return fabs(gs + r);
}

/* ======================================================================== */
/* The primal-dual loop                                                       */
/* tv.cpp:99-200, tgv2.cpp:180-338, ictgv2.cpp:294-492;                       */
/* tv_dmri.m:161-227, tgv2_dmri.m:176-255, ictgv2_dmri.m:187-282.             */
/* ======================================================================== */

static int recon_core(const orc_params *p, const orc_op *op, const double *datar,
                      const double *b1, double *x_inout, double *x_out_all,
                      double *sig_log, double *tau_log, double *gap_log, int *n_gap,
                      int nthreads) {
  const int W = op->W, H = op->H, C = op->C, T = op->T;
  const double *mask = op->mask;
  (void)C;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#else
  (void)nthreads;
#endif
  int method = p->method == ORC_TGV2_3D ? ORC_TGV2 : p->method;  /* TGV2 on a volume */
  long N = (long)W * H * T, CN = op->kn;
  int P = orc_primal_components(method), D = orc_dual_components(method);
  double d6[6] = {1, 1, 1, 1, 1, 1};
  if (method == ORC_TVTEMP) {  /* dt as given: no time/space weights (tv_temp.cpp:98) */
    if (!(p->dt > 0.0)) return -1;
    d6[2] = p->dt;
  } else if (p->method == ORC_TGV2_3D) {  /* dx, dy, dz as given */
    if (!(p->dx > 0.0 && p->dy > 0.0 && p->dz > 0.0)) return -1;
    d6[0] = p->dx;
    d6[1] = p->dy;
    d6[2] = p->dz;
  } else {
    if (orc_time_space_weights(p->tsw, &d6[0], &d6[2]) != 0) return -1;
    d6[1] = d6[0];
  }
  if (method == ORC_ICTGV2 || method == ORC_ICTV) {
    if (orc_time_space_weights(p->tsw2, &d6[3], &d6[5]) != 0) return -1;
    d6[4] = d6[3];
  }
  const dc *data = (const dc *)datar;
  dc *x = calloc((size_t)P * N, sizeof(dc)), *ext = calloc((size_t)P * N, sizeof(dc));
  dc *xn = calloc((size_t)P * N, sizeof(dc));
  dc *y = calloc((size_t)D * N, sizeof(dc)), *ky = calloc((size_t)D * N, sizeof(dc));
  dc *z = calloc((size_t)CN, sizeof(dc)), *kz = calloc((size_t)CN, sizeof(dc));
  dc *img = calloc((size_t)N, sizeof(dc)), *tmp = calloc((size_t)N, sizeof(dc));
  dc *delta = calloc((size_t)P * N, sizeof(dc));
  memcpy(x, x_inout, sizeof(dc) * N);
  memcpy(ext, x_inout, sizeof(dc) * N);
  double sigma = p->sigma, tau = p->tau, lambda = p->lambda;
  int ng = 0, k;
  for (k = 0; k < p->maxIt; k++) {
    /* dual ascent: y += sigma K_reg ext ; z += sigma K ext_u */
    reg_K(method, ext, W, H, T, d6, ky, tmp);
    for (long i = 0; i < (long)D * N; i++) y[i] += sigma * ky[i];
    op_K(op, (double *)ext, b1, (double *)kz);
    for (long i = 0; i < CN; i++) z[i] += sigma * kz[i];
    /* proximal maps */
    prox_all(method, y, N, p->alpha0, p->alpha1, p->alpha);
    double zs = 1.0 / (1.0 + sigma / lambda);
    for (long i = 0; i < CN; i++) z[i] = (z[i] - sigma * data[i]) * zs;
    /* primal descent: x+ = x - tau (K_reg^T y + [K^H z, 0...]) */
    reg_KT(method, y, W, H, T, d6, xn);
    op_KH(op, (double *)z, b1, (double *)img);
    for (long i = 0; i < N; i++) xn[i] += img[i];
    for (long i = 0; i < (long)P * N; i++) xn[i] = x[i] - tau * xn[i];
    /* extragradient: ext = 2 x+ - x ; x = x+ */
    for (long i = 0; i < (long)P * N; i++) {
      ext[i] = 2.0 * xn[i] - x[i];
      x[i] = xn[i];
    }
    int adapt = (p->schedule == 1) || (p->schedule == 0 && ((k < 10) || (k % 50 == 0)));
    if (adapt) {
      for (long i = 0; i < (long)P * N; i++) delta[i] = ext[i] - x[i];
      adapt_step_op(method, delta, b1, op, d6, p->sigmaTauRatio, &sigma, &tau, NULL, NULL);
    }
    if (sig_log) sig_log[k] = sigma;
    if (tau_log) tau_log[k] = tau;
    if (p->gapEvery > 0 && k % p->gapEvery == 0 && gap_log && !op->nc && !op->vol) {
      double g = orc_pdgap(method, (double *)x, (double *)y, (double *)z, datar,
                           b1, mask, W, H, C, T, p->centered, lambda, p->alpha0,
                           p->alpha1, p->alpha, d6);
      gap_log[ng++] = g / (double)N;
    }
  }
  memcpy(x_inout, x, sizeof(dc) * N);
  if (x_out_all) memcpy(x_out_all, x, sizeof(dc) * P * N);
  if (n_gap) *n_gap = ng;
  free(x); free(ext); free(xn); free(y); free(ky); free(z); free(kz);
  free(img); free(tmp); free(delta);
  return k;
}

int orc_recon(const orc_params *p, int W, int H, int C, int T,
              const double *datar, const double *b1, const double *mask,
              double *x_inout, double *x_out_all, double *sig_log,
              double *tau_log, double *gap_log, int *n_gap, int nthreads) {
  orc_op op = {W, H, C, T, (long)C * W * H * T, p->centered, mask, 0, 0, 0, 0.0, NULL, NULL};
  return recon_core(p, &op, datar, b1, x_inout, x_out_all, sig_log, tau_log, gap_log, n_gap,
                    nthreads);
}

int orc_recon3(const orc_params *p, int W, int H, int D, int C, const double *data,
               const double *b1, const double *mask, double *x_inout, double *x_out_all,
               double *sig_log, double *tau_log, int nthreads) {
  orc_op op = {W, H, C, D, (long)C * W * H * D, 0, mask, 0, 0, 0, 0.0, NULL, NULL, 1};
  return recon_core(p, &op, data, b1, x_inout, x_out_all, sig_log, tau_log, NULL, NULL,
                    nthreads);
}

int orc_recon_nc(const orc_params *p, int W, int H, int C, int T, int M, int J,
                 double osf, const double *traj, const double *dcf,
                 const double *datar, const double *b1, double *x_inout,
                 double *x_out_all, double *sig_log, double *tau_log,
                 int nthreads) {
  orc_op op = {W, H, C, T, (long)C * M * T, 0, NULL, 1, M, J, osf, traj, dcf};
  return recon_core(p, &op, datar, b1, x_inout, x_out_all, sig_log, tau_log, NULL, NULL,
                    nthreads);
}
