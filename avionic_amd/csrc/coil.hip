// Coil-sensitivity estimation on the GPU: CoilConstruction
// (CUDA/src/coil_construction.cpp:22-557) with the Cartesian
// TimeAveragedReconstruction (cartesian_coil_construction.cpp:23-87).
//
//   1. time-averaged coil images crec_c = F(sum_t d_{t,c} / max(1, sum_t M_t))
//      and u0 = sqrt(sum_c |crec_c|^2) (phase of sum_c crec_c);
//   2. |b1_c| from (mu |u0|^2 + L) x_c = mu |u0| |crec_c| (L the periodic 2D
//      Laplacian, cg_forward_operation.h:46-67), all coils in one batched
//      conjugate-gradient solve with per-coil scalars on the device, then
//      |x_c| / sqrt(sum_c |x_c|^2) (B1FromUH1 :363-446);
//   3. a greedy coil order (FindMaximumSum :52-73, then the coil overlapping
//      most with the previous one); per coil a TGV2 image estimate from the
//      coils so far (UTGV2Recon :75-216) and the phase of the next map from a
//      quadratically regularised fit (B1Recon :217-362);
//   4. a final B1Recon over all coils (:553-556).
// Images are [coils][height][width] complex float; arithmetic is float (the
// reference's precision), reductions are double in a fixed order.  Each PD
// iteration is two fused kernels (dual ascent + projection, primal descent +
// extragradient) over all pixels (and coils).
#include <algorithm>
#include <cmath>
#include <memory>
#include <vector>

#include "coil.hpp"

namespace avionic {

namespace {

constexpr int kTC = 256;

inline unsigned nblocks(long n) { return (unsigned)std::max<long>(1, std::min<long>(4096, (n + kTC - 1) / kTC)); }

// Non-periodic 2D differences (utils::Gradient2D / Divergence2D etc.).
__device__ inline float2 fdx(const float2 *a, int i, int x, int W) { return x < W - 1 ? csub(a[i + 1], a[i]) : czero(); }
__device__ inline float2 fdy(const float2 *a, int i, int y, int W, int H) {
  return y < H - 1 ? csub(a[i + W], a[i]) : czero();
}
__device__ inline float2 bdx(const float2 *a, int i, int x, int W) {
  if (W == 1) return czero();
  if (x == 0) return a[i];
  if (x == W - 1) return cscale(a[i - 1], -1.f);
  return csub(a[i], a[i - 1]);
}
__device__ inline float2 bdy(const float2 *a, int i, int y, int W, int H) {
  if (H == 1) return czero();
  if (y == 0) return a[i];
  if (y == H - 1) return cscale(a[i - W], -1.f);
  return csub(a[i], a[i - W]);
}

// ------------------------------------------------------------ time average
// davg[c] = sum_t d[t][c] / max(1, sum_t M_t), mask from the compact rows.
__global__ void k_tar_average(const CartGeom g, const float2 *__restrict__ kd,
                              float2 *__restrict__ avg) {
  const long HW = (long)g.H * g.W, n = HW * g.C;
  for (long e = (long)blockIdx.x * kTC + threadIdx.x; e < n; e += (long)gridDim.x * kTC) {
    const long c = e / HW, p = e % HW;
    const int y = (int)(p / g.W), x = (int)(p % g.W);
    float ms = 0.f;
    float2 s = czero();
    for (int t = 0; t < g.T; ++t) {
      const int k = g.rowmap[t * g.H + y];
      ms += k >= 0 ? g.mrows[((long)g.row_off[t] + k) * g.W + x] : 0.f;
      s = cadd(s, kd[((long)t * g.C + c) * HW + p]);
    }
    avg[e] = cscale(s, 1.f / fmaxf(1.f, ms));
  }
}
// u = sqrt(sum_c |crec_c|^2) (times the phase of sum_c crec_c)
__global__ void k_tar_combine(const float2 *__restrict__ crec, float2 *__restrict__ u, int C,
                              long N, int apply_phase) {
  for (long i = (long)blockIdx.x * kTC + threadIdx.x; i < N; i += (long)gridDim.x * kTC) {
    float2 s = czero();
    float m = 0.f;
    for (int c = 0; c < C; ++c) {
      const float2 v = crec[c * N + i];
      s = cadd(s, v);
      m += cabs2(v);
    }
    const float a = sqrtf(m);
    if (apply_phase) {
      const float ph = atan2f(s.y, s.x);
      u[i] = make_float2(a * cosf(ph), a * sinf(ph));
    } else {
      u[i] = make_float2(a, 0.f);
    }
  }
}

// --------------------------------------------------------- reductions
// block-wide sum of v, stored by thread 0 at *out
__device__ inline void block_sum_store(double v, double *out) {
  __shared__ double red[kTC];
  red[threadIdx.x] = v;
  __syncthreads();
  for (int s = kTC / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = red[0];
}
// per coil: sum |a| (b == null) or sum |a| |b| (b: one image)
__global__ void k_coil_abs_sum(const float2 *__restrict__ a, const float2 *__restrict__ b, long N,
                               double *__restrict__ partial) {
  const int c = blockIdx.y, nb = gridDim.x;
  double acc = 0.0;
  for (long i = (long)blockIdx.x * kTC + threadIdx.x; i < N; i += (long)nb * kTC) {
    const float2 v = a[c * N + i];
    double t = sqrt((double)cabs2(v));
    if (b) t *= sqrt((double)cabs2(b[i]));
    acc += t;
  }
  block_sum_store(acc, partial + (long)c * nb + blockIdx.x);
}
// max |a| over one image (block partials)
__global__ void k_abs_max(const float2 *__restrict__ a, long N, double *__restrict__ partial) {
  __shared__ double red[kTC];
  double m = 0.0;
  for (long i = (long)blockIdx.x * kTC + threadIdx.x; i < N; i += (long)gridDim.x * kTC)
    m = fmax(m, sqrt((double)cabs2(a[i])));
  red[threadIdx.x] = m;
  __syncthreads();
  for (int s = kTC / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// ------------------------------------------------------------- B1FromUH1
// Batched CG over the coils: A = diag(mu |u|^2) + L, L the periodic 2D
// Laplacian (D^T D with wrap-around forward differences).
struct CgScal {
  double rho, rho0, alpha;
  int active;
};
__global__ void k_cg_init(const float2 *__restrict__ u, const float2 *__restrict__ crec, float mu,
                          float2 *__restrict__ x, float2 *__restrict__ r, float2 *__restrict__ p,
                          long N, int C) {
  const long n = N * C;
  for (long e = (long)blockIdx.x * kTC + threadIdx.x; e < n; e += (long)gridDim.x * kTC) {
    const long i = e % N;
    const float a = mu * sqrtf(cabs2(crec[e]));
    const float2 uu = u[i];
    const float2 rhs = make_float2(uu.x * a, -uu.y * a);  // conj(u) mu |crec|
    x[e] = czero();
    r[e] = rhs;
    p[e] = rhs;
  }
}
__device__ inline float2 lap_periodic(const float2 *a, int x, int y, int W, int H) {
  const int i = y * W + x;
  const int xp = x + 1 < W ? x + 1 : 0, xm = x > 0 ? x - 1 : W - 1;
  const int yp = y + 1 < H ? y + 1 : 0, ym = y > 0 ? y - 1 : H - 1;
  const float2 c = a[i];
  const float2 bx = cadd(csub(a[y * W + xp], cscale(c, 2.f)), a[y * W + xm]);
  const float2 by = cadd(csub(a[yp * W + x], cscale(c, 2.f)), a[ym * W + x]);
  return cscale(cadd(bx, by), -1.f);
}
// Ap = (mu |u|^2 + L) p for active coils; partial[c][b] = sum Re(conj(p) Ap)
__global__ void k_cg_apply(const float2 *__restrict__ u, float mu, const float2 *__restrict__ p,
                           float2 *__restrict__ ap, const CgScal *__restrict__ sc, int W, int H,
                           double *__restrict__ partial) {
  const int c = blockIdx.y, nb = gridDim.x;
  const long N = (long)W * H;
  double acc = 0.0;
  if (sc[c].active) {
    const float2 *pc = p + c * N;
    for (long i = (long)blockIdx.x * kTC + threadIdx.x; i < N; i += (long)nb * kTC) {
      const int x = (int)(i % W), y = (int)(i / W);
      const float2 v = cadd(lap_periodic(pc, x, y, W, H), cscale(pc[i], mu * cabs2(u[i])));
      ap[c * N + i] = v;
      acc += (double)pc[i].x * v.x + (double)pc[i].y * v.y;
    }
  }
  block_sum_store(acc, partial + (long)c * nb + blockIdx.x);
}
// alpha = rho / <p, Ap>
__global__ void k_cg_alpha(CgScal *sc, const double *__restrict__ partial, int nb) {
  const int c = blockIdx.x;
  if (threadIdx.x != 0 || !sc[c].active) return;
  double s = 0.0;
  for (int b = 0; b < nb; ++b) s += partial[(long)c * nb + b];
  sc[c].alpha = sc[c].rho / s;
}
// x += alpha p; r -= alpha Ap; partial = sum |r|^2
__global__ void k_cg_update(float2 *__restrict__ x, float2 *__restrict__ r,
                            const float2 *__restrict__ p, const float2 *__restrict__ ap,
                            const CgScal *__restrict__ sc, long N, double *__restrict__ partial) {
  const int c = blockIdx.y, nb = gridDim.x;
  double acc = 0.0;
  if (sc[c].active) {
    const float al = (float)sc[c].alpha;
    for (long i = (long)blockIdx.x * kTC + threadIdx.x; i < N; i += (long)nb * kTC) {
      const long e = c * N + i;
      x[e] = cadd(x[e], cscale(p[e], al));
      const float2 rn = csub(r[e], cscale(ap[e], al));
      r[e] = rn;
      acc += (double)cabs2(rn);
    }
  }
  block_sum_store(acc, partial + (long)c * nb + blockIdx.x);
}
// convergence (squared residual <= 1e-12 rho0 or <= 1e-12) and beta
__global__ void k_cg_beta(CgScal *sc, const double *__restrict__ partial, int nb, double *beta) {
  const int c = blockIdx.x;
  if (threadIdx.x != 0) return;
  beta[c] = 0.0;
  if (!sc[c].active) return;
  double rn = 0.0;
  for (int b = 0; b < nb; ++b) rn += partial[(long)c * nb + b];
  if (rn <= 1e-12 * sc[c].rho0 || rn <= 1e-12) {
    sc[c].active = 0;
    return;
  }
  beta[c] = rn / sc[c].rho;
  sc[c].rho = rn;
}
__global__ void k_cg_direction(float2 *__restrict__ p, const float2 *__restrict__ r,
                               const CgScal *__restrict__ sc, const double *__restrict__ beta,
                               long N) {
  const int c = blockIdx.y;
  if (!sc[c].active) return;
  const float be = (float)beta[c];
  for (long i = (long)blockIdx.x * kTC + threadIdx.x; i < N; i += (long)gridDim.x * kTC) {
    const long e = c * N + i;
    p[e] = cadd(r[e], cscale(p[e], be));
  }
}
// b1_c = |x_c| / sqrt(sum_c |x_c|^2)
__global__ void k_cg_finish(const float2 *__restrict__ x, float2 *__restrict__ b1, int C, long N) {
  for (long i = (long)blockIdx.x * kTC + threadIdx.x; i < N; i += (long)gridDim.x * kTC) {
    float s = 0.f;
    for (int c = 0; c < C; ++c) s += cabs2(x[c * N + i]);
    const float inv = 1.f / sqrtf(s);
    for (int c = 0; c < C; ++c) b1[c * N + i] = make_float2(sqrtf(cabs2(x[c * N + i])) * inv, 0.f);
  }
}

// ------------------------------------------------------------- UTGV2Recon
// Lw = sum_{c in inds} conj(b1_c) crec_c; ib = 1 / (1 + tau sum |b1_c|^2)
__global__ void k_ut_pre(const float2 *__restrict__ b1, const float2 *__restrict__ crec,
                         const int *__restrict__ inds, int n_inds, float tau, long N,
                         float2 *__restrict__ lw, float *__restrict__ ib) {
  for (long i = (long)blockIdx.x * kTC + threadIdx.x; i < N; i += (long)gridDim.x * kTC) {
    float2 l = czero();
    float s = 0.f;
    for (int k = 0; k < n_inds; ++k) {
      const long o = (long)inds[k] * N + i;
      l = cadd(l, cmulc(b1[o], crec[o]));
      s += cabs2(b1[o]);
    }
    lw[i] = l;
    ib[i] = 1.f / (1.f + tau * s);
  }
}
struct UtState {
  float2 *x1, *ext1, *x2, *ext2, *y1, *y2;  // x2 / ext2 / y1: 2 N, y2: 3 N
};
// y1 += sigma (grad ext1 - ext2); y2 += sigma E ext2; projections (radii
// alpha1 nu and alpha0 nu; |y2| with the off-diagonal weight 2)
__global__ void k_ut_dual(UtState s, int W, int H, float sigma, float r1, float r2) {
  const long N = (long)W * H;
  for (long e = (long)blockIdx.x * kTC + threadIdx.x; e < N; e += (long)gridDim.x * kTC) {
    const int i = (int)e, x = i % W, y = i / W;
    const float2 *v0 = s.ext2, *v1 = s.ext2 + N;
    float2 p0 = cadd(s.y1[i], cscale(csub(fdx(s.ext1, i, x, W), v0[i]), sigma));
    float2 p1 = cadd(s.y1[N + i], cscale(csub(fdy(s.ext1, i, y, W, H), v1[i]), sigma));
    const float2 e0 = bdx(v0, i, x, W), e1 = bdy(v1, i, y, W, H);
    const float2 e2 = cscale(cadd(bdy(v0, i, y, W, H), bdx(v1, i, x, W)), 0.5f);
    float2 q0 = cadd(s.y2[i], cscale(e0, sigma));
    float2 q1 = cadd(s.y2[N + i], cscale(e1, sigma));
    float2 q2 = cadd(s.y2[2 * N + i], cscale(e2, sigma));
    const float s1 = fmaxf(1.f, sqrtf(cabs2(p0) + cabs2(p1)) / r1);
    const float s2 = fmaxf(1.f, sqrtf(cabs2(q0) + cabs2(q1) + 2.f * cabs2(q2)) / r2);
    s.y1[i] = cscale(p0, 1.f / s1);
    s.y1[N + i] = cscale(p1, 1.f / s1);
    s.y2[i] = cscale(q0, 1.f / s2);
    s.y2[N + i] = cscale(q1, 1.f / s2);
    s.y2[2 * N + i] = cscale(q2, 1.f / s2);
  }
}
// x1+ = (x1 + tau div y1 + tau Lw) ib; x2+ = x2 + tau (y1 + divsym y2);
// ext = 2 x+ - x; x = x+
__global__ void k_ut_primal(UtState s, const float2 *__restrict__ lw, const float *__restrict__ ib,
                            int W, int H, float tau) {
  const long N = (long)W * H;
  for (long e = (long)blockIdx.x * kTC + threadIdx.x; e < N; e += (long)gridDim.x * kTC) {
    const int i = (int)e, x = i % W, y = i / W;
    const float2 *y10 = s.y1, *y11 = s.y1 + N, *q0 = s.y2, *q1 = s.y2 + N, *q2 = s.y2 + 2 * N;
    const float2 dv = cadd(bdx(y10, i, x, W), bdy(y11, i, y, W, H));
    const float2 sd0 = cadd(fdx(q0, i, x, W), fdy(q2, i, y, W, H));
    const float2 sd1 = cadd(fdx(q2, i, x, W), fdy(q1, i, y, W, H));
    const float2 xo = s.x1[i];
    const float2 xn = cscale(cadd(cadd(xo, cscale(dv, tau)), cscale(lw[i], tau)), ib[i]);
    s.ext1[i] = csub(cscale(xn, 2.f), xo);
    s.x1[i] = xn;
    const float2 a0 = s.x2[i], a1 = s.x2[N + i];
    const float2 n0 = cadd(a0, cscale(cadd(y10[i], sd0), tau));
    const float2 n1 = cadd(a1, cscale(cadd(y11[i], sd1), tau));
    s.ext2[i] = csub(cscale(n0, 2.f), a0);
    s.ext2[N + i] = csub(cscale(n1, 2.f), a1);
    s.x2[i] = n0;
    s.x2[N + i] = n1;
  }
}

// --------------------------------------------------------------- B1Recon
// un = u0 / max|u0|: lw_c = conj(un) crec_c, iu = 1 / (1 + tau |un|^2)
__global__ void k_b1_pre(const float2 *__restrict__ u0, const float2 *__restrict__ crec,
                         const double *__restrict__ mx, float tau, int coils, long N,
                         float2 *__restrict__ lw, float *__restrict__ iu) {
  const float inv = (float)(1.0 / *mx);
  for (long i = (long)blockIdx.x * kTC + threadIdx.x; i < N; i += (long)gridDim.x * kTC) {
    const float2 un = cscale(u0[i], inv);
    iu[i] = 1.f / (1.f + tau * cabs2(un));
    for (int c = 0; c < coils; ++c) lw[c * N + i] = cmulc(un, crec[c * N + i]);
  }
}
// y_c = (y_c + sigma grad ext_c) / (1 + sigma / mu)
__global__ void k_b1_dual(const float2 *__restrict__ ext, float2 *__restrict__ y, int W, int H,
                          int coils, float sigma, float shrink) {
  const long N = (long)W * H, n = N * coils;
  for (long e = (long)blockIdx.x * kTC + threadIdx.x; e < n; e += (long)gridDim.x * kTC) {
    const long c = e / N;
    const int i = (int)(e % N), x = i % W, y0 = i / W;
    const float2 *ec = ext + c * N;
    float2 *yc = y + 2 * c * N;
    yc[i] = cscale(cadd(yc[i], cscale(fdx(ec, i, x, W), sigma)), shrink);
    yc[N + i] = cscale(cadd(yc[N + i], cscale(fdy(ec, i, y0, W, H), sigma)), shrink);
  }
}
// x+ = (x + tau div y_c + tau lw_c) iu; ext = 2 x+ - x; x = x+
__global__ void k_b1_primal(float2 *__restrict__ x1, float2 *__restrict__ ext,
                            const float2 *__restrict__ y, const float2 *__restrict__ lw,
                            const float *__restrict__ iu, int W, int H, int coils, float tau) {
  const long N = (long)W * H, n = N * coils;
  for (long e = (long)blockIdx.x * kTC + threadIdx.x; e < n; e += (long)gridDim.x * kTC) {
    const long c = e / N;
    const int i = (int)(e % N), x = i % W, y0 = i / W;
    const float2 *yc = y + 2 * c * N;
    const float2 dv = cadd(bdx(yc, i, x, W), bdy(yc + N, i, y0, W, H));
    const float2 xo = x1[e];
    const float2 xn = cscale(cadd(cadd(xo, cscale(dv, tau)), cscale(lw[e], tau)), iu[i]);
    ext[e] = csub(cscale(xn, 2.f), xo);
    x1[e] = xn;
  }
}
// x_c /= sqrt(sum_c |x_c|^2)
__global__ void k_b1_normalize(float2 *__restrict__ x, int coils, long N) {
  for (long i = (long)blockIdx.x * kTC + threadIdx.x; i < N; i += (long)gridDim.x * kTC) {
    float s = 0.f;
    for (int c = 0; c < coils; ++c) s += cabs2(x[c * N + i]);
    const float inv = 1.f / sqrtf(s);
    for (int c = 0; c < coils; ++c) x[c * N + i] = cscale(x[c * N + i], inv);
  }
}

// ------------------------------------------------------ outer loop helpers
// w = sqrt(w^2 + |b|^2)
__global__ void k_weight_update(float *__restrict__ w, const float2 *__restrict__ b, long N) {
  for (long i = (long)blockIdx.x * kTC + threadIdx.x; i < N; i += (long)gridDim.x * kTC)
    w[i] = sqrtf(w[i] * w[i] + cabs2(b[i]));
}
// bm = |b1_m| u_prev w; cm = crec_m w
__global__ void k_weighted_inputs(const float2 *__restrict__ absb, const float2 *__restrict__ u,
                                  const float2 *__restrict__ crec, const float *__restrict__ w,
                                  float2 *__restrict__ bm, float2 *__restrict__ cm, long N) {
  for (long i = (long)blockIdx.x * kTC + threadIdx.x; i < N; i += (long)gridDim.x * kTC) {
    bm[i] = cscale(cmul(absb[i], u[i]), w[i]);
    cm[i] = cscale(crec[i], w[i]);
  }
}
// b1_m = |b1_m| exp(i arg x)
__global__ void k_apply_phase(const float2 *__restrict__ absb, const float2 *__restrict__ x,
                              float2 *__restrict__ out, long N) {
  for (long i = (long)blockIdx.x * kTC + threadIdx.x; i < N; i += (long)gridDim.x * kTC) {
    const float ph = atan2f(x[i].y, x[i].x);
    out[i] = cmul(absb[i], make_float2(cosf(ph), sinf(ph)));
  }
}

double host_sum(const double *d, long n, hipStream_t s, long stride = 1, long off = 0) {
  std::vector<double> h((size_t)n * stride);
  AV_HIP(hipMemcpyAsync(h.data(), d, sizeof(double) * h.size(), hipMemcpyDeviceToHost, s));
  AV_HIP(hipStreamSynchronize(s));
  double a = 0.0;
  for (long k = 0; k < n; ++k) a += h[(size_t)(k * stride + off)];
  return a;
}

}  // namespace

// ======================================================================= host
// CoilConstruction::InitParams (coil_construction.cpp:22-46)
void coil_default_params(avionic_coil_params *p) {
  *p = avionic_coil_params{};
  p->uH1mu = 1e-5;
  p->uReg = 0.2;
  p->uNrIt = 100;
  p->uSigma = p->uTau = std::sqrt(1.0 / 12.0);
  p->uAlpha0 = std::sqrt(2.0);
  p->uAlpha1 = 1.0;
  p->b1Reg = 2.0;
  p->b1NrIt = 500;
  p->b1Sigma = p->b1Tau = std::sqrt(1.0 / 8.0);
  p->b1FinalReg = 0.1;
  p->b1FinalNrIt = 1000;
}

// sum over each coil of |crec_c| (or |crec_c| |ref|), fixed-order
std::vector<double> coil_abs_sums(const CoilCtx &k, const float2 *crec, const float2 *ref,
                                  int coils) {
  const int nb = (int)std::min<long>(256, (k.N + kTC - 1) / kTC);
  DevBuf part;
  part.ensure(sizeof(double) * (size_t)nb * coils);
  hipLaunchKernelGGL(k_coil_abs_sum, dim3(nb, coils), dim3(kTC), 0, k.s, crec, ref, k.N,
                     part.as<double>());
  AV_LAUNCH_CHECK();
  std::vector<double> h((size_t)nb * coils), out(coils, 0.0);
  AV_HIP(hipMemcpyAsync(h.data(), part.p, sizeof(double) * h.size(), hipMemcpyDeviceToHost, k.s));
  AV_HIP(hipStreamSynchronize(k.s));
  for (int c = 0; c < coils; ++c)
    for (int b = 0; b < nb; ++b) out[c] += h[(size_t)c * nb + b];
  return out;
}

int coil_find_max_sum(const CoilCtx &k, const float2 *crec) {
  const std::vector<double> sums = coil_abs_sums(k, crec, nullptr, k.C);
  int best = -1;
  for (int c = 0; c < k.C; ++c)
    if (best < 0 || sums[c] > sums[best]) best = c;
  return best;
}

void coil_time_averaged(const Operator &op, const float2 *kdata, float2 *u, float2 *crec,
                        bool apply_phase, hipStream_t s) {
  const long HW = (long)op.W * op.H;
  if (const auto *c = dynamic_cast<const CartesianOperator *>(&op)) {
    // cartesian_coil_construction.cpp:23-87: per coil the sampled average
    // over frames, then the image transform of the operator
    const CartGeom &g = c->plan().g;
    hipLaunchKernelGGL(k_tar_average, dim3(nblocks(HW * g.C)), dim3(kTC), 0, s, g, kdata, crec);
    AV_LAUNCH_CHECK();
    launch_fft2(crec, g.pw, g.ph, g.H, g.W, g.C, -1, g.centered, g.scale, s);
  } else {
    // noncartesian_coil_construction.cpp:21-123: per coil, every frame's
    // points gridded together into one image (kernel width 3, oversampling 2)
    std::unique_ptr<Operator> all(make_noncartesian_all_frames(op));
    if (!all)
      throw Error(AVIONIC_E_UNSUPPORTED, "time-averaged coil images need a 2D+t operator");
    const long M = op.dense_elems() / ((long)op.T * op.C);  // points per frame
    DevBuf z, ones;
    z.ensure(sizeof(float2) * M * op.T);
    ones.ensure(sizeof(float2) * HW);
    const std::vector<float2> one((size_t)HW, make_float2(1.f, 0.f));
    AV_HIP(hipMemcpyAsync(ones.p, one.data(), sizeof(float2) * HW, hipMemcpyHostToDevice, s));
    for (int c = 0; c < op.C; ++c) {
      // [frames][coils][M] -> coil c's [frames * M]
      AV_HIP(hipMemcpy2DAsync(z.p, sizeof(float2) * M, kdata + (long)c * M,
                              sizeof(float2) * M * op.C, sizeof(float2) * M, op.T,
                              hipMemcpyDeviceToDevice, s));
      all->forward(z.as<float2>(), crec + (long)c * HW, ones.as<float2>(), s);
    }
    AV_HIP(hipStreamSynchronize(s));  // `one` and the operator end with this call
  }
  hipLaunchKernelGGL(k_tar_combine, dim3(nblocks(HW)), dim3(kTC), 0, s, crec, u, op.C, HW,
                     apply_phase ? 1 : 0);
  AV_LAUNCH_CHECK();
}

void coil_b1_from_uh1(const CoilCtx &k, const float2 *u, const float2 *crec, double mu,
                      float2 *b1) {
  const long n = k.N * k.C;
  DevBuf x, r, p, ap, sc, part, beta;
  x.ensure(sizeof(float2) * n);
  r.ensure(sizeof(float2) * n);
  p.ensure(sizeof(float2) * n);
  ap.ensure(sizeof(float2) * n);
  sc.ensure(sizeof(CgScal) * k.C);
  beta.ensure(sizeof(double) * k.C);
  const int nb = (int)std::min<long>(256, (k.N + kTC - 1) / kTC);
  part.ensure(sizeof(double) * (size_t)nb * k.C);
  hipLaunchKernelGGL(k_cg_init, dim3(nblocks(n)), dim3(kTC), 0, k.s, u, crec, (float)mu,
                     x.as<float2>(), r.as<float2>(), p.as<float2>(), k.N, k.C);
  AV_LAUNCH_CHECK();
  // rho0 = |r|^2 per coil
  {
    std::vector<CgScal> hs(k.C);
    const float2 *rr = r.as<float2>();
    std::vector<float2> hr((size_t)n);
    AV_HIP(hipMemcpyAsync(hr.data(), rr, sizeof(float2) * n, hipMemcpyDeviceToHost, k.s));
    AV_HIP(hipStreamSynchronize(k.s));
    for (int c = 0; c < k.C; ++c) {
      double a = 0.0;
      for (long i = 0; i < k.N; ++i) {
        const float2 v = hr[(size_t)(c * k.N + i)];
        a += (double)v.x * v.x + (double)v.y * v.y;
      }
      hs[c].rho = hs[c].rho0 = a;
      hs[c].alpha = 0.0;
      hs[c].active = a > 1e-12 ? 1 : 0;
    }
    AV_HIP(hipMemcpyAsync(sc.p, hs.data(), sizeof(CgScal) * k.C, hipMemcpyHostToDevice, k.s));
  }
  const dim3 grid(nb, k.C);
  for (int it = 0; it < 800; ++it) {
    hipLaunchKernelGGL(k_cg_apply, grid, dim3(kTC), 0, k.s, u, (float)mu, p.as<float2>(),
                       ap.as<float2>(), sc.as<CgScal>(), k.W, k.H, part.as<double>());
    hipLaunchKernelGGL(k_cg_alpha, dim3(k.C), dim3(64), 0, k.s, sc.as<CgScal>(), part.as<double>(),
                       nb);
    hipLaunchKernelGGL(k_cg_update, grid, dim3(kTC), 0, k.s, x.as<float2>(), r.as<float2>(),
                       p.as<float2>(), ap.as<float2>(), sc.as<CgScal>(), k.N, part.as<double>());
    hipLaunchKernelGGL(k_cg_beta, dim3(k.C), dim3(64), 0, k.s, sc.as<CgScal>(), part.as<double>(),
                       nb, beta.as<double>());
    hipLaunchKernelGGL(k_cg_direction, grid, dim3(kTC), 0, k.s, p.as<float2>(), r.as<float2>(),
                       sc.as<CgScal>(), beta.as<double>(), k.N);
    AV_LAUNCH_CHECK();
    if ((it & 15) == 15) {  // stop once every coil has converged
      std::vector<CgScal> hs(k.C);
      AV_HIP(hipMemcpyAsync(hs.data(), sc.p, sizeof(CgScal) * k.C, hipMemcpyDeviceToHost, k.s));
      AV_HIP(hipStreamSynchronize(k.s));
      bool any = false;
      for (const CgScal &c : hs) any = any || c.active;
      if (!any) break;
    }
  }
  hipLaunchKernelGGL(k_cg_finish, dim3(nblocks(k.N)), dim3(kTC), 0, k.s, x.as<float2>(), b1, k.C,
                     k.N);
  AV_LAUNCH_CHECK();
}

void coil_utgv2(const CoilCtx &k, const float2 *b1, const float2 *crec, const std::vector<int> &inds,
                const avionic_coil_params &p, float2 *x_inout) {
  const long N = k.N;
  DevBuf bi, lw, ib, st;
  bi.ensure(sizeof(int) * inds.size());
  AV_HIP(hipMemcpyAsync(bi.p, inds.data(), sizeof(int) * inds.size(), hipMemcpyHostToDevice, k.s));
  lw.ensure(sizeof(float2) * N);
  ib.ensure(sizeof(float) * N);
  st.ensure(sizeof(float2) * N * 10);  // ext1, x2 (2), ext2 (2), y1 (2), y2 (3)
  AV_HIP(hipMemsetAsync(st.p, 0, st.bytes, k.s));
  UtState s;
  s.x1 = x_inout;
  s.ext1 = st.as<float2>();
  s.x2 = s.ext1 + N;
  s.ext2 = s.x2 + 2 * N;
  s.y1 = s.ext2 + 2 * N;
  s.y2 = s.y1 + 2 * N;
  AV_HIP(hipMemcpyAsync(s.ext1, x_inout, sizeof(float2) * N, hipMemcpyDeviceToDevice, k.s));
  const float tau = (float)p.uTau, sigma = (float)p.uSigma;
  hipLaunchKernelGGL(k_ut_pre, dim3(nblocks(N)), dim3(kTC), 0, k.s, b1, crec, bi.as<int>(),
                     (int)inds.size(), tau, N, lw.as<float2>(), ib.as<float>());
  AV_LAUNCH_CHECK();
  const float r1 = (float)(p.uAlpha1 * p.uReg), r2 = (float)(p.uAlpha0 * p.uReg);
  for (unsigned it = 0; it < p.uNrIt; ++it) {
    hipLaunchKernelGGL(k_ut_dual, dim3(nblocks(N)), dim3(kTC), 0, k.s, s, k.W, k.H, sigma, r1, r2);
    hipLaunchKernelGGL(k_ut_primal, dim3(nblocks(N)), dim3(kTC), 0, k.s, s, lw.as<float2>(),
                       ib.as<float>(), k.W, k.H, tau);
  }
  AV_LAUNCH_CHECK();
  AV_HIP(hipStreamSynchronize(k.s));  // the index table stays alive until here
}

void coil_b1_recon(const CoilCtx &k, const float2 *u0, const float2 *crec, int coils,
                   unsigned maxit, double mu, const avionic_coil_params &p, float2 *x1) {
  const long N = k.N, n = N * coils;
  DevBuf ext, y, lw, iu, part;
  ext.ensure(sizeof(float2) * n);
  y.ensure(sizeof(float2) * 2 * n);
  lw.ensure(sizeof(float2) * n);
  iu.ensure(sizeof(float) * N);
  const int nb = (int)std::min<long>(256, (N + kTC - 1) / kTC);
  part.ensure(sizeof(double) * (nb + 1));
  AV_HIP(hipMemsetAsync(x1, 0, sizeof(float2) * n, k.s));
  AV_HIP(hipMemsetAsync(ext.p, 0, ext.bytes, k.s));
  AV_HIP(hipMemsetAsync(y.p, 0, y.bytes, k.s));
  hipLaunchKernelGGL(k_abs_max, dim3(nb), dim3(kTC), 0, k.s, u0, N, part.as<double>());
  AV_LAUNCH_CHECK();
  std::vector<double> hm(nb);
  AV_HIP(hipMemcpyAsync(hm.data(), part.p, sizeof(double) * nb, hipMemcpyDeviceToHost, k.s));
  AV_HIP(hipStreamSynchronize(k.s));
  double mx = 0.0;
  for (double v : hm) mx = std::max(mx, v);
  AV_CHECK(mx > 0.0, "coil construction: the image estimate is zero");
  AV_HIP(hipMemcpyAsync(part.as<double>() + nb, &mx, sizeof(double), hipMemcpyHostToDevice, k.s));
  const float tau = (float)p.b1Tau, sigma = (float)p.b1Sigma;
  hipLaunchKernelGGL(k_b1_pre, dim3(nblocks(N)), dim3(kTC), 0, k.s, u0, crec,
                     part.as<double>() + nb, tau, coils, N, lw.as<float2>(), iu.as<float>());
  AV_LAUNCH_CHECK();
  const float shrink = (float)(1.0 / (1.0 + p.b1Sigma / mu));
  for (unsigned it = 0; it < maxit; ++it) {
    hipLaunchKernelGGL(k_b1_dual, dim3(nblocks(n)), dim3(kTC), 0, k.s, ext.as<float2>(),
                       y.as<float2>(), k.W, k.H, coils, sigma, shrink);
    hipLaunchKernelGGL(k_b1_primal, dim3(nblocks(n)), dim3(kTC), 0, k.s, x1, ext.as<float2>(),
                       y.as<float2>(), lw.as<float2>(), iu.as<float>(), k.W, k.H, coils, tau);
  }
  AV_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_b1_normalize, dim3(nblocks(N)), dim3(kTC), 0, k.s, x1, coils, N);
  AV_LAUNCH_CHECK();
  AV_HIP(hipStreamSynchronize(k.s));  // `mx` is read by the copy above
}

// PerformCoilConstruction (coil_construction.cpp:447-557) from the
// time-averaged images.  u_out, b1_out: [C][H][W]; order: the coil order.
void coil_construction_from(const CoilCtx &k, const float2 *crec, const float2 *u0c,
                            const avionic_coil_params &p, float2 *u, float2 *b1,
                            std::vector<int> *order) {
  const long N = k.N;
  DevBuf u0, absb1, w, bm, cm, x1, umax;
  u0.ensure(sizeof(float2) * N);
  absb1.ensure(sizeof(float2) * N * k.C);
  w.ensure(sizeof(float) * N);
  bm.ensure(sizeof(float2) * N);
  cm.ensure(sizeof(float2) * N);
  x1.ensure(sizeof(float2) * N);
  umax.ensure(sizeof(float2) * N);
  // |u0| (coil_construction.cpp:459-461)
  hipLaunchKernelGGL(k_tar_combine, dim3(nblocks(N)), dim3(kTC), 0, k.s, u0c, u0.as<float2>(), 1,
                     N, 0);
  AV_LAUNCH_CHECK();
  coil_b1_from_uh1(k, u0.as<float2>(), crec, p.uH1mu, absb1.as<float2>());
  AV_HIP(hipMemsetAsync(b1, 0, sizeof(float2) * N * k.C, k.s));
  AV_HIP(hipMemsetAsync(u, 0, sizeof(float2) * N * k.C, k.s));
  AV_HIP(hipMemsetAsync(w.p, 0, w.bytes, k.s));
  AV_HIP(hipMemsetAsync(umax.p, 0, umax.bytes, k.s));
  std::vector<int> inds{coil_find_max_sum(k, crec)};
  AV_HIP(hipMemcpyAsync(b1 + inds[0] * N, absb1.as<float2>() + inds[0] * N, sizeof(float2) * N,
                        hipMemcpyDeviceToDevice, k.s));
  coil_utgv2(k, b1, crec, inds, p, umax.as<float2>());
  AV_HIP(hipMemcpyAsync(u, umax.p, sizeof(float2) * N, hipMemcpyDeviceToDevice, k.s));
  for (int cnt = 1; cnt < k.C; ++cnt) {
    const int prev = inds[cnt - 1];
    hipLaunchKernelGGL(k_weight_update, dim3(nblocks(N)), dim3(kTC), 0, k.s, w.as<float>(),
                       absb1.as<float2>() + prev * N, N);
    AV_LAUNCH_CHECK();
    // the unused coil overlapping most with the previous one
    const std::vector<double> ov = coil_abs_sums(k, crec, crec + prev * N, k.C);
    int best = -1;
    for (int c = 0; c < k.C; ++c) {
      if (std::find(inds.begin(), inds.end(), c) != inds.end()) continue;
      if (best < 0 || ov[c] > ov[best]) best = c;
    }
    inds.push_back(best);
    hipLaunchKernelGGL(k_weighted_inputs, dim3(nblocks(N)), dim3(kTC), 0, k.s,
                       absb1.as<float2>() + best * N, u + (cnt - 1) * N, crec + best * N,
                       w.as<float>(), bm.as<float2>(), cm.as<float2>(), N);
    AV_LAUNCH_CHECK();
    coil_b1_recon(k, bm.as<float2>(), cm.as<float2>(), 1, p.b1NrIt, p.b1Reg, p, x1.as<float2>());
    hipLaunchKernelGGL(k_apply_phase, dim3(nblocks(N)), dim3(kTC), 0, k.s,
                       absb1.as<float2>() + best * N, x1.as<float2>(), b1 + best * N, N);
    AV_LAUNCH_CHECK();
    AV_HIP(hipMemcpyAsync(umax.p, u + (cnt - 1) * N, sizeof(float2) * N, hipMemcpyDeviceToDevice,
                          k.s));
    coil_utgv2(k, b1, crec, inds, p, umax.as<float2>());
    AV_HIP(hipMemcpyAsync(u + cnt * N, umax.p, sizeof(float2) * N, hipMemcpyDeviceToDevice, k.s));
  }
  coil_b1_recon(k, umax.as<float2>(), crec, k.C, p.b1FinalNrIt, p.b1FinalReg, p, b1);
  AV_HIP(hipStreamSynchronize(k.s));
  if (order) *order = inds;
}

}  // namespace avionic
