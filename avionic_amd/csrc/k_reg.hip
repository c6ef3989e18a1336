// Regulariser kernels (see reg.hpp): the per-reconstruction launchers used by
// the API entry points and the gap / functional reductions.  The hot-loop
// versions over task lists live in k_tasks.hip; both share the bodies in
// reg_device.hpp.
#include <algorithm>
#include <cmath>
#include <type_traits>
#include <vector>

#include "reg_device.hpp"

namespace avionic {

// One voxel per lane (VPL = 1), or a pair of voxels of a row (VPL = 2, even
// widths), as in the task-list kernels.
__device__ inline bool block_voxels(const Grid3 &g, Vox &v) {
  return vox_of_block<false>(g, blockIdx.x, threadIdx.x, v);
}
__device__ inline bool block_voxels(const Grid3 &g, Vox2 &v) {
  return vox2_of_block(g, blockIdx.x, threadIdx.x, v);
}

template <int M, typename DT, int VPL>
__global__ __launch_bounds__(RT) void k_dual(Grid3 g, RegParams rp, const float2 *__restrict__ ext,
                                             DT *__restrict__ y, const DevScalars *__restrict__ sc) {
  typename std::conditional<VPL == 2, Vox2, Vox>::type v;
  if (!block_voxels(g, v)) return;
  dual_voxel<M, DT>(g, rp, ext, y, (float)sc->sigma, v);
}

template <int M, typename DT, int VPL>
__global__ __launch_bounds__(RT) void k_primal(Grid3 g, RegParams rp, float2 *__restrict__ x,
                                               float2 *__restrict__ ext, const DT *__restrict__ y,
                                               const float2 *__restrict__ khz,
                                               const DevScalars *__restrict__ sc) {
  typename std::conditional<VPL == 2, Vox2, Vox>::type v;
  if (!block_voxels(g, v)) return;
  primal_voxel<M, DT>(g, rp, x, ext, y, khz, (float)sc->tau, v);
}

template <int M>
__global__ __launch_bounds__(RT) void k_step_reg(Grid3 g, RegParams rp, const float2 *__restrict__ ext,
                                                 const float2 *__restrict__ x,
                                                 double *__restrict__ partial) {
  Vox v;
  double ak = 0.0, ax = 0.0;
  if (vox_of_block<true>(g, blockIdx.x, threadIdx.x, v)) step_voxel<M>(g, rp, ext, x, v, ak, ax);
  block_reduce2(ak, ax, partial + 2 * blockIdx.x);
}

// Deterministic final reduction and the step rule (pd_recon.cpp:40-61).
__global__ __launch_bounds__(RT) void k_step_finalize(const double *__restrict__ rp, int nreg,
                                                      const double *__restrict__ kp, int nk,
                                                      DevScalars *sc, double *sig_log,
                                                      double *tau_log, int log_index) {
  __shared__ double ra[RT], rb[RT];
  const int tid = threadIdx.x;
  double a = 0.0, b = 0.0;
  for (int j = tid; j < nreg; j += RT) {
    a += rp[2 * j];
    b += rp[2 * j + 1];
  }
  for (int j = tid; j < nk; j += RT) a += kp[j];
  ra[tid] = a;
  rb[tid] = b;
  __syncthreads();
  for (int s = RT / 2; s > 0; s >>= 1) {
    if (tid < s) {
      ra[tid] += ra[tid + s];
      rb[tid] += rb[tid + s];
    }
    __syncthreads();
  }
  if (tid == 0) {
    apply_step_rule(sc, ra[0], rb[0]);
    const int k = log_index >= 0 ? log_index : sc->iter;
    if (sig_log) sig_log[k] = sc->sigma;
    if (tau_log) tau_log[k] = sc->tau;
  }
}

__global__ void k_iter_bump(DevScalars *sc) {
  if (threadIdx.x == 0) sc->iter += 1;
}

// ---------------------------------------------------------------------------
// Functional value R(x): utils.cpp:1119-1287; get_tv.m, get_tgv2.m, get_ictgv2.m.
template <int M>
__global__ __launch_bounds__(RT) void k_reg_norm(Grid3 g, RegParams rp, double a0, double a1,
                                                 double wa, double wb,
                                                 const float2 *__restrict__ x,
                                                 double *__restrict__ partial) {
  const long i = (long)blockIdx.x * RT + threadIdx.x;
  double acc = 0.0;
  if (i < g.N) {
    const Vox v = vox_of(g, i);
    const long N = g.N;
    const Inv3 d1{rp.idx, rp.idy, rp.idt}, d2{rp.idx2, rp.idy2, rp.idt2};
    PlainAcc X{x};
    float2 gr[3], e[6];
    if (M == kTV) {
      grad3(X, 0, v, g, d1, gr);
      acc = norm3f(gr);
    } else if (M == kICTV) {  // utils::ICTVNorm (utils.cpp:1221-1243)
      grad3(DiffAcc{x, x + N}, 0, v, g, d1, gr);
      const double t1 = norm3f(gr);
      grad3(X, N, v, g, d2, gr);
      acc = a1 * (wa * t1 + wb * (double)norm3f(gr));
    } else {
      if (M == kTGV2) {
        grad3(X, 0, v, g, d1, gr);
      } else {
        DiffAcc UW{x, x + 4 * N};
        grad3(UW, 0, v, g, d1, gr);
      }
      for (int c = 0; c < 3; ++c) gr[c] = csub(gr[c], x[(1 + c) * N + i]);
      symgrad6(X, N, 2 * N, 3 * N, v, g, d1, e);
      double t1 = a1 * (double)norm3f(gr) + a0 * (double)norm6f(e);
      if (M == kTGV2) {
        acc = t1;
      } else {
        grad3(X, 4 * N, v, g, d2, gr);
        for (int c = 0; c < 3; ++c) gr[c] = csub(gr[c], x[(5 + c) * N + i]);
        symgrad6(X, 5 * N, 6 * N, 7 * N, v, g, d2, e);
        double t2 = a1 * (double)norm3f(gr) + a0 * (double)norm6f(e);
        acc = wa * t1 + wb * t2;
      }
    }
  }
  block_reduce2(acc, 0.0, partial + 2 * blockIdx.x);
}

// G* dual-residual l1 terms (tv.cpp:216-221, tgv2.cpp:146-161,
// ictgv2.cpp:195-227): |khz - div1 y1| (+ sum_c |divsym1 y2 + y1|_c
// + |div1 y1 - div2 y3| + sum_c |divsym2 y4 + y3|_c).  The duals are read in
// their storage format, each component group decoded with its own projection
// radius (y1: r1, y2: r2, y3: r3, y4: r4; ICTV keeps y1 and y3 only).
template <int M, typename DT>
__global__ __launch_bounds__(RT) void k_gstar_terms(Grid3 g, RegParams rp,
                                                    const float2 *__restrict__ khz,
                                                    const DT *__restrict__ y,
                                                    double *__restrict__ partial) {
  const long i = (long)blockIdx.x * RT + threadIdx.x;
  double acc = 0.0;
  if (i < g.N) {
    const Vox v = vox_of(g, i);
    const long N = g.N;
    const Inv3 d1{rp.idx, rp.idy, rp.idt}, d2{rp.idx2, rp.idy2, rp.idt2};
    const DualAcc<DT> Y1{y, rp.r1}, Y2{y, rp.r2}, Y3{y, rp.r3}, Y4{y, rp.r4};
    const unsigned ui = (unsigned)i;
    const float2 dv1 = div3(Y1, 0, N, 2 * N, v, g, d1);
    acc += sqrt((double)cabs2(csub(khz[i], dv1)));
    if (M == kICTV) {  // ictv.cpp:178-183: + |div1 y1 - div2 y3|
      const float2 dv3 = div3(Y3, 3 * N, 4 * N, 5 * N, v, g, d2);
      acc += sqrt((double)cabs2(csub(dv1, dv3)));
    } else if (M != kTV) {
      float2 sd[3];
      symdiv3(Y2, 3 * N, N, v, g, d1, sd);
      for (int c = 0; c < 3; ++c) acc += sqrt((double)cabs2(cadd(sd[c], Y1.at(c * N)(ui))));
      if (M == kICTGV2) {
        const float2 dv3 = div3(Y3, 9 * N, 10 * N, 11 * N, v, g, d2);
        acc += sqrt((double)cabs2(csub(dv1, dv3)));
        symdiv3(Y4, 12 * N, N, v, g, d2, sd);
        for (int c = 0; c < 3; ++c)
          acc += sqrt((double)cabs2(cadd(sd[c], Y3.at((9 + c) * N)(ui))));
      }
    }
  }
  block_reduce2(acc, 0.0, partial + 2 * blockIdx.x);
}

// Re(conj(d) z) and |z|^2 (the F*(z) part of G*).
__global__ __launch_bounds__(RT) void k_kspace_dots(const float2 *__restrict__ d,
                                                    const float2 *__restrict__ z, long n,
                                                    double *__restrict__ partial) {
  double a = 0.0, b = 0.0;
  for (long i = (long)blockIdx.x * RT + threadIdx.x; i < n; i += (long)gridDim.x * RT) {
    const float2 dd = d[i], zz = z[i];
    a += (double)dd.x * zz.x + (double)dd.y * zz.y;
    b += (double)zz.x * zz.x + (double)zz.y * zz.y;
  }
  block_reduce2(a, b, partial + 2 * blockIdx.x);
}

// ---------------------------------------------------------------------------
// Standalone stencils and projections.
__global__ void k_gradient(Grid3 g, Inv3 d, const float2 *u, float2 *out) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= g.N) return;
  const Vox v = vox_of(g, i);
  float2 gr[3];
  grad3(PlainAcc{u}, 0, v, g, d, gr);
  for (int c = 0; c < 3; ++c) out[c * g.N + i] = gr[c];
}
__global__ void k_symgrad(Grid3 g, Inv3 d, const float2 *vv, float2 *out) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= g.N) return;
  const Vox v = vox_of(g, i);
  float2 e[6];
  symgrad6(PlainAcc{vv}, 0, g.N, 2 * g.N, v, g, d, e);
  for (int k = 0; k < 6; ++k) out[k * g.N + i] = e[k];
}
__global__ void k_divergence(Grid3 g, Inv3 d, const float2 *p, float2 *out) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= g.N) return;
  const Vox v = vox_of(g, i);
  out[i] = div3(PlainAcc{p}, 0, g.N, 2 * g.N, v, g, d);
}
__global__ void k_symdiv(Grid3 g, Inv3 d, const float2 *q, float2 *out) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= g.N) return;
  const Vox v = vox_of(g, i);
  float2 sd[3];
  symdiv3(PlainAcc{q}, 0, g.N, v, g, d, sd);
  for (int c = 0; c < 3; ++c) out[c * g.N + i] = sd[c];
}
// y /= max(1, |y| * scale), utils.cpp:1297-1305.
__global__ void k_prox(int ncomp, float2 *y, long n, float scale) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float2 p[6];
  for (int k = 0; k < ncomp; ++k) p[k] = y[k * n + i];
  float nn;
  if (ncomp == 3)
    nn = norm3f(p);
  else
    nn = norm6f(p);
  const float s = fmaxf(1.f, nn * scale);
  for (int k = 0; k < ncomp; ++k) y[k * n + i] = cscale(p[k], 1.f / s);
}

// ---------------------------------------------------------------------------
// host side

RegParams make_reg_params(int method, const double d6[6], double alpha0, double alpha1,
                          double alpha) {
  RegParams rp{};
  rp.idx = (float)(1.0 / d6[0]);
  rp.idy = (float)(1.0 / d6[1]);
  rp.idt = (float)(1.0 / d6[2]);
  rp.idx2 = (float)(1.0 / d6[3]);
  rp.idy2 = (float)(1.0 / d6[4]);
  rp.idt2 = (float)(1.0 / d6[5]);
  if (method == kTV) {
    rp.r1 = 1.f;  // tv.cpp:150
  } else if (method == kTGV2) {
    rp.r1 = (float)alpha1;  // tgv2.cpp:258-259
    rp.r2 = (float)alpha0;
  } else {
    const double m = std::min(alpha, 1.0 - alpha);  // ictgv2.cpp:367-381
    rp.r1 = (float)(alpha1 * (alpha / m));
    rp.r2 = (float)(alpha0 * (alpha / m));
    rp.r3 = (float)(alpha1 * ((1.0 - alpha) / m));
    rp.r4 = (float)(alpha0 * ((1.0 - alpha) / m));
  }
  return rp;
}

static inline unsigned nblk(long n) { return (unsigned)((n + RT - 1) / RT); }

void launch_dual(int method, int ds, const Grid3 &g, const RegParams &rp, const float2 *ext, void *y,
                 const DevScalars *sc, hipStream_t s) {
  check_grid(g);
  const bool pairs = g.W % 2 == 0;
  dim3 grid((unsigned)(pairs ? pair_blocks(g) : step_blocks(g))), block(RT);
#define AV_DUAL(M, DT)                                                                      \
  do {                                                                                      \
    if (pairs)                                                                              \
      hipLaunchKernelGGL((k_dual<M, DT, 2>), grid, block, 0, s, g, rp, ext, (DT *)y, sc);   \
    else                                                                                    \
      hipLaunchKernelGGL((k_dual<M, DT, 1>), grid, block, 0, s, g, rp, ext, (DT *)y, sc);   \
  } while (0)
#define AV_DUAL_M(DT)                        \
  do {                                       \
    if (method == kTV) AV_DUAL(kTV, DT);     \
    else if (method == kTGV2) AV_DUAL(kTGV2, DT); \
    else if (method == kICTGV2) AV_DUAL(kICTGV2, DT);  \
    else AV_DUAL(kICTV, DT);                           \
  } while (0)
  if (ds == kDualF32) AV_DUAL_M(float2);
  else if (ds == kDualBF16) AV_DUAL_M(bf16x2);
  else AV_DUAL_M(q16x2);
#undef AV_DUAL_M
#undef AV_DUAL
  AV_LAUNCH_CHECK();
}

void launch_primal(int method, int ds, const Grid3 &g, const RegParams &rp, float2 *x, float2 *ext,
                   const void *y, const float2 *khz, const DevScalars *sc, hipStream_t s) {
  check_grid(g);
  const bool pairs = g.W % 2 == 0;
  dim3 grid((unsigned)(pairs ? pair_blocks(g) : step_blocks(g))), block(RT);
#define AV_PRIMAL(M, DT)                                                                     \
  do {                                                                                       \
    if (pairs)                                                                               \
      hipLaunchKernelGGL((k_primal<M, DT, 2>), grid, block, 0, s, g, rp, x, ext, (const DT *)y, \
                         khz, sc);                                                           \
    else                                                                                     \
      hipLaunchKernelGGL((k_primal<M, DT, 1>), grid, block, 0, s, g, rp, x, ext, (const DT *)y, \
                         khz, sc);                                                           \
  } while (0)
#define AV_PRIMAL_M(DT)                          \
  do {                                           \
    if (method == kTV) AV_PRIMAL(kTV, DT);       \
    else if (method == kTGV2) AV_PRIMAL(kTGV2, DT); \
    else if (method == kICTGV2) AV_PRIMAL(kICTGV2, DT);  \
    else AV_PRIMAL(kICTV, DT);                           \
  } while (0)
  if (ds == kDualF32) AV_PRIMAL_M(float2);
  else if (ds == kDualBF16) AV_PRIMAL_M(bf16x2);
  else AV_PRIMAL_M(q16x2);
#undef AV_PRIMAL_M
#undef AV_PRIMAL
  AV_LAUNCH_CHECK();
}

int step_blocks(const Grid3 &g) { return (int)(tiles_per_frame(g) * g.T); }
int pair_blocks(const Grid3 &g) { return (int)(pair_tiles_per_frame(g) * g.T); }
int norm_blocks(const Grid3 &g) { return (int)nblk(g.N); }

void launch_step_reg(int method, const Grid3 &g, const RegParams &rp, const float2 *ext,
                     const float2 *x, double *partial, hipStream_t s) {
  check_grid(g);
  dim3 grid((unsigned)step_blocks(g)), block(RT);
  if (method == kTV) hipLaunchKernelGGL(k_step_reg<kTV>, grid, block, 0, s, g, rp, ext, x, partial);
  else if (method == kTGV2) hipLaunchKernelGGL(k_step_reg<kTGV2>, grid, block, 0, s, g, rp, ext, x, partial);
  else if (method == kICTGV2) hipLaunchKernelGGL(k_step_reg<kICTGV2>, grid, block, 0, s, g, rp, ext, x, partial);
  else hipLaunchKernelGGL(k_step_reg<kICTV>, grid, block, 0, s, g, rp, ext, x, partial);
  AV_LAUNCH_CHECK();
}

void launch_step_finalize(const double *reg_partial, int nreg, const double *k_partial, int nk,
                          DevScalars *sc, double *sig_log, double *tau_log, int log_index,
                          hipStream_t s) {
  hipLaunchKernelGGL(k_step_finalize, dim3(1), dim3(RT), 0, s, reg_partial, nreg, k_partial, nk, sc,
                     sig_log, tau_log, log_index);
  AV_LAUNCH_CHECK();
}

void launch_iter_bump(DevScalars *sc, hipStream_t s) {
  hipLaunchKernelGGL(k_iter_bump, dim3(1), dim3(64), 0, s, sc);
  AV_LAUNCH_CHECK();
}

void launch_reg_norm(int method, const Grid3 &g, const RegParams &rp, double alpha0, double alpha1,
                     double alpha, const float2 *x, double *partial, hipStream_t s) {
  check_grid(g);
  dim3 grid(nblk(g.N)), block(RT);
  double wa = 1.0, wb = 1.0;
  if (method == kICTGV2 || method == kICTV) {
    const double m = std::min(alpha, 1.0 - alpha);
    wa = alpha / m;
    wb = (1.0 - alpha) / m;
  }
  if (method == kTV) hipLaunchKernelGGL(k_reg_norm<kTV>, grid, block, 0, s, g, rp, alpha0, alpha1, wa, wb, x, partial);
  else if (method == kTGV2) hipLaunchKernelGGL(k_reg_norm<kTGV2>, grid, block, 0, s, g, rp, alpha0, alpha1, wa, wb, x, partial);
  else if (method == kICTGV2) hipLaunchKernelGGL(k_reg_norm<kICTGV2>, grid, block, 0, s, g, rp, alpha0, alpha1, wa, wb, x, partial);
  else hipLaunchKernelGGL(k_reg_norm<kICTV>, grid, block, 0, s, g, rp, alpha0, alpha1, wa, wb, x, partial);
  AV_LAUNCH_CHECK();
}

void launch_gstar_terms(int method, int ds, const Grid3 &g, const RegParams &rp,
                        const float2 *khz, const void *y, double *partial, hipStream_t s) {
  check_grid(g);
  dim3 grid(nblk(g.N)), block(RT);
#define AV_GS(M, DT) \
  hipLaunchKernelGGL((k_gstar_terms<M, DT>), grid, block, 0, s, g, rp, khz, (const DT *)y, partial)
#define AV_GS_M(DT)                                  \
  do {                                               \
    if (method == kTV) AV_GS(kTV, DT);               \
    else if (method == kTGV2) AV_GS(kTGV2, DT);      \
    else if (method == kICTGV2) AV_GS(kICTGV2, DT);  \
    else AV_GS(kICTV, DT);                           \
  } while (0)
  if (ds == kDualF32) AV_GS_M(float2);
  else if (ds == kDualBF16) AV_GS_M(bf16x2);
  else AV_GS_M(q16x2);
#undef AV_GS_M
#undef AV_GS
  AV_LAUNCH_CHECK();
}

int reduce_blocks(long n) {
  long b = (n + RT - 1) / RT;
  return (int)std::max(1L, std::min(b, 2048L));
}

void launch_kspace_dots(const float2 *d, const float2 *z, long n, double *partial, hipStream_t s) {
  hipLaunchKernelGGL(k_kspace_dots, dim3(reduce_blocks(n)), dim3(RT), 0, s, d, z, n, partial);
  AV_LAUNCH_CHECK();
}

double reduce_host(const double *partial, int n, int stride, int offset, hipStream_t s) {
  std::vector<double> h((size_t)n * stride);
  AV_HIP(hipMemcpyAsync(h.data(), partial, sizeof(double) * h.size(), hipMemcpyDeviceToHost, s));
  AV_HIP(hipStreamSynchronize(s));
  double acc = 0.0;
  for (int i = 0; i < n; ++i) acc += h[(size_t)i * stride + offset];
  return acc;
}

void launch_gradient(const Grid3 &g, float idx, float idy, float idt, const float2 *u, float2 *out,
                     hipStream_t s) {
  check_grid(g);
  hipLaunchKernelGGL(k_gradient, dim3(nblk(g.N)), dim3(RT), 0, s, g, Inv3{idx, idy, idt}, u, out);
  AV_LAUNCH_CHECK();
}
void launch_symgrad(const Grid3 &g, float idx, float idy, float idt, const float2 *v, float2 *out,
                    hipStream_t s) {
  check_grid(g);
  hipLaunchKernelGGL(k_symgrad, dim3(nblk(g.N)), dim3(RT), 0, s, g, Inv3{idx, idy, idt}, v, out);
  AV_LAUNCH_CHECK();
}
void launch_divergence(const Grid3 &g, float idx, float idy, float idt, const float2 *p,
                       float2 *out, hipStream_t s) {
  check_grid(g);
  hipLaunchKernelGGL(k_divergence, dim3(nblk(g.N)), dim3(RT), 0, s, g, Inv3{idx, idy, idt}, p, out);
  AV_LAUNCH_CHECK();
}
void launch_symdiv(const Grid3 &g, float idx, float idy, float idt, const float2 *q, float2 *out,
                   hipStream_t s) {
  check_grid(g);
  hipLaunchKernelGGL(k_symdiv, dim3(nblk(g.N)), dim3(RT), 0, s, g, Inv3{idx, idy, idt}, q, out);
  AV_LAUNCH_CHECK();
}
void launch_prox(int ncomp, float2 *y, long n, float scale, hipStream_t s) {
  hipLaunchKernelGGL(k_prox, dim3(nblk(n)), dim3(RT), 0, s, ncomp, y, n, scale);
  AV_LAUNCH_CHECK();
}

}  // namespace avionic
