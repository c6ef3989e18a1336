// FP64 operator-norm power iteration (opnorm.hpp): FP64 stencils of the
// regulariser operator and its adjoint, an FP64 Stockham FFT, the FP64
// Cartesian K^H K, deterministic FP64 reductions and the host driver.
#include <cmath>
#include <map>
#include <mutex>
#include <utility>
#include <vector>

#include "opnorm.hpp"
#include "solver.hpp"

namespace avionic {

namespace {

constexpr int kT64 = 256;

__host__ __device__ inline double2 dadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__host__ __device__ inline double2 dsub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__host__ __device__ inline double2 dscale(double2 a, double s) { return make_double2(a.x * s, a.y * s); }
__host__ __device__ inline double2 dmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__host__ __device__ inline double2 dmulc(double2 a, double2 b) {  // conj(a) b
  return make_double2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}
__device__ inline double2 up(float2 a) { return make_double2((double)a.x, (double)a.y); }

inline int blocks_of(long n) { return (int)std::max<long>(1, std::min<long>(8192, (n + kT64 - 1) / kT64)); }

// ---------------------------------------------------------------- start vector
__device__ inline double hash_unit(unsigned long long j) {
  unsigned long long z = j + kOpnormSeed;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  z ^= z >> 31;
  return (double)(z >> 11) * (2.0 / 9007199254740992.0) - 1.0;  // [-1, 1)
}
__global__ void k_start(double2 *x, long n) {
  for (long i = (long)blockIdx.x * kT64 + threadIdx.x; i < n; i += (long)gridDim.x * kT64)
    x[i] = make_double2(hash_unit(2 * (unsigned long long)i), hash_unit(2 * (unsigned long long)i + 1));
}

// ---------------------------------------------------------------- stencils
// Differences with the reference's boundary rules (matlab/get_sp_fdif.m,
// get_sp_bdif.m): forward 0 at the last index; backward u_0 first,
// u_i - u_{i-1} inside, -u_{n-2} last.
struct G64 {
  int W, H, T;
  long N;
};
__device__ inline double2 fd(const double2 *a, long i, long s, int pos, int len) {
  return pos < len - 1 ? dsub(a[i + s], a[i]) : make_double2(0.0, 0.0);
}
__device__ inline double2 bd(const double2 *a, long i, long s, int pos, int len) {
  if (len == 1) return make_double2(0.0, 0.0);
  if (pos == 0) return a[i];
  if (pos == len - 1) return dscale(a[i - s], -1.0);
  return dsub(a[i], a[i - s]);
}
struct W3 {
  double x, y, t;  // 1/dx, 1/dy, 1/dt
};
// gradient of the field at a (+ optional subtraction field b: grad(a - b))
__device__ inline void grad64(const double2 *a, const double2 *b, long i, int x, int y, int t,
                              const G64 &g, W3 w, double2 *o) {
  double2 gx = fd(a, i, 1, x, g.W), gy = fd(a, i, g.W, y, g.H), gt = fd(a, i, (long)g.W * g.H, t, g.T);
  if (b) {
    gx = dsub(gx, fd(b, i, 1, x, g.W));
    gy = dsub(gy, fd(b, i, g.W, y, g.H));
    gt = dsub(gt, fd(b, i, (long)g.W * g.H, t, g.T));
  }
  o[0] = dscale(gx, w.x);
  o[1] = dscale(gy, w.y);
  o[2] = dscale(gt, w.t);
}
// symmetrised gradient (xx, yy, tt, xy, xt, yt) of v0, v1, v2
__device__ inline void symgrad64(const double2 *v0, const double2 *v1, const double2 *v2, long i,
                                 int x, int y, int t, const G64 &g, W3 w, double2 *e) {
  const long sy = g.W, st = (long)g.W * g.H;
  e[0] = dscale(bd(v0, i, 1, x, g.W), w.x);
  e[1] = dscale(bd(v1, i, sy, y, g.H), w.y);
  e[2] = dscale(bd(v2, i, st, t, g.T), w.t);
  e[3] = dscale(dadd(dscale(bd(v0, i, sy, y, g.H), w.y), dscale(bd(v1, i, 1, x, g.W), w.x)), 0.5);
  e[4] = dscale(dadd(dscale(bd(v0, i, st, t, g.T), w.t), dscale(bd(v2, i, 1, x, g.W), w.x)), 0.5);
  e[5] = dscale(dadd(dscale(bd(v1, i, st, t, g.T), w.t), dscale(bd(v2, i, sy, y, g.H), w.y)), 0.5);
}
// div (= -grad^T) of the 3-vector p0, p1, p2
__device__ inline double2 div64(const double2 *p0, const double2 *p1, const double2 *p2, long i,
                                int x, int y, int t, const G64 &g, W3 w) {
  return dadd(dadd(dscale(bd(p0, i, 1, x, g.W), w.x), dscale(bd(p1, i, g.W, y, g.H), w.y)),
              dscale(bd(p2, i, (long)g.W * g.H, t, g.T), w.t));
}
// symmetric divergence of the 6-vector q (components q + k N)
__device__ inline void symdiv64(const double2 *q, long N, long i, int x, int y, int t,
                                const G64 &g, W3 w, double2 *o) {
  const long sy = g.W, st = (long)g.W * g.H;
  const double2 *q0 = q, *q1 = q + N, *q2 = q + 2 * N, *q3 = q + 3 * N, *q4 = q + 4 * N,
                *q5 = q + 5 * N;
  o[0] = dadd(dadd(dscale(fd(q0, i, 1, x, g.W), w.x), dscale(fd(q3, i, sy, y, g.H), w.y)),
              dscale(fd(q4, i, st, t, g.T), w.t));
  o[1] = dadd(dadd(dscale(fd(q3, i, 1, x, g.W), w.x), dscale(fd(q1, i, sy, y, g.H), w.y)),
              dscale(fd(q5, i, st, t, g.T), w.t));
  o[2] = dadd(dadd(dscale(fd(q4, i, 1, x, g.W), w.x), dscale(fd(q5, i, sy, y, g.H), w.y)),
              dscale(fd(q2, i, st, t, g.T), w.t));
}

// K_reg x (D components), as the dual update applies it (tv.cpp:140-144,
// tgv2.cpp:238-253, ictgv2.cpp:338-361, ictv.cpp:266-271).
template <int M>
__global__ __launch_bounds__(kT64) void k_reg_K64(G64 g, W3 w1, W3 w2, const double2 *__restrict__ x,
                                                  double2 *__restrict__ y) {
  const long N = g.N;
  for (long i = (long)blockIdx.x * kT64 + threadIdx.x; i < N; i += (long)gridDim.x * kT64) {
    const int xx = (int)(i % g.W), yy = (int)((i / g.W) % g.H), tt = (int)(i / ((long)g.W * g.H));
    double2 o[6];
    if (M == kTV) {
      grad64(x, nullptr, i, xx, yy, tt, g, w1, o);
      for (int c = 0; c < 3; ++c) y[c * N + i] = o[c];
    } else if (M == kICTV) {
      grad64(x, x + N, i, xx, yy, tt, g, w1, o);
      for (int c = 0; c < 3; ++c) y[c * N + i] = o[c];
      grad64(x + N, nullptr, i, xx, yy, tt, g, w2, o);
      for (int c = 0; c < 3; ++c) y[(3 + c) * N + i] = o[c];
    } else {
      grad64(x, M == kICTGV2 ? x + 4 * N : nullptr, i, xx, yy, tt, g, w1, o);
      for (int c = 0; c < 3; ++c) y[c * N + i] = dsub(o[c], x[(1 + c) * N + i]);
      symgrad64(x + N, x + 2 * N, x + 3 * N, i, xx, yy, tt, g, w1, o);
      for (int c = 0; c < 6; ++c) y[(3 + c) * N + i] = o[c];
      if (M == kICTGV2) {
        grad64(x + 4 * N, nullptr, i, xx, yy, tt, g, w2, o);
        for (int c = 0; c < 3; ++c) y[(9 + c) * N + i] = dsub(o[c], x[(5 + c) * N + i]);
        symgrad64(x + 5 * N, x + 6 * N, x + 7 * N, i, xx, yy, tt, g, w2, o);
        for (int c = 0; c < 6; ++c) y[(12 + c) * N + i] = o[c];
      }
    }
  }
}

// K_reg^* y (P components), the operator of the primal descent (tv.cpp:156-159,
// tgv2.cpp:271-285, ictgv2.cpp:389-416, ictv.cpp:301-310): adjoint of K_reg
// for the pairing with off-diagonal weight 2.
template <int M>
__global__ __launch_bounds__(kT64) void k_reg_KT64(G64 g, W3 w1, W3 w2, const double2 *__restrict__ y,
                                                   double2 *__restrict__ x) {
  const long N = g.N;
  for (long i = (long)blockIdx.x * kT64 + threadIdx.x; i < N; i += (long)gridDim.x * kT64) {
    const int xx = (int)(i % g.W), yy = (int)((i / g.W) % g.H), tt = (int)(i / ((long)g.W * g.H));
    const double2 d1 = div64(y, y + N, y + 2 * N, i, xx, yy, tt, g, w1);
    x[i] = dscale(d1, -1.0);
    if (M == kICTV) {
      const double2 d3 = div64(y + 3 * N, y + 4 * N, y + 5 * N, i, xx, yy, tt, g, w2);
      x[N + i] = dsub(d1, d3);
    } else if (M != kTV) {
      double2 sd[3];
      symdiv64(y + 3 * N, N, i, xx, yy, tt, g, w1, sd);
      for (int c = 0; c < 3; ++c) x[(1 + c) * N + i] = dscale(dadd(y[c * N + i], sd[c]), -1.0);
      if (M == kICTGV2) {
        const double2 d3 = div64(y + 9 * N, y + 10 * N, y + 11 * N, i, xx, yy, tt, g, w2);
        x[4 * N + i] = dsub(d1, d3);
        symdiv64(y + 12 * N, N, i, xx, yy, tt, g, w2, sd);
        for (int c = 0; c < 3; ++c)
          x[(5 + c) * N + i] = dscale(dadd(y[(9 + c) * N + i], sd[c]), -1.0);
      }
    }
  }
}

// ---------------------------------------------------------------- FFT (FP64)
// One block transforms one length-L sequence along an axis of a strided
// array: element i of transform b sits at base(b) + i * stride, with
// base(b) = (b / inner) * outer_stride + (b % inner) * inner_stride.  The
// sequence is staged in LDS and each Stockham stage writes the other half of
// a ping-pong pair.  centered: fftshift(F(ifftshift(.))).
constexpr int kMaxL64 = 1024;
struct Axis64 {
  int L, nst, centered;
  int radix[kMaxStages], ns[kMaxStages];
  long stride, inner, inner_stride, outer_stride;
  const double2 *tw;  // exp(-2 pi i m / L), m < L
};

template <int R>
__device__ inline void stage64(const double2 *in, double2 *out, const double2 *tw, int L, int Ns,
                               double sign, int tid) {
  const int nbf = L / R;
  for (int j = tid; j < nbf; j += kT64) {
    const int k = j % Ns;
    double2 v[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      double2 a = in[j + r * nbf];
      if (r > 0 && Ns > 1) {
        double2 w = tw[(r * k * (L / (Ns * R))) % L];
        w.y *= -sign;
        a = dmul(a, w);
      }
      v[r] = a;
    }
    const int o = (j / Ns) * Ns * R + k;
#pragma unroll
    for (int q = 0; q < R; ++q) {
      double2 acc = v[0];
#pragma unroll
      for (int r = 1; r < R; ++r) {
        double2 w = tw[((r * q) % R) * nbf];
        w.y *= -sign;
        acc = dadd(acc, dmul(v[r], w));
      }
      out[o + q * Ns] = acc;
    }
  }
}

__global__ __launch_bounds__(kT64) void k_fft64_axis(double2 *data, Axis64 a, double sign,
                                                     double scale) {
  __shared__ double2 buf[2][kMaxL64];
  __shared__ double2 tws[kMaxL64];
  const int tid = threadIdx.x, L = a.L;
  const long b = blockIdx.x;
  const long base = (b / a.inner) * a.outer_stride + (b % a.inner) * a.inner_stride;
  for (int i = tid; i < L; i += kT64) {
    const int src = a.centered ? (i + L / 2) % L : i;
    buf[0][i] = data[base + (long)src * a.stride];
    tws[i] = a.tw[i];
  }
  __syncthreads();
  int cur = 0;
  for (int s = 0; s < a.nst; ++s) {
    const double2 *in = buf[cur];
    double2 *out = buf[1 - cur];
    switch (a.radix[s]) {
      case 16: stage64<16>(in, out, tws, L, a.ns[s], sign, tid); break;
      case 8: stage64<8>(in, out, tws, L, a.ns[s], sign, tid); break;
      case 4: stage64<4>(in, out, tws, L, a.ns[s], sign, tid); break;
      case 2: stage64<2>(in, out, tws, L, a.ns[s], sign, tid); break;
      case 3: stage64<3>(in, out, tws, L, a.ns[s], sign, tid); break;
      case 5: stage64<5>(in, out, tws, L, a.ns[s], sign, tid); break;
      case 7: stage64<7>(in, out, tws, L, a.ns[s], sign, tid); break;
      case 11: stage64<11>(in, out, tws, L, a.ns[s], sign, tid); break;
      case 13: stage64<13>(in, out, tws, L, a.ns[s], sign, tid); break;
      default: break;
    }
    __syncthreads();
    cur = 1 - cur;
  }
  for (int i = tid; i < L; i += kT64) {
    const int src = a.centered ? (i + (L + 1) / 2) % L : i;
    data[base + (long)i * a.stride] = dscale(buf[cur][src], scale);
  }
}

// Plans (radices from make_fft_plan, double twiddles) cached per device and
// length for the life of the process.
struct Plan64 {
  FftPlan p{};  // radices / strides
  double2 *tw = nullptr;
};
std::mutex g_plan_mu;
std::map<std::pair<int, int>, Plan64> g_plans;

const Plan64 &plan64(int L) {
  AV_CHECK(L >= 1 && L <= kMaxL64, "FP64 transform length out of range (1..1024)");
  int dev = 0;
  AV_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lock(g_plan_mu);
  auto it = g_plans.find({dev, L});
  if (it != g_plans.end()) return it->second;
  Plan64 pl;
  void *tw32 = nullptr;
  make_fft_plan(L, pl.p, &tw32);  // radices and strides; its float twiddles are not used
  (void)hipFree(tw32);
  std::vector<double2> tw(L);
  for (int m = 0; m < L; ++m) {
    const double a = -2.0 * M_PI * (double)m / (double)L;
    tw[m] = make_double2(std::cos(a), std::sin(a));
  }
  AV_HIP(hipMalloc((void **)&pl.tw, sizeof(double2) * L));
  AV_HIP(hipMemcpy(pl.tw, tw.data(), sizeof(double2) * L, hipMemcpyHostToDevice));
  return g_plans.emplace(std::make_pair(dev, L), pl).first->second;
}

// Transform along one axis: `count` sequences of length L.
void fft64_axis(double2 *data, int L, long stride, long count, long inner, long inner_stride,
                long outer_stride, int centered, double sign, double scale, hipStream_t s) {
  if (L == 1 && scale == 1.0) return;
  const Plan64 &pl = plan64(L);
  Axis64 a{};
  a.L = L;
  a.nst = pl.p.nst;
  a.centered = centered;
  for (int k = 0; k < pl.p.nst; ++k) {
    a.radix[k] = pl.p.radix[k];
    a.ns[k] = pl.p.ns[k];
  }
  a.stride = stride;
  a.inner = inner;
  a.inner_stride = inner_stride;
  a.outer_stride = outer_stride;
  a.tw = pl.tw;
  AV_CHECK(count < (1L << 31), "too many FP64 transforms");
  hipLaunchKernelGGL(k_fft64_axis, dim3((unsigned)count), dim3(kT64), 0, s, data, a, sign, scale);
  AV_LAUNCH_CHECK();
}

// Unitary 2D transform of `planes` H x W planes (rows then columns).
void fft64_2d(double2 *v, int H, int W, long planes, int centered, double sign, hipStream_t s) {
  const long HW = (long)H * W;
  fft64_axis(v, W, 1, planes * H, planes * H, W, 0, centered, sign, 1.0, s);       // rows
  fft64_axis(v, H, W, planes * W, W, 1, HW, centered, sign, 1.0 / std::sqrt((double)HW), s);
}

// ---------------------------------------------------------------- data term
// Coil images of the Cartesian 2D+t operator: z[t][c] = b1_c u_t.
__global__ void k_expand64(const double2 *__restrict__ u, const float2 *__restrict__ b1,
                           double2 *__restrict__ z, int C, long HW, long n) {
  for (long e = (long)blockIdx.x * kT64 + threadIdx.x; e < n; e += (long)gridDim.x * kT64) {
    const long p = e % HW, tc = e / HW, t = tc / C, c = tc % C;
    z[e] = dmul(up(b1[c * HW + p]), u[t * HW + p]);
  }
}
// z[t][c][y][x] *= mask(t, y, x)^2 (the mask of K and the mask K^H applies
// to its input), mask from the compact-row tables.
__global__ void k_mask2_64(const CartGeom g, double2 *__restrict__ z, long n) {
  const long HW = (long)g.H * g.W;
  for (long e = (long)blockIdx.x * kT64 + threadIdx.x; e < n; e += (long)gridDim.x * kT64) {
    const long p = e % HW, t = e / HW / g.C;
    const int y = (int)(p / g.W), x = (int)(p % g.W);
    const int k = g.rowmap[t * g.H + y];
    const double m = k >= 0 ? (double)g.mrows[((long)g.row_off[t] + k) * g.W + x] : 0.0;
    z[e] = dscale(z[e], m * m);
  }
}
// out_t += sum_c conj(b1_c) z[t][c]
__global__ void k_coil_sum64(const double2 *__restrict__ z, const float2 *__restrict__ b1,
                             double2 *__restrict__ out, int C, long HW, long N) {
  for (long i = (long)blockIdx.x * kT64 + threadIdx.x; i < N; i += (long)gridDim.x * kT64) {
    const long p = i % HW, t = i / HW;
    double2 acc = out[i];
    for (int c = 0; c < C; ++c) acc = dadd(acc, dmulc(up(b1[c * HW + p]), z[(t * C + c) * HW + p]));
    out[i] = acc;
  }
}
// volumetric: z[c][i] = b1[c][i] u[i]; z *= mask^2; out[i] += sum_c conj(b1[c][i]) z[c][i]
__global__ void k3_expand64(const double2 *__restrict__ u, const float2 *__restrict__ b1,
                            double2 *__restrict__ z, long N, long n) {
  for (long e = (long)blockIdx.x * kT64 + threadIdx.x; e < n; e += (long)gridDim.x * kT64)
    z[e] = dmul(up(b1[e]), u[e % N]);
}
__global__ void k3_mask2_64(const float *__restrict__ mask, double2 *__restrict__ z, long N,
                            long n) {
  for (long e = (long)blockIdx.x * kT64 + threadIdx.x; e < n; e += (long)gridDim.x * kT64) {
    const double m = mask ? (double)mask[e % N] : 1.0;
    z[e] = dscale(z[e], m * m);
  }
}
__global__ void k3_coil_sum64(const double2 *__restrict__ z, const float2 *__restrict__ b1,
                              double2 *__restrict__ out, int C, long N) {
  for (long i = (long)blockIdx.x * kT64 + threadIdx.x; i < N; i += (long)gridDim.x * kT64) {
    double2 acc = out[i];
    for (int c = 0; c < C; ++c) acc = dadd(acc, dmulc(up(b1[c * N + i]), z[c * N + i]));
    out[i] = acc;
  }
}

// Non-Cartesian K^H K in double precision, on the oversampled grid the
// gridding passes use (noncart.hpp): image (y, x) sits at grid
// ((y - H/2) mod Gy, (x - W/2) mod Gx), deapodised; K interpolates each point
// from its kNcTaps x kNcTaps neighbourhood; K^H gathers every grid cell from
// its cell -> point list (the adjoint of the interpolation), inverse
// transforms and crops.
__global__ void k_nc64_fill(const Nc64 g, const double2 *__restrict__ u,
                            const float2 *__restrict__ b1, double2 *__restrict__ grid, long n) {
  const long HW = (long)g.H * g.W, G2 = (long)g.Gx * g.Gy;
  for (long e = (long)blockIdx.x * kT64 + threadIdx.x; e < n; e += (long)gridDim.x * kT64) {
    const long plane = e / G2, cell = e % G2;  // plane = t C + c
    const int gy = (int)(cell / g.Gx), gx = (int)(cell % g.Gx);
    const int y = (gy + g.H / 2) % g.Gy, x = (gx + g.W / 2) % g.Gx;
    double2 v = make_double2(0.0, 0.0);
    if (y < g.H && x < g.W) {
      const long t = plane / g.C, c = plane % g.C, p = (long)y * g.W + x;
      v = dscale(dmul(up(b1[c * HW + p]), u[t * HW + p]), g.dxi[x] * g.dyi[y]);
    }
    grid[e] = v;
  }
}
constexpr int kTaps64 = 4;
__global__ void k_nc64_interp(const Nc64 g, const double2 *__restrict__ grid,
                              double2 *__restrict__ z, long n) {
  const long G2 = (long)g.Gx * g.Gy;
  for (long e = (long)blockIdx.x * kT64 + threadIdx.x; e < n; e += (long)gridDim.x * kT64) {
    const long plane = e / g.M, m = e % g.M, t = plane / g.C, pm = t * g.M + m;
    const int2 q = g.q[pm];
    const double *w = g.w + pm * 2 * kTaps64;
    const double2 *gp = grid + plane * G2;
    double2 acc = make_double2(0.0, 0.0);
    for (int dy = 0; dy < kTaps64; ++dy) {
      const int gy = (q.y + dy) % g.Gy;
      double2 row = make_double2(0.0, 0.0);
      for (int dx = 0; dx < kTaps64; ++dx)
        row = dadd(row, dscale(gp[(long)gy * g.Gx + (q.x + dx) % g.Gx], w[dx]));
      acc = dadd(acc, dscale(row, w[kTaps64 + dy]));
    }
    z[e] = dscale(acc, g.sw[pm]);
  }
}
__global__ void k_nc64_spread(const Nc64 g, const double2 *__restrict__ z,
                              double2 *__restrict__ grid, long n) {
  const long G2 = (long)g.Gx * g.Gy;
  for (long e = (long)blockIdx.x * kT64 + threadIdx.x; e < n; e += (long)gridDim.x * kT64) {
    const long plane = e / G2, cell = e % G2, t = plane / g.C;
    const int *cp = g.cptr + t * (G2 + 1);
    const double2 *zp = z + plane * g.M;
    double2 acc = make_double2(0.0, 0.0);
    for (int k = cp[cell]; k < cp[cell + 1]; ++k) acc = dadd(acc, dscale(zp[g.ce[k].x], g.ce_w[k]));
    grid[e] = acc;
  }
}
__global__ void k_nc64_crop(const Nc64 g, const double2 *__restrict__ grid,
                            const float2 *__restrict__ b1, double2 *__restrict__ out, long N) {
  const long HW = (long)g.H * g.W, G2 = (long)g.Gx * g.Gy;
  for (long i = (long)blockIdx.x * kT64 + threadIdx.x; i < N; i += (long)gridDim.x * kT64) {
    const long t = i / HW, p = i % HW;
    const int y = (int)(p / g.W), x = (int)(p % g.W);
    const long cell = (long)((y - g.H / 2 + g.Gy) % g.Gy) * g.Gx + (x - g.W / 2 + g.Gx) % g.Gx;
    double2 acc = make_double2(0.0, 0.0);
    for (int c = 0; c < g.C; ++c)
      acc = dadd(acc, dmulc(up(b1[c * HW + p]), grid[(t * g.C + c) * G2 + cell]));
    out[i] = dadd(out[i], dscale(acc, g.dxi[x] * g.dyi[y]));
  }
}

// ---------------------------------------------------------------- reductions
// partial[b] = (sum Re(conj(a) b), sum |a|^2) over block b's strided range;
// the host adds the partials in block order.
__global__ __launch_bounds__(kT64) void k_dot64(const double2 *__restrict__ a,
                                                const double2 *__restrict__ b, long n,
                                                double *__restrict__ partial) {
  __shared__ double sa[kT64], sb[kT64];
  double da = 0.0, db = 0.0;
  for (long i = (long)blockIdx.x * kT64 + threadIdx.x; i < n; i += (long)gridDim.x * kT64) {
    const double2 p = a[i], q = b[i];
    da += p.x * q.x + p.y * q.y;
    db += p.x * p.x + p.y * p.y;
  }
  sa[threadIdx.x] = da;
  sb[threadIdx.x] = db;
  __syncthreads();
  for (int st = kT64 / 2; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st) {
      sa[threadIdx.x] += sa[threadIdx.x + st];
      sb[threadIdx.x] += sb[threadIdx.x + st];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    partial[2 * blockIdx.x] = sa[0];
    partial[2 * blockIdx.x + 1] = sb[0];
  }
}
__global__ void k_scale64(double2 *x, const double2 *y, double s, long n) {
  for (long i = (long)blockIdx.x * kT64 + threadIdx.x; i < n; i += (long)gridDim.x * kT64)
    x[i] = dscale(y[i], s);
}

// (Re<a, b>, |a|^2)
std::pair<double, double> dot64(const double2 *a, const double2 *b, long n, Fp64Scratch &ws,
                                hipStream_t s) {
  const int nb = std::min(blocks_of(n), 1024);
  ws.ensure_red(sizeof(double) * 2 * nb);
  hipLaunchKernelGGL(k_dot64, dim3(nb), dim3(kT64), 0, s, a, b, n, (double *)ws.red);
  AV_LAUNCH_CHECK();
  std::vector<double> h(2 * nb);
  AV_HIP(hipMemcpyAsync(h.data(), ws.red, sizeof(double) * h.size(), hipMemcpyDeviceToHost, s));
  AV_HIP(hipStreamSynchronize(s));
  double da = 0.0, db = 0.0;
  for (int k = 0; k < nb; ++k) {
    da += h[2 * k];
    db += h[2 * k + 1];
  }
  return {da, db};
}

}  // namespace

Fp64Scratch::~Fp64Scratch() {
  if (k) (void)hipFree(k);
  if (red) (void)hipFree(red);
}
void Fp64Scratch::ensure_k(size_t bytes) {
  if (bytes <= k_bytes && k) return;
  if (k) (void)hipFree(k);
  k = nullptr;
  AV_HIP(hipMalloc(&k, bytes));
  k_bytes = bytes;
}
void Fp64Scratch::ensure_red(size_t bytes) {
  if (bytes <= red_bytes && red) return;
  if (red) (void)hipFree(red);
  red = nullptr;
  AV_HIP(hipMalloc(&red, bytes));
  red_bytes = bytes;
}

void cart_khk_fp64(const CartGeom &g, const double2 *u, double2 *out, const float2 *b1,
                   Fp64Scratch &ws, hipStream_t s) {
  const long HW = (long)g.H * g.W, planes = (long)g.T * g.C, n = planes * HW;
  ws.ensure_k(sizeof(double2) * (size_t)n);
  double2 *z = (double2 *)ws.k;
  hipLaunchKernelGGL(k_expand64, dim3(blocks_of(n)), dim3(kT64), 0, s, u, b1, z, g.C, HW, n);
  AV_LAUNCH_CHECK();
  fft64_2d(z, g.H, g.W, planes, g.centered, +1.0, s);  // K: inverse unitary transform
  hipLaunchKernelGGL(k_mask2_64, dim3(blocks_of(n)), dim3(kT64), 0, s, g, z, n);
  AV_LAUNCH_CHECK();
  fft64_2d(z, g.H, g.W, planes, g.centered, -1.0, s);  // K^H: forward unitary transform
  hipLaunchKernelGGL(k_coil_sum64, dim3(blocks_of((long)g.T * HW)), dim3(kT64), 0, s, z, b1, out,
                     g.C, HW, (long)g.T * HW);
  AV_LAUNCH_CHECK();
}

void nc_khk_fp64(const Nc64 &g, const double2 *u, double2 *out, const float2 *b1,
                 Fp64Scratch &ws, hipStream_t s) {
  const long planes = (long)g.T * g.C, G2 = (long)g.Gx * g.Gy, ng = planes * G2;
  const long nz = planes * g.M;
  // grid planes, then the k-space samples after them in the same buffer
  ws.ensure_k(sizeof(double2) * (size_t)(ng + nz));
  double2 *grid = (double2 *)ws.k, *z = grid + ng;
  hipLaunchKernelGGL(k_nc64_fill, dim3(blocks_of(ng)), dim3(kT64), 0, s, g, u, b1, grid, ng);
  AV_LAUNCH_CHECK();
  // K: e^{-i} transforms of the grid (unscaled; 1 / sqrt(HW) is in the weights)
  fft64_axis(grid, g.Gx, 1, planes * g.Gy, planes * g.Gy, g.Gx, 0, 0, -1.0, 1.0, s);
  fft64_axis(grid, g.Gy, g.Gx, planes * g.Gx, g.Gx, 1, G2, 0, -1.0, 1.0, s);
  hipLaunchKernelGGL(k_nc64_interp, dim3(blocks_of(nz)), dim3(kT64), 0, s, g, grid, z, nz);
  AV_LAUNCH_CHECK();
  // K^H: spreading (gather per cell), e^{+i} transforms, crop and coil sum
  hipLaunchKernelGGL(k_nc64_spread, dim3(blocks_of(ng)), dim3(kT64), 0, s, g, z, grid, ng);
  AV_LAUNCH_CHECK();
  fft64_axis(grid, g.Gx, 1, planes * g.Gy, planes * g.Gy, g.Gx, 0, 0, +1.0, 1.0, s);
  fft64_axis(grid, g.Gy, g.Gx, planes * g.Gx, g.Gx, 1, G2, 0, +1.0, 1.0, s);
  const long N = (long)g.T * g.H * g.W;
  hipLaunchKernelGGL(k_nc64_crop, dim3(blocks_of(N)), dim3(kT64), 0, s, g, grid, b1, out, N);
  AV_LAUNCH_CHECK();
}

void cart3_khk_fp64(int W, int H, int D, int C, const float *mask, const double2 *u, double2 *out,
                    const float2 *b1, Fp64Scratch &ws, hipStream_t s) {
  const long HW = (long)H * W, N = HW * D, n = N * C;
  ws.ensure_k(sizeof(double2) * (size_t)n);
  double2 *z = (double2 *)ws.k;
  hipLaunchKernelGGL(k3_expand64, dim3(blocks_of(n)), dim3(kT64), 0, s, u, b1, z, N, n);
  AV_LAUNCH_CHECK();
  for (int pass = 0; pass < 2; ++pass) {
    const double sign = pass == 0 ? +1.0 : -1.0;  // K: e^{+i}; K^H: e^{-i}
    fft64_axis(z, W, 1, (long)C * D * H, (long)C * D * H, W, 0, 0, sign, 1.0, s);
    fft64_axis(z, H, W, (long)C * D * W, W, 1, HW, 0, sign, 1.0, s);
    fft64_axis(z, D, HW, (long)C * HW, HW, 1, N, 0, sign, 1.0 / std::sqrt((double)N), s);
    if (pass == 0) {
      hipLaunchKernelGGL(k3_mask2_64, dim3(blocks_of(n)), dim3(kT64), 0, s, mask, z, N, n);
      AV_LAUNCH_CHECK();
    }
  }
  hipLaunchKernelGGL(k3_coil_sum64, dim3(blocks_of(N)), dim3(kT64), 0, s, z, b1, out, C, N);
  AV_LAUNCH_CHECK();
}

double operator_norm_fp64(int method, int W, int H, int T, const double d6[6],
                          const Operator &op, const float2 *b1, int iterations, hipStream_t s) {
  AV_CHECK(iterations >= 1, "power iteration needs at least one step");
  const G64 g{W, H, T, (long)W * H * T};
  const long N = g.N;
  const int P = primal_components(method), D = dual_components(method);
  const W3 w1{1.0 / d6[0], 1.0 / d6[1], 1.0 / d6[2]}, w2{1.0 / d6[3], 1.0 / d6[4], 1.0 / d6[5]};
  DevBuf bx, by, bd;
  bx.ensure(sizeof(double2) * (size_t)P * N);
  by.ensure(sizeof(double2) * (size_t)P * N);
  bd.ensure(sizeof(double2) * (size_t)D * N);
  double2 *x = bx.as<double2>(), *y = by.as<double2>(), *dual = bd.as<double2>();
  Fp64Scratch ws;
  const int nb = blocks_of(N), nbp = blocks_of((long)P * N);
  auto apply_M = [&](const double2 *in, double2 *out) {
    switch (method) {
      case kTV:
        hipLaunchKernelGGL(k_reg_K64<kTV>, dim3(nb), dim3(kT64), 0, s, g, w1, w2, in, dual);
        hipLaunchKernelGGL(k_reg_KT64<kTV>, dim3(nb), dim3(kT64), 0, s, g, w1, w2, dual, out);
        break;
      case kTGV2:
        hipLaunchKernelGGL(k_reg_K64<kTGV2>, dim3(nb), dim3(kT64), 0, s, g, w1, w2, in, dual);
        hipLaunchKernelGGL(k_reg_KT64<kTGV2>, dim3(nb), dim3(kT64), 0, s, g, w1, w2, dual, out);
        break;
      case kICTGV2:
        hipLaunchKernelGGL(k_reg_K64<kICTGV2>, dim3(nb), dim3(kT64), 0, s, g, w1, w2, in, dual);
        hipLaunchKernelGGL(k_reg_KT64<kICTGV2>, dim3(nb), dim3(kT64), 0, s, g, w1, w2, dual, out);
        break;
      default:
        hipLaunchKernelGGL(k_reg_K64<kICTV>, dim3(nb), dim3(kT64), 0, s, g, w1, w2, in, dual);
        hipLaunchKernelGGL(k_reg_KT64<kICTV>, dim3(nb), dim3(kT64), 0, s, g, w1, w2, dual, out);
        break;
    }
    AV_LAUNCH_CHECK();
    AV_CHECK(op.khk_fp64(in, out, b1, ws, s), "no FP64 operator-norm path for this data operator");
  };
  // x1 = start; y1 = M x1; then (x1 = y1 / |y1|, y1 = M x1, l = <y1, x1>)
  // `iterations` times (matlab/getopnorm.m:4-13)
  hipLaunchKernelGGL(k_start, dim3(nbp), dim3(kT64), 0, s, x, (long)P * N);
  AV_LAUNCH_CHECK();
  apply_M(x, y);
  double l = 0.0;
  for (int it = 0; it < iterations; ++it) {
    const double n2 = dot64(y, y, (long)P * N, ws, s).second;
    const double inv = n2 > 0.0 ? 1.0 / std::sqrt(n2) : 1.0;
    hipLaunchKernelGGL(k_scale64, dim3(nbp), dim3(kT64), 0, s, x, y, inv, (long)P * N);
    AV_LAUNCH_CHECK();
    apply_M(x, y);
    l = dot64(y, x, (long)P * N, ws, s).first;
  }
  AV_CHECK(std::isfinite(l), "power iteration produced a non-finite value");
  return std::sqrt(std::fabs(l));
}

}  // namespace avionic
