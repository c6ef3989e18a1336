// Device-side building blocks of the regulariser kernels: voxel indexing,
// stencils with the reference's boundary conventions, projections, block
// reductions, and the per-voxel bodies of the dual / primal / step updates.
//   forward difference  d+ u_i = u_{i+1} - u_i, 0 at the last index
//       (matlab/get_sp_fdif.m:8-22; AGILE diff3 at CUDA/src/utils.cpp:104-117)
//   backward difference d- = -(d+)^T: u_0 first, u_i - u_{i-1} inside,
//       -u_{n-2} last (matlab/get_sp_bdif.m:8-24, sym_bgrad_3_3.m:17-27)
#pragma once

#include "reg.hpp"

namespace avionic {

constexpr int RT = kRedThreads;

struct Vox {
  int x, y, t;
  long i;
};

__device__ inline Vox vox_of(const Grid3 &g, long i) {
  Vox v;
  v.i = i;
  v.x = (int)(i % g.W);
  const long r = i / g.W;
  v.y = (int)(r % g.H);
  v.t = (int)(r / g.H);
  return v;
}

// Voxel blocks: each frame is cut into tiles of RT consecutive voxels.
// Frame-major order (tile lb % tiles of frame lb / tiles) streams memory
// linearly; the interleaved order (tile lb / T of frame lb % T) has
// consecutive blocks work on the same tile of consecutive frames, so the
// t-neighbours one block reads are the centre voxels the next block reads.
// Measured on MI355X: the dual and primal updates are faster frame-major, the
// step-rule reduction (which reads two primal states) interleaved.
// Returns false past the end of the frame.
__host__ __device__ inline long tiles_per_frame(const Grid3 &g) {
  return ((long)g.W * g.H + RT - 1) / RT;
}
template <bool INTERLEAVE>
__device__ inline bool vox_of_block(const Grid3 &g, long lb, int tid, Vox &v) {
  const long HW = (long)g.W * g.H;
  const long tiles = (HW + RT - 1) / RT;
  const int t = (int)(INTERLEAVE ? lb % g.T : lb / tiles);
  const long p = (INTERLEAVE ? lb / g.T : lb % tiles) * RT + tid;
  if (p >= HW) return false;
  v.t = t;
  v.y = (int)(p / g.W);
  v.x = (int)(p % g.W);
  v.i = (long)t * HW + p;
  return true;
}

struct PlainAcc {
  const float2 *p;
  __device__ float2 operator()(long j) const { return p[j]; }
};
template <typename DT>
struct DualAcc {
  const DT *p;
  float r;  // projection radius of the component group read through it
  __device__ float2 operator()(long j) const { return DualIO<DT>::load(p, j, r); }
};
// a[j] - b[j]
struct DiffAcc {
  const float2 *a, *b;
  __device__ float2 operator()(long j) const { return csub(a[j], b[j]); }
};
// a[j] - a[j + off]
struct SubAcc {
  const float2 *a;
  long off;
  __device__ float2 operator()(long j) const { return csub(a[j], a[j + off]); }
};
// (a[j] - b[j]) - (a[j + off] - b[j + off])
struct DiffSubAcc {
  const float2 *a, *b;
  long off;
  __device__ float2 operator()(long j) const {
    return csub(csub(a[j], b[j]), csub(a[j + off], b[j + off]));
  }
};

template <class A>
__device__ inline float2 fdiff(const A &a, long i, long s, int pos, int len) {
  return (pos < len - 1) ? csub(a(i + s), a(i)) : czero();
}
template <class A>
__device__ inline float2 bdiff(const A &a, long i, long s, int pos, int len) {
  if (len == 1) return czero();
  if (pos == 0) return a(i);
  if (pos == len - 1) return cscale(a(i - s), -1.f);
  return csub(a(i), a(i - s));
}

struct Inv3 {
  float x, y, t;
};

// grad of the scalar field at component offset o.
template <class A>
__device__ inline void grad3(const A &a, long o, const Vox &v, const Grid3 &g, Inv3 d, float2 *out) {
  const long i = o + v.i, sy = g.W, st = (long)g.W * g.H;
  out[0] = cscale(fdiff(a, i, 1, v.x, g.W), d.x);
  out[1] = cscale(fdiff(a, i, sy, v.y, g.H), d.y);
  out[2] = cscale(fdiff(a, i, st, v.t, g.T), d.t);
}

// symmetrised gradient (xx, yy, tt, xy, xt, yt) of the 3-vector at o0,o1,o2.
template <class A>
__device__ inline void symgrad6(const A &a, long o0, long o1, long o2, const Vox &v, const Grid3 &g,
                                Inv3 d, float2 *e) {
  const long sy = g.W, st = (long)g.W * g.H;
  const long i0 = o0 + v.i, i1 = o1 + v.i, i2 = o2 + v.i;
  e[0] = cscale(bdiff(a, i0, 1, v.x, g.W), d.x);
  e[1] = cscale(bdiff(a, i1, sy, v.y, g.H), d.y);
  e[2] = cscale(bdiff(a, i2, st, v.t, g.T), d.t);
  e[3] = cscale(cadd(cscale(bdiff(a, i0, sy, v.y, g.H), d.y), cscale(bdiff(a, i1, 1, v.x, g.W), d.x)), 0.5f);
  e[4] = cscale(cadd(cscale(bdiff(a, i0, st, v.t, g.T), d.t), cscale(bdiff(a, i2, 1, v.x, g.W), d.x)), 0.5f);
  e[5] = cscale(cadd(cscale(bdiff(a, i1, st, v.t, g.T), d.t), cscale(bdiff(a, i2, sy, v.y, g.H), d.y)), 0.5f);
}

// div of the 3-vector at o0,o1,o2 (= -grad^T).
template <class A>
__device__ inline float2 div3(const A &a, long o0, long o1, long o2, const Vox &v, const Grid3 &g,
                              Inv3 d) {
  const long sy = g.W, st = (long)g.W * g.H;
  return cadd(cadd(cscale(bdiff(a, o0 + v.i, 1, v.x, g.W), d.x),
                   cscale(bdiff(a, o1 + v.i, sy, v.y, g.H), d.y)),
              cscale(bdiff(a, o2 + v.i, st, v.t, g.T), d.t));
}

// symmetric divergence of the 6-vector starting at component offset o
// (components o + k N), k = xx, yy, tt, xy, xt, yt.
template <class A>
__device__ inline void symdiv3(const A &a, long o, long N, const Vox &v, const Grid3 &g, Inv3 d,
                               float2 *out) {
  const long sy = g.W, st = (long)g.W * g.H;
  const long q0 = o + v.i, q1 = q0 + N, q2 = q0 + 2 * N, q3 = q0 + 3 * N, q4 = q0 + 4 * N,
             q5 = q0 + 5 * N;
  out[0] = cadd(cadd(cscale(fdiff(a, q0, 1, v.x, g.W), d.x), cscale(fdiff(a, q3, sy, v.y, g.H), d.y)),
                cscale(fdiff(a, q4, st, v.t, g.T), d.t));
  out[1] = cadd(cadd(cscale(fdiff(a, q3, 1, v.x, g.W), d.x), cscale(fdiff(a, q1, sy, v.y, g.H), d.y)),
                cscale(fdiff(a, q5, st, v.t, g.T), d.t));
  out[2] = cadd(cadd(cscale(fdiff(a, q4, 1, v.x, g.W), d.x), cscale(fdiff(a, q5, sy, v.y, g.H), d.y)),
                cscale(fdiff(a, q2, st, v.t, g.T), d.t));
}

__device__ inline void proj3(float2 *p, float r) {
  const float n = sqrtf(cabs2(p[0]) + cabs2(p[1]) + cabs2(p[2]));
  const float s = fmaxf(1.f, n / r);
  const float is = 1.f / s;
  p[0] = cscale(p[0], is); p[1] = cscale(p[1], is); p[2] = cscale(p[2], is);
}
__device__ inline void proj6(float2 *q, float r) {
  const float n2 = cabs2(q[0]) + cabs2(q[1]) + cabs2(q[2]) +
                   2.f * (cabs2(q[3]) + cabs2(q[4]) + cabs2(q[5]));
  const float s = fmaxf(1.f, sqrtf(n2) / r);
  const float is = 1.f / s;
#pragma unroll
  for (int k = 0; k < 6; ++k) q[k] = cscale(q[k], is);
}

__device__ inline float norm3f(const float2 *p) { return sqrtf(cabs2(p[0]) + cabs2(p[1]) + cabs2(p[2])); }
__device__ inline float norm6f(const float2 *q) {
  return sqrtf(cabs2(q[0]) + cabs2(q[1]) + cabs2(q[2]) + 2.f * (cabs2(q[3]) + cabs2(q[4]) + cabs2(q[5])));
}

// Block-wide double reduction of two values; thread 0 writes them.
__device__ inline void block_reduce2(double a, double b, double *out2) {
  __shared__ double ra[RT], rb[RT];
  const int tid = threadIdx.x;
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
    out2[0] = ra[0];
    out2[1] = rb[0];
  }
}


// ---------------------------------------------------------------------------
// Per-voxel bodies.

// Dual update at voxel v: y += sigma K_reg(ext), then the projections.
// TV: tv.cpp:140-150; TGV2: tgv2.cpp:238-259; ICTGV2: ictgv2.cpp:338-381.
template <int M, typename DT>
__device__ inline void dual_voxel(const Grid3 &g, const RegParams &rp, const float2 *__restrict__ ext,
                                  DT *__restrict__ y, float sig, const Vox &v) {
  const long N = g.N, i = v.i;
  const Inv3 d1{rp.idx, rp.idy, rp.idt}, d2{rp.idx2, rp.idy2, rp.idt2};
  PlainAcc E{ext};
  float2 gr[3], e[6], p[6];
  if (M == kTV) {
    grad3(E, 0, v, g, d1, gr);
#pragma unroll
    for (int c = 0; c < 3; ++c) p[c] = cadd(DualIO<DT>::load(y, c * N + i, rp.r1), cscale(gr[c], sig));
    proj3(p, rp.r1);
#pragma unroll
    for (int c = 0; c < 3; ++c) DualIO<DT>::store(y, c * N + i, p[c], rp.r1);
    return;
  }
  if (M == kTGV2) {
    grad3(E, 0, v, g, d1, gr);
  } else {
    SubAcc UW{ext, 4 * N};  // u - w
    grad3(UW, 0, v, g, d1, gr);
  }
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const float2 t = csub(gr[c], ext[(1 + c) * N + i]);
    p[c] = cadd(DualIO<DT>::load(y, c * N + i, rp.r1), cscale(t, sig));
  }
  proj3(p, rp.r1);
#pragma unroll
  for (int c = 0; c < 3; ++c) DualIO<DT>::store(y, c * N + i, p[c], rp.r1);
  symgrad6(E, N, 2 * N, 3 * N, v, g, d1, e);
#pragma unroll
  for (int k = 0; k < 6; ++k) p[k] = cadd(DualIO<DT>::load(y, (3 + k) * N + i, rp.r2), cscale(e[k], sig));
  proj6(p, rp.r2);
#pragma unroll
  for (int k = 0; k < 6; ++k) DualIO<DT>::store(y, (3 + k) * N + i, p[k], rp.r2);
  if (M == kICTGV2) {
    grad3(E, 4 * N, v, g, d2, gr);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const float2 t = csub(gr[c], ext[(5 + c) * N + i]);
      p[c] = cadd(DualIO<DT>::load(y, (9 + c) * N + i, rp.r3), cscale(t, sig));
    }
    proj3(p, rp.r3);
#pragma unroll
    for (int c = 0; c < 3; ++c) DualIO<DT>::store(y, (9 + c) * N + i, p[c], rp.r3);
    symgrad6(E, 5 * N, 6 * N, 7 * N, v, g, d2, e);
#pragma unroll
    for (int k = 0; k < 6; ++k) p[k] = cadd(DualIO<DT>::load(y, (12 + k) * N + i, rp.r4), cscale(e[k], sig));
    proj6(p, rp.r4);
#pragma unroll
    for (int k = 0; k < 6; ++k) DualIO<DT>::store(y, (12 + k) * N + i, p[k], rp.r4);
  }
}

// Primal update + extragradient at voxel v.
// TV: tv.cpp:155-169; TGV2: tgv2.cpp:268-303; ICTGV2: ictgv2.cpp:388-445.
template <int M, typename DT>
__device__ inline void primal_voxel(const Grid3 &g, const RegParams &rp, float2 *__restrict__ x,
                                    float2 *__restrict__ ext, const DT *__restrict__ y,
                                    const float2 *__restrict__ khz, float tau, const Vox &v) {
  const long N = g.N, i = v.i;
  const Inv3 d1{rp.idx, rp.idy, rp.idt}, d2{rp.idx2, rp.idy2, rp.idt2};
  const DualAcc<DT> Y1{y, rp.r1}, Y2{y, rp.r2}, Y3{y, rp.r3}, Y4{y, rp.r4};
  float2 upd[8];
  const float2 dv1 = div3(Y1, 0, N, 2 * N, v, g, d1);
  upd[0] = csub(x[i], cscale(csub(khz[i], dv1), tau));
  int P = 1;
  if (M != kTV) {
    float2 sd[3];
    symdiv3(Y2, 3 * N, N, v, g, d1, sd);
#pragma unroll
    for (int c = 0; c < 3; ++c)
      upd[1 + c] = cadd(x[(1 + c) * N + i], cscale(cadd(Y1(c * N + i), sd[c]), tau));
    P = 4;
    if (M == kICTGV2) {
      const float2 dv3 = div3(Y3, 9 * N, 10 * N, 11 * N, v, g, d2);
      upd[4] = csub(x[4 * N + i], cscale(csub(dv1, dv3), tau));
      symdiv3(Y4, 12 * N, N, v, g, d2, sd);
#pragma unroll
      for (int c = 0; c < 3; ++c)
        upd[5 + c] = cadd(x[(5 + c) * N + i], cscale(cadd(Y3((9 + c) * N + i), sd[c]), tau));
      P = 8;
    }
  }
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    if (c < P) {
      const long j = c * N + i;
      const float2 xo = x[j];
      ext[j] = csub(cscale(upd[c], 2.f), xo);
      x[j] = upd[c];
    }
  }
}

// Step-rule terms at voxel v: weighted |K_reg (ext - x)|^2 and |ext - x|^2.
// tv.cpp:79-91, tgv2.cpp:89-122, ictgv2.cpp:97-155.
template <int M>
__device__ inline void step_voxel(const Grid3 &g, const RegParams &rp, const float2 *__restrict__ ext,
                                  const float2 *__restrict__ x, const Vox &v, double &ak,
                                  double &ax) {
  const long N = g.N, i = v.i;
  const Inv3 d1{rp.idx, rp.idy, rp.idt}, d2{rp.idx2, rp.idy2, rp.idt2};
  DiffAcc D{ext, x};
  float2 gr[3], e[6];
  const int P = (M == kTV) ? 1 : (M == kTGV2 ? 4 : 8);
  for (int c = 0; c < P; ++c) ax += (double)cabs2(D(c * N + i));
  if (M == kTV) {
    grad3(D, 0, v, g, d1, gr);
    for (int c = 0; c < 3; ++c) ak += (double)cabs2(gr[c]);
    return;
  }
  if (M == kTGV2) {
    grad3(D, 0, v, g, d1, gr);
  } else {
    DiffSubAcc DUW{ext, x, 4 * N};
    grad3(DUW, 0, v, g, d1, gr);
  }
  for (int c = 0; c < 3; ++c) ak += (double)cabs2(csub(gr[c], D((1 + c) * N + i)));
  symgrad6(D, N, 2 * N, 3 * N, v, g, d1, e);
  for (int k = 0; k < 6; ++k) ak += (k < 3 ? 1.0 : 2.0) * (double)cabs2(e[k]);
  if (M == kICTGV2) {
    grad3(D, 4 * N, v, g, d2, gr);
    for (int c = 0; c < 3; ++c) ak += (double)cabs2(csub(gr[c], D((5 + c) * N + i)));
    symgrad6(D, 5 * N, 6 * N, 7 * N, v, g, d2, e);
    for (int k = 0; k < 6; ++k) ak += (k < 3 ? 1.0 : 2.0) * (double)cabs2(e[k]);
  }
}

// The step rule (pd_recon.cpp:40-61) on FP64 reductions.
__device__ inline void apply_step_rule(DevScalars *sc, double nK2, double nx2) {
  sc->nKx2 = nK2;
  sc->nx2 = nx2;
  if (!isfinite(nK2) || !isfinite(nx2)) sc->diverged = 1;
  const double nKx = sqrt(nK2), nx = sqrt(nx2);
  const double tmp = nx / nKx, theta = 0.95;
  double s = sc->sigma, t = sc->tau;
  if (s * t > tmp * tmp) {
    if (theta * theta * s * t < tmp * tmp) {
      s *= theta;
      t *= theta;
    } else {
      s = tmp * sc->ratio;
      t = tmp / sc->ratio;
    }
  }
  sc->sigma = s;
  sc->tau = t;
}

}  // namespace avionic
