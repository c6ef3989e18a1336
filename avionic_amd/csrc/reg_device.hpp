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

// A voxel and its linear index within one component plane stack
// (t HW + y W + x).  Component c of a state lives at base + c N; kernels add
// c N to the (uniform) base pointer and index with the 32-bit i, so every
// access is a scalar base plus a 32-bit lane offset (N < 2^29 is checked on
// the host, which keeps byte offsets within 32 bits).
struct Vox {
  int x, y, t;
  unsigned i;
};

__device__ inline Vox vox_of(const Grid3 &g, long i) {
  Vox v;
  v.i = (unsigned)i;
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
  const int HW = g.W * g.H;
  const int tiles = (HW + RT - 1) / RT;
  const int t = (int)(INTERLEAVE ? lb % g.T : lb / tiles);
  const int p = (int)(INTERLEAVE ? lb / g.T : lb % tiles) * RT + tid;
  if (p >= HW) return false;
  v.t = t;
  v.y = p / g.W;
  v.x = p - v.y * g.W;
  v.i = (unsigned)t * (unsigned)HW + (unsigned)p;
  return true;
}

// Accessors: a(j) reads element j of a component; a.at(o) is the accessor of
// the component at element offset o.
struct PlainAcc {
  const float2 *p;
  __device__ PlainAcc at(long o) const { return {p + o}; }
  __device__ float2 operator()(unsigned j) const { return gld(p, j); }
};
template <typename DT>
struct DualAcc {
  const DT *p;
  float r;  // projection radius of the component group read through it
  __device__ DualAcc at(long o) const { return {p + o, r}; }
  __device__ float2 operator()(unsigned j) const { return DualIO<DT>::load(p, j, r); }
};
// a[j] - b[j]
struct DiffAcc {
  const float2 *a, *b;
  __device__ DiffAcc at(long o) const { return {a + o, b + o}; }
  __device__ float2 operator()(unsigned j) const { return csub(gld(a, j), gld(b, j)); }
};
// (a[j] - b[j]) - (c[j] - d[j])
struct Diff2Acc {
  const float2 *a, *b, *c, *d;
  __device__ Diff2Acc at(long o) const { return {a + o, b + o, c + o, d + o}; }
  __device__ float2 operator()(unsigned j) const {
    return csub(csub(gld(a, j), gld(b, j)), csub(gld(c, j), gld(d, j)));
  }
};

// Branch-free, so that every load of a voxel is issued unconditionally (a
// load under a divergent branch whose value is used after it gets its own
// wait, which serialises the voxel's loads).  The neighbour index is clamped
// to i at the boundary: the forward difference is then exactly 0 there, and
// the backward difference c mc - p mp with the 0/1 factors mc = [pos != len-1],
// mp = [pos != 0] gives c, -p and c - p exactly (finite values).
template <class A>
__device__ inline float2 fdiff(const A &a, unsigned i, unsigned s, int pos, int len) {
  return csub(a(pos < len - 1 ? i + s : i), a(i));
}
template <class A>
__device__ inline float2 bdiff(const A &a, unsigned i, unsigned s, int pos, int len) {
  const float mc = pos != len - 1 ? 1.f : 0.f, mp = pos != 0 ? 1.f : 0.f;
  return csub(cscale(a(i), mc), cscale(a(pos > 0 ? i - s : i), mp));
}

struct Inv3 {
  float x, y, t;
};

// grad of the scalar field at component offset o.
template <class A>
__device__ inline void grad3(const A &a0, long o, const Vox &v, const Grid3 &g, Inv3 d, float2 *out) {
  const A a = a0.at(o);
  const unsigned i = v.i, sy = g.W, st = (unsigned)g.W * g.H;
  out[0] = cscale(fdiff(a, i, 1u, v.x, g.W), d.x);
  out[1] = cscale(fdiff(a, i, sy, v.y, g.H), d.y);
  out[2] = cscale(fdiff(a, i, st, v.t, g.T), d.t);
}

// symmetrised gradient (xx, yy, tt, xy, xt, yt) of the 3-vector at o0,o1,o2.
template <class A>
__device__ inline void symgrad6(const A &a, long o0, long o1, long o2, const Vox &v, const Grid3 &g,
                                Inv3 d, float2 *e) {
  const unsigned i = v.i, sy = g.W, st = (unsigned)g.W * g.H;
  const A a0 = a.at(o0), a1 = a.at(o1), a2 = a.at(o2);
  e[0] = cscale(bdiff(a0, i, 1u, v.x, g.W), d.x);
  e[1] = cscale(bdiff(a1, i, sy, v.y, g.H), d.y);
  e[2] = cscale(bdiff(a2, i, st, v.t, g.T), d.t);
  e[3] = cscale(cadd(cscale(bdiff(a0, i, sy, v.y, g.H), d.y), cscale(bdiff(a1, i, 1u, v.x, g.W), d.x)), 0.5f);
  e[4] = cscale(cadd(cscale(bdiff(a0, i, st, v.t, g.T), d.t), cscale(bdiff(a2, i, 1u, v.x, g.W), d.x)), 0.5f);
  e[5] = cscale(cadd(cscale(bdiff(a1, i, st, v.t, g.T), d.t), cscale(bdiff(a2, i, sy, v.y, g.H), d.y)), 0.5f);
}

// div of the 3-vector at o0,o1,o2 (= -grad^T).
template <class A>
__device__ inline float2 div3(const A &a, long o0, long o1, long o2, const Vox &v, const Grid3 &g,
                              Inv3 d) {
  const unsigned i = v.i, sy = g.W, st = (unsigned)g.W * g.H;
  return cadd(cadd(cscale(bdiff(a.at(o0), i, 1u, v.x, g.W), d.x),
                   cscale(bdiff(a.at(o1), i, sy, v.y, g.H), d.y)),
              cscale(bdiff(a.at(o2), i, st, v.t, g.T), d.t));
}

// symmetric divergence of the 6-vector starting at component offset o
// (components o + k N), k = xx, yy, tt, xy, xt, yt.
template <class A>
__device__ inline void symdiv3(const A &a, long o, long N, const Vox &v, const Grid3 &g, Inv3 d,
                               float2 *out) {
  const unsigned i = v.i, sy = g.W, st = (unsigned)g.W * g.H;
  const A q0 = a.at(o), q1 = a.at(o + N), q2 = a.at(o + 2 * N), q3 = a.at(o + 3 * N),
          q4 = a.at(o + 4 * N), q5 = a.at(o + 5 * N);
  out[0] = cadd(cadd(cscale(fdiff(q0, i, 1u, v.x, g.W), d.x), cscale(fdiff(q3, i, sy, v.y, g.H), d.y)),
                cscale(fdiff(q4, i, st, v.t, g.T), d.t));
  out[1] = cadd(cadd(cscale(fdiff(q3, i, 1u, v.x, g.W), d.x), cscale(fdiff(q1, i, sy, v.y, g.H), d.y)),
                cscale(fdiff(q5, i, st, v.t, g.T), d.t));
  out[2] = cadd(cadd(cscale(fdiff(q4, i, 1u, v.x, g.W), d.x), cscale(fdiff(q5, i, sy, v.y, g.H), d.y)),
                cscale(fdiff(q2, i, st, v.t, g.T), d.t));
}

// p / max(1, |p| / r) (norm_3.m, norm_6.m; utils.cpp:1297-1351) evaluated as
// p * min(1, r / |p|): one reciprocal square root instead of a square root and
// two divisions (the scale differs from the quotient form by at most a few
// ulp).  n2 = |p|^2 with the off-diagonal weight 2 for norm_6.
__device__ inline float proj_scale(float n2, float r) {
  return n2 > r * r ? r * __builtin_amdgcn_rsqf(n2) : 1.f;
}
__device__ inline void proj3(float2 *p, float r) {
  const float is = proj_scale(cabs2(p[0]) + cabs2(p[1]) + cabs2(p[2]), r);
  p[0] = cscale(p[0], is); p[1] = cscale(p[1], is); p[2] = cscale(p[2], is);
}
__device__ inline void proj6(float2 *q, float r) {
  const float n2 = cabs2(q[0]) + cabs2(q[1]) + cabs2(q[2]) +
                   2.f * (cabs2(q[3]) + cabs2(q[4]) + cabs2(q[5]));
  const float is = proj_scale(n2, r);
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
  const long N = g.N;
  const unsigned i = v.i;
  const Inv3 d1{rp.idx, rp.idy, rp.idt}, d2{rp.idx2, rp.idy2, rp.idt2};
  PlainAcc E{ext};
  float2 gr[3], e[6], p[6];
  // components [c0, c0 + n) of y: p += sigma * k, project, store
  auto ascend = [&](int c0, int n, const float2 *k, float r) {
#pragma unroll
    for (int c = 0; c < 6; ++c)
      if (c < n) p[c] = cadd(DualIO<DT>::load_stream(y + (c0 + c) * N, i, r), cscale(k[c], sig));
    if (n == 3) proj3(p, r); else proj6(p, r);
#pragma unroll
    for (int c = 0; c < 6; ++c)
      if (c < n) DualIO<DT>::store(y + (c0 + c) * N, i, p[c], r);
  };
  if (M == kTV) {
    grad3(E, 0, v, g, d1, gr);
    ascend(0, 3, gr, rp.r1);
    return;
  }
  if (M == kICTV) {  // ictv.cpp:268-290: y1 <- grad1(u - w), y3 <- grad2 w
    grad3(DiffAcc{ext, ext + N}, 0, v, g, d1, gr);
    ascend(0, 3, gr, rp.r1);
    grad3(E, N, v, g, d2, gr);
    ascend(3, 3, gr, rp.r3);
    return;
  }
  if (M == kTGV2) {
    grad3(E, 0, v, g, d1, gr);
  } else {
    grad3(DiffAcc{ext, ext + 4 * N}, 0, v, g, d1, gr);  // u - w
  }
#pragma unroll
  for (int c = 0; c < 3; ++c) gr[c] = csub(gr[c], gld(ext + (1 + c) * N, i));
  ascend(0, 3, gr, rp.r1);
  symgrad6(E, N, 2 * N, 3 * N, v, g, d1, e);
  ascend(3, 6, e, rp.r2);
  if (M == kICTGV2) {
    grad3(E, 4 * N, v, g, d2, gr);
#pragma unroll
    for (int c = 0; c < 3; ++c) gr[c] = csub(gr[c], gld(ext + (5 + c) * N, i));
    ascend(9, 3, gr, rp.r3);
    symgrad6(E, 5 * N, 6 * N, 7 * N, v, g, d2, e);
    ascend(12, 6, e, rp.r4);
  }
}

// Primal update + extragradient at voxel v.
// TV: tv.cpp:155-169; TGV2: tgv2.cpp:268-303; ICTGV2: ictgv2.cpp:388-445.
template <int M, typename DT>
__device__ inline void primal_voxel(const Grid3 &g, const RegParams &rp, float2 *__restrict__ x,
                                    float2 *__restrict__ ext, const DT *__restrict__ y,
                                    const float2 *__restrict__ khz, float tau, const Vox &v) {
  const long N = g.N;
  const unsigned i = v.i;
  const Inv3 d1{rp.idx, rp.idy, rp.idt}, d2{rp.idx2, rp.idy2, rp.idt2};
  const DualAcc<DT> Y1{y, rp.r1}, Y2{y, rp.r2}, Y3{y, rp.r3}, Y4{y, rp.r4};
  float2 upd[8];
  const float2 dv1 = div3(Y1, 0, N, 2 * N, v, g, d1);
  upd[0] = csub(gld(x, i), cscale(csub(gld(khz, i), dv1), tau));
  int P = 1;
  if (M == kICTV) {  // ictv.cpp:301-310: w <- w - tau (div1 y1 - div2 y3)
    const float2 dv3 = div3(Y3, 3 * N, 4 * N, 5 * N, v, g, d2);
    upd[1] = csub(gld(x + N, i), cscale(csub(dv1, dv3), tau));
    P = 2;
  } else if (M != kTV) {
    float2 sd[3];
    symdiv3(Y2, 3 * N, N, v, g, d1, sd);
#pragma unroll
    for (int c = 0; c < 3; ++c)
      upd[1 + c] = cadd(gld(x + (1 + c) * N, i), cscale(cadd(Y1.at(c * N)(i), sd[c]), tau));
    P = 4;
    if (M == kICTGV2) {
      const float2 dv3 = div3(Y3, 9 * N, 10 * N, 11 * N, v, g, d2);
      upd[4] = csub(gld(x + 4 * N, i), cscale(csub(dv1, dv3), tau));
      symdiv3(Y4, 12 * N, N, v, g, d2, sd);
#pragma unroll
      for (int c = 0; c < 3; ++c)
        upd[5 + c] = cadd(gld(x + (5 + c) * N, i), cscale(cadd(Y3.at((9 + c) * N)(i), sd[c]), tau));
      P = 8;
    }
  }
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    if (c < P) {
      float2 *xc = x + c * N, *ec = ext + c * N;
      const float2 xo = gld(xc, i);
      gst_stream(ec, i, csub(cscale(upd[c], 2.f), xo));
      gst_stream(xc, i, upd[c]);
    }
  }
}

// Step-rule terms at voxel v: weighted |K_reg (ext - x)|^2 and |ext - x|^2.
// tv.cpp:79-91, tgv2.cpp:89-122, ictgv2.cpp:97-155.
template <int M>
__device__ inline void step_voxel(const Grid3 &g, const RegParams &rp, const float2 *__restrict__ ext,
                                  const float2 *__restrict__ x, const Vox &v, double &ak,
                                  double &ax) {
  const long N = g.N;
  const unsigned i = v.i;
  const Inv3 d1{rp.idx, rp.idy, rp.idt}, d2{rp.idx2, rp.idy2, rp.idt2};
  const DiffAcc D{ext, x};
  float2 gr[3], e[6];
  const int P = (M == kTV) ? 1 : (M == kTGV2 ? 4 : (M == kICTV ? 2 : 8));
  for (int c = 0; c < P; ++c) ax += (double)cabs2(D.at(c * N)(i));
  // ICTV's AdaptStepSize (ictv.cpp:96-131) overwrites its regulariser sum
  // before use: nKx holds the data term alone, so nothing is added here.
  if (M == kICTV) return;
  if (M == kTV) {
    grad3(D, 0, v, g, d1, gr);
    for (int c = 0; c < 3; ++c) ak += (double)cabs2(gr[c]);
    return;
  }
  if (M == kTGV2) {
    grad3(D, 0, v, g, d1, gr);
  } else {
    grad3(Diff2Acc{ext, x, ext + 4 * N, x + 4 * N}, 0, v, g, d1, gr);
  }
  for (int c = 0; c < 3; ++c) ak += (double)cabs2(csub(gr[c], D.at((1 + c) * N)(i)));
  symgrad6(D, N, 2 * N, 3 * N, v, g, d1, e);
  for (int k = 0; k < 6; ++k) ak += (k < 3 ? 1.0 : 2.0) * (double)cabs2(e[k]);
  if (M == kICTGV2) {
    grad3(D, 4 * N, v, g, d2, gr);
    for (int c = 0; c < 3; ++c) ak += (double)cabs2(csub(gr[c], D.at((5 + c) * N)(i)));
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
