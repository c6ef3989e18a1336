// Device-side building blocks of the regulariser kernels: voxel indexing,
// stencils with the reference's boundary conventions, projections, block
// reductions, and the per-voxel bodies of the dual / primal / step updates.
//   forward difference  d+ u_i = u_{i+1} - u_i, 0 at the last index
//       (matlab/get_sp_fdif.m:8-22; AGILE diff3 at CUDA/src/utils.cpp:104-117)
//   backward difference d- = -(d+)^T: u_0 first, u_i - u_{i-1} inside,
//       -u_{n-2} last (matlab/get_sp_bdif.m:8-24, sym_bgrad_3_3.m:17-27)
#pragma once

#include "reg.hpp"

// The arithmetic of the bodies below is not contracted into multiply-adds
// (radd / rsub / rscale / rabs2 and every operation written here): the
// one-voxel and the two-voxel forms of an update then round identically, and
// a reconstruction gives the same bits whichever form runs it.  These kernels
// are bound by memory traffic, not by arithmetic.
#pragma clang fp contract(off)

namespace avionic {

__device__ inline float2 radd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ inline float2 rsub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ inline float2 rscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
__device__ inline float rabs2(float2 a) { return a.x * a.x + a.y * a.y; }
__device__ inline C2 radd(C2 p, C2 q) { return {radd(p.a, q.a), radd(p.b, q.b)}; }
__device__ inline C2 rsub(C2 p, C2 q) { return {rsub(p.a, q.a), rsub(p.b, q.b)}; }
__device__ inline C2 rscale(C2 p, float s) { return {rscale(p.a, s), rscale(p.b, s)}; }

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

// Two voxels per lane: the pair (x, x + 1) of one row with x even (rows of
// even width only, so a pair never straddles two rows and every component
// plane starts 16-byte aligned).  The lane loads and stores both voxels of a
// component with one 16-byte (complex64) or 8-byte (4-byte duals) access;
// the arithmetic is the one-voxel arithmetic applied to each of the two.
struct Vox2 {
  int x, y, t;  // x: the even one of the pair
  unsigned i;
};
constexpr int kPairTile = 2 * RT;  // voxels of one two-voxel block
__host__ __device__ inline long pair_tiles_per_frame(const Grid3 &g) {
  return ((long)g.W * g.H + kPairTile - 1) / kPairTile;
}
// Pair tile lb (frame-major, like vox_of_block<false>); false past the end
// of the frame or of the grid.
__device__ inline bool vox2_of_block(const Grid3 &g, long lb, int tid, Vox2 &v) {
  const int HW = g.W * g.H;
  const int tiles = (HW + kPairTile - 1) / kPairTile;
  const int t = (int)(lb / tiles);
  const int p = (int)(lb % tiles) * kPairTile + 2 * tid;
  if (p >= HW || t >= g.T) return false;
  v.t = t;
  v.y = p / g.W;
  v.x = p - v.y * g.W;
  v.i = (unsigned)t * (unsigned)HW + (unsigned)p;
  return true;
}

// Accessors: a(j) reads element j of a component, a.pair(j) elements j and
// j + 1 (j even); a.at(o) is the accessor of the component at element
// offset o.
struct PlainAcc {
  const float2 *p;
  __device__ PlainAcc at(long o) const { return {p + o}; }
  __device__ float2 operator()(unsigned j) const { return gld(p, j); }
  __device__ C2 pair(unsigned j) const { return gld2(p, j); }
};
template <typename DT>
struct DualAcc {
  const DT *p;
  float r;  // projection radius of the component group read through it
  __device__ DualAcc at(long o) const { return {p + o, r}; }
  __device__ float2 operator()(unsigned j) const { return DualIO<DT>::load(p, j, r); }
  __device__ C2 pair(unsigned j) const { return DualIO<DT>::load2(p, j, r); }
};
// a[j] - b[j]
struct DiffAcc {
  const float2 *a, *b;
  __device__ DiffAcc at(long o) const { return {a + o, b + o}; }
  __device__ float2 operator()(unsigned j) const { return rsub(gld(a, j), gld(b, j)); }
  __device__ C2 pair(unsigned j) const { return rsub(gld2(a, j), gld2(b, j)); }
};
// (a[j] - b[j]) - (c[j] - d[j])
struct Diff2Acc {
  const float2 *a, *b, *c, *d;
  __device__ Diff2Acc at(long o) const { return {a + o, b + o, c + o, d + o}; }
  __device__ float2 operator()(unsigned j) const {
    return rsub(rsub(gld(a, j), gld(b, j)), rsub(gld(c, j), gld(d, j)));
  }
};

// The lane's value of a component: one complex value (Vox) or a pair (Vox2).
template <class V>
struct LaneT {
  typedef float2 type;
};
template <>
struct LaneT<Vox2> {
  typedef C2 type;
};
template <class A>
__device__ inline float2 ldv(const A &a, const Vox &, unsigned j) { return a(j); }
template <class A>
__device__ inline C2 ldv(const A &a, const Vox2 &, unsigned j) { return a.pair(j); }

// Branch-free, so that every load of a voxel is issued unconditionally (a
// load under a divergent branch whose value is used after it gets its own
// wait, which serialises the voxel's loads).  The neighbour index is clamped
// to i at the boundary: the forward difference is then exactly 0 there, and
// the backward difference c mc - p mp with the 0/1 factors mc = [pos != len-1],
// mp = [pos != 0] gives c, -p and c - p exactly (finite values).
template <class A>
__device__ inline float2 fdiff(const A &a, unsigned i, unsigned s, int pos, int len) {
  return rsub(a(pos < len - 1 ? i + s : i), a(i));
}
template <class A>
__device__ inline float2 bdiff(const A &a, unsigned i, unsigned s, int pos, int len) {
  const float mc = pos != len - 1 ? 1.f : 0.f, mp = pos != 0 ? 1.f : 0.f;
  return rsub(rscale(a(i), mc), rscale(a(pos > 0 ? i - s : i), mp));
}

struct Inv3 {
  float x, y, t;
};

// Differences along x (fdx, bdx) and along a stride s of y or t (fds, bds),
// with the boundary rules of fdiff / bdiff.  For a pair the x differences
// need one more element beside the pair: x + 2 (forward) or x - 1 (backward),
// clamped into the row like the one-voxel neighbour; the pair's two results
// are bitwise those of fdiff / bdiff at x and x + 1.
template <class A>
__device__ inline float2 fdx(const A &a, const Vox &v, int W) { return fdiff(a, v.i, 1u, v.x, W); }
template <class A>
__device__ inline C2 fdx(const A &a, const Vox2 &v, int W) {
  const C2 c = a.pair(v.i);
  const float2 n = a(v.x + 2 < W ? v.i + 2 : v.i + 1);  // = c.b at the row end: difference 0
  return {rsub(c.b, c.a), rsub(n, c.b)};
}
template <class A>
__device__ inline float2 bdx(const A &a, const Vox &v, int W) { return bdiff(a, v.i, 1u, v.x, W); }
template <class A>
__device__ inline C2 bdx(const A &a, const Vox2 &v, int W) {
  const C2 c = a.pair(v.i);
  const float2 p = a(v.x > 0 ? v.i - 1 : v.i);
  // the 0/1 factors of both elements as the one-voxel form computes them (the
  // compiler then builds the same operations, including its multiply-add
  // contractions, for each element)
  const float mc0 = v.x != W - 1 ? 1.f : 0.f, mp0 = v.x != 0 ? 1.f : 0.f;
  const float mc1 = v.x + 1 != W - 1 ? 1.f : 0.f, mp1 = v.x + 1 != 0 ? 1.f : 0.f;
  return {rsub(rscale(c.a, mc0), rscale(p, mp0)), rsub(rscale(c.b, mc1), rscale(c.a, mp1))};
}
template <class A, class V>
__device__ inline typename LaneT<V>::type fds(const A &a, const V &v, unsigned s, int pos, int len) {
  return rsub(ldv(a, v, pos < len - 1 ? v.i + s : v.i), ldv(a, v, v.i));
}
template <class A, class V>
__device__ inline typename LaneT<V>::type bds(const A &a, const V &v, unsigned s, int pos, int len) {
  const float mc = pos != len - 1 ? 1.f : 0.f, mp = pos != 0 ? 1.f : 0.f;
  return rsub(rscale(ldv(a, v, v.i), mc), rscale(ldv(a, v, pos > 0 ? v.i - s : v.i), mp));
}

// grad of the scalar field at component offset o.
template <class A, class V, class L>
__device__ inline void grad3(const A &a0, long o, const V &v, const Grid3 &g, Inv3 d, L *out) {
  const A a = a0.at(o);
  const unsigned sy = g.W, st = (unsigned)g.W * g.H;
  out[0] = rscale(fdx(a, v, g.W), d.x);
  out[1] = rscale(fds(a, v, sy, v.y, g.H), d.y);
  out[2] = rscale(fds(a, v, st, v.t, g.T), d.t);
}

// symmetrised gradient (xx, yy, tt, xy, xt, yt) of the 3-vector at o0,o1,o2.
template <class A, class V, class L>
__device__ inline void symgrad6(const A &a, long o0, long o1, long o2, const V &v, const Grid3 &g,
                                Inv3 d, L *e) {
  const unsigned sy = g.W, st = (unsigned)g.W * g.H;
  const A a0 = a.at(o0), a1 = a.at(o1), a2 = a.at(o2);
  e[0] = rscale(bdx(a0, v, g.W), d.x);
  e[1] = rscale(bds(a1, v, sy, v.y, g.H), d.y);
  e[2] = rscale(bds(a2, v, st, v.t, g.T), d.t);
  e[3] = rscale(radd(rscale(bds(a0, v, sy, v.y, g.H), d.y), rscale(bdx(a1, v, g.W), d.x)), 0.5f);
  e[4] = rscale(radd(rscale(bds(a0, v, st, v.t, g.T), d.t), rscale(bdx(a2, v, g.W), d.x)), 0.5f);
  e[5] = rscale(radd(rscale(bds(a1, v, st, v.t, g.T), d.t), rscale(bds(a2, v, sy, v.y, g.H), d.y)), 0.5f);
}

// div of the 3-vector at o0,o1,o2 (= -grad^T).
template <class A, class V>
__device__ inline typename LaneT<V>::type div3(const A &a, long o0, long o1, long o2, const V &v,
                                               const Grid3 &g, Inv3 d) {
  const unsigned sy = g.W, st = (unsigned)g.W * g.H;
  return radd(radd(rscale(bdx(a.at(o0), v, g.W), d.x),
                   rscale(bds(a.at(o1), v, sy, v.y, g.H), d.y)),
              rscale(bds(a.at(o2), v, st, v.t, g.T), d.t));
}

// symmetric divergence of the 6-vector starting at component offset o
// (components o + k N), k = xx, yy, tt, xy, xt, yt.
template <class A, class V, class L>
__device__ inline void symdiv3(const A &a, long o, long N, const V &v, const Grid3 &g, Inv3 d,
                               L *out) {
  const unsigned sy = g.W, st = (unsigned)g.W * g.H;
  const A q0 = a.at(o), q1 = a.at(o + N), q2 = a.at(o + 2 * N), q3 = a.at(o + 3 * N),
          q4 = a.at(o + 4 * N), q5 = a.at(o + 5 * N);
  out[0] = radd(radd(rscale(fdx(q0, v, g.W), d.x), rscale(fds(q3, v, sy, v.y, g.H), d.y)),
                rscale(fds(q4, v, st, v.t, g.T), d.t));
  out[1] = radd(radd(rscale(fdx(q3, v, g.W), d.x), rscale(fds(q1, v, sy, v.y, g.H), d.y)),
                rscale(fds(q5, v, st, v.t, g.T), d.t));
  out[2] = radd(radd(rscale(fdx(q4, v, g.W), d.x), rscale(fds(q5, v, sy, v.y, g.H), d.y)),
                rscale(fds(q2, v, st, v.t, g.T), d.t));
}

// p / max(1, |p| / r) (norm_3.m, norm_6.m; utils.cpp:1297-1351) evaluated as
// p * min(1, r / |p|): one reciprocal square root instead of a square root and
// two divisions (the scale differs from the quotient form by at most a few
// ulp).  n2 = |p|^2 with the off-diagonal weight 2 for norm_6.
__device__ inline float proj_scale(float n2, float r) {
  return n2 > r * r ? r * __builtin_amdgcn_rsqf(n2) : 1.f;
}
__device__ inline void proj3(float2 *p, float r) {
  const float is = proj_scale(rabs2(p[0]) + rabs2(p[1]) + rabs2(p[2]), r);
  p[0] = rscale(p[0], is); p[1] = rscale(p[1], is); p[2] = rscale(p[2], is);
}
__device__ inline void proj6(float2 *q, float r) {
  const float n2 = rabs2(q[0]) + rabs2(q[1]) + rabs2(q[2]) +
                   2.f * (rabs2(q[3]) + rabs2(q[4]) + rabs2(q[5]));
  const float is = proj_scale(n2, r);
#pragma unroll
  for (int k = 0; k < 6; ++k) q[k] = rscale(q[k], is);
}

// Pairs: each voxel of the pair is projected on its own.
__device__ inline void proj3(C2 *p, float r) {
  float2 a[3] = {p[0].a, p[1].a, p[2].a}, b[3] = {p[0].b, p[1].b, p[2].b};
  proj3(a, r);
  proj3(b, r);
#pragma unroll
  for (int k = 0; k < 3; ++k) p[k] = {a[k], b[k]};
}
__device__ inline void proj6(C2 *q, float r) {
  float2 a[6], b[6];
#pragma unroll
  for (int k = 0; k < 6; ++k) { a[k] = q[k].a; b[k] = q[k].b; }
  proj6(a, r);
  proj6(b, r);
#pragma unroll
  for (int k = 0; k < 6; ++k) q[k] = {a[k], b[k]};
}

__device__ inline float norm3f(const float2 *p) { return sqrtf(rabs2(p[0]) + rabs2(p[1]) + rabs2(p[2])); }
__device__ inline float norm6f(const float2 *q) {
  return sqrtf(rabs2(q[0]) + rabs2(q[1]) + rabs2(q[2]) + 2.f * (rabs2(q[3]) + rabs2(q[4]) + rabs2(q[5])));
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

// Loads and stores of the lane's value of a dual or primal component.
template <typename DT>
__device__ inline float2 dual_ld_stream(const DT *p, const Vox &v, float r) {
  return DualIO<DT>::load_stream(p, v.i, r);
}
template <typename DT>
__device__ inline C2 dual_ld_stream(const DT *p, const Vox2 &v, float r) {
  return DualIO<DT>::load2_stream(p, v.i, r);
}
template <typename DT>
__device__ inline void dual_st(DT *p, const Vox &v, float2 val, float r) {
  DualIO<DT>::store(p, v.i, val, r);
}
template <typename DT>
__device__ inline void dual_st(DT *p, const Vox2 &v, C2 val, float r) {
  DualIO<DT>::store2(p, v.i, val, r);
}
__device__ inline void st_stream(float2 *p, const Vox &v, float2 val) { gst_stream(p, v.i, val); }
__device__ inline void st_stream(float2 *p, const Vox2 &v, C2 val) { gst2_stream(p, v.i, val); }

// Dual update at voxel v (or voxel pair): y += sigma K_reg(ext), then the
// projections.  TV: tv.cpp:140-150; TGV2: tgv2.cpp:238-259; ICTGV2:
// ictgv2.cpp:338-381.
template <int M, typename DT, class V>
__device__ inline void dual_voxel(const Grid3 &g, const RegParams &rp, const float2 *__restrict__ ext,
                                  DT *__restrict__ y, float sig, const V &v) {
  typedef typename LaneT<V>::type L;
  const long N = g.N;
  const unsigned i = v.i;
  const Inv3 d1{rp.idx, rp.idy, rp.idt}, d2{rp.idx2, rp.idy2, rp.idt2};
  PlainAcc E{ext};
  L gr[3], e[6], p[6];
  // components [c0, c0 + n) of y: p += sigma * k, project, store
  auto ascend = [&](int c0, int n, const L *k, float r) {
#pragma unroll
    for (int c = 0; c < 6; ++c)
      if (c < n) p[c] = radd(dual_ld_stream(y + (c0 + c) * N, v, r), rscale(k[c], sig));
    if (n == 3) proj3(p, r); else proj6(p, r);
#pragma unroll
    for (int c = 0; c < 6; ++c)
      if (c < n) dual_st(y + (c0 + c) * N, v, p[c], r);
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
  for (int c = 0; c < 3; ++c) gr[c] = rsub(gr[c], ldv(E.at((1 + c) * N), v, i));
  ascend(0, 3, gr, rp.r1);
  symgrad6(E, N, 2 * N, 3 * N, v, g, d1, e);
  ascend(3, 6, e, rp.r2);
  if (M == kICTGV2) {
    grad3(E, 4 * N, v, g, d2, gr);
#pragma unroll
    for (int c = 0; c < 3; ++c) gr[c] = rsub(gr[c], ldv(E.at((5 + c) * N), v, i));
    ascend(9, 3, gr, rp.r3);
    symgrad6(E, 5 * N, 6 * N, 7 * N, v, g, d2, e);
    ascend(12, 6, e, rp.r4);
  }
}

// Primal update + extragradient at voxel v (or voxel pair).
// TV: tv.cpp:155-169; TGV2: tgv2.cpp:268-303; ICTGV2: ictgv2.cpp:388-445.
template <int M, typename DT, class V>
__device__ inline void primal_voxel(const Grid3 &g, const RegParams &rp, float2 *__restrict__ x,
                                    float2 *__restrict__ ext, const DT *__restrict__ y,
                                    const float2 *__restrict__ khz, float tau, const V &v) {
  typedef typename LaneT<V>::type L;
  constexpr int P = M == kTV ? 1 : (M == kICTV ? 2 : (M == kTGV2 ? 4 : 8));
  const long N = g.N;
  const unsigned i = v.i;
  const Inv3 d1{rp.idx, rp.idy, rp.idt}, d2{rp.idx2, rp.idy2, rp.idt2};
  const DualAcc<DT> Y1{y, rp.r1}, Y2{y, rp.r2}, Y3{y, rp.r3}, Y4{y, rp.r4};
  const PlainAcc X{x}, KHZ{khz};
  // x is read once, before any store: the extragradient needs the old values
  L xo[P], upd[P];
#pragma unroll
  for (int c = 0; c < P; ++c) xo[c] = ldv(X.at(c * N), v, i);
  const L dv1 = div3(Y1, 0, N, 2 * N, v, g, d1);
  upd[0] = rsub(xo[0], rscale(rsub(ldv(KHZ, v, i), dv1), tau));
  if constexpr (M == kICTV) {  // ictv.cpp:301-310: w <- w - tau (div1 y1 - div2 y3)
    const L dv3 = div3(Y3, 3 * N, 4 * N, 5 * N, v, g, d2);
    upd[1] = rsub(xo[1], rscale(rsub(dv1, dv3), tau));
  } else if constexpr (M != kTV) {
    L sd[3];
    symdiv3(Y2, 3 * N, N, v, g, d1, sd);
#pragma unroll
    for (int c = 0; c < 3; ++c)
      upd[1 + c] = radd(xo[1 + c], rscale(radd(ldv(Y1.at(c * N), v, i), sd[c]), tau));
    if constexpr (M == kICTGV2) {
      const L dv3 = div3(Y3, 9 * N, 10 * N, 11 * N, v, g, d2);
      upd[4] = rsub(xo[4], rscale(rsub(dv1, dv3), tau));
      symdiv3(Y4, 12 * N, N, v, g, d2, sd);
#pragma unroll
      for (int c = 0; c < 3; ++c)
        upd[5 + c] = radd(xo[5 + c], rscale(radd(ldv(Y3.at((9 + c) * N), v, i), sd[c]), tau));
    }
  }
#pragma unroll
  for (int c = 0; c < P; ++c) {
    st_stream(ext + c * N, v, rsub(rscale(upd[c], 2.f), xo[c]));
    st_stream(x + c * N, v, upd[c]);
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
  for (int c = 0; c < P; ++c) ax += (double)rabs2(D.at(c * N)(i));
  // ICTV's AdaptStepSize (ictv.cpp:96-131) overwrites its regulariser sum
  // before use: nKx holds the data term alone, so nothing is added here.
  if (M == kICTV) return;
  if (M == kTV) {
    grad3(D, 0, v, g, d1, gr);
    for (int c = 0; c < 3; ++c) ak += (double)rabs2(gr[c]);
    return;
  }
  if (M == kTGV2) {
    grad3(D, 0, v, g, d1, gr);
  } else {
    grad3(Diff2Acc{ext, x, ext + 4 * N, x + 4 * N}, 0, v, g, d1, gr);
  }
  for (int c = 0; c < 3; ++c) ak += (double)rabs2(rsub(gr[c], D.at((1 + c) * N)(i)));
  symgrad6(D, N, 2 * N, 3 * N, v, g, d1, e);
  for (int k = 0; k < 6; ++k) ak += (k < 3 ? 1.0 : 2.0) * (double)rabs2(e[k]);
  if (M == kICTGV2) {
    grad3(D, 4 * N, v, g, d2, gr);
    for (int c = 0; c < 3; ++c) ak += (double)rabs2(rsub(gr[c], D.at((5 + c) * N)(i)));
    symgrad6(D, 5 * N, 6 * N, 7 * N, v, g, d2, e);
    for (int k = 0; k < 6; ++k) ak += (k < 3 ? 1.0 : 2.0) * (double)rabs2(e[k]);
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

#pragma clang fp contract(fast)
