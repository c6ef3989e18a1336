// Fused Cartesian operator passes of the hot loop (task-list kernels).
//
// Column passes: the first radix stage reads the image (times the coil map)
// straight from global memory and the last stage hands each output to its
// consumer: the forward pass keeps only the sampled phase-encode rows, the
// adjoint pass accumulates conj(b1_c) * value over the coils in registers and
// writes K^H z after the last coil.  Row pass: forward row transform, z dual
// update and prox per output, the masked z staged in LDS for the adjoint row
// transform, whose last stage writes zhat.  Only the exchanges between radix
// stages go through LDS; twiddles are staged in LDS once per block.
#pragma once

#include "cart_device.hpp"

namespace avionic {

#ifndef AVIONIC_COL_THREADS
#define AVIONIC_COL_THREADS 256
#endif

constexpr int RTILE = kRowTileElems;
constexpr int NTC = AVIONIC_COL_THREADS;  // threads of a column-pass block
// Outputs per lane of a column-pass stage (narrow / wide plans).
constexpr int col_slots(bool wide) { return max_slots(kColTileElems, NTC, wide); }

// (i + n/2) mod n and (a + b) mod n for 0 <= i, a < n, 0 <= b <= n, without a
// division.
__device__ __forceinline__ int wrap_half(int i, int n) {
  const int v = i + n / 2;
  return v >= n ? v - n : v;
}
__device__ __forceinline__ int wrap_add(int a, int b, int n) {
  const int v = a + b;
  return v >= n ? v - n : v;
}

// ------------------------------------------------------------ column plans
// A column pass runs on the geometry's runtime plan (g.ph, g.XBF) or on one
// of a few compile-time plans.  With a compile-time plan every index of the
// transform (butterfly, Stockham position, twiddle, centring shift) is a
// constant expression and the integer divisions and remainders that the
// runtime plan evaluates per element fold into shifts and masks: measured on
// MI355X, the runtime-plan passes spend more VALU instructions on that index
// arithmetic than on the transform.
template <bool W>
struct ColPlanRtT {
  static constexpr int H = 0, XB = 0;
  static constexpr bool WIDE = W;
};
using ColPlanRt = ColPlanRtT<false>;
using ColPlanRtWide = ColPlanRtT<true>;
template <int H_, int XB_, int R0_, int R1_, int R2_ = 1>
struct ColPlanCt {
  static constexpr int H = H_, XB = XB_, R0 = R0_, R1 = R1_, R2 = R2_;
  static constexpr bool WIDE = false;
  static_assert(H_ * XB_ <= kColTileElems, "column tile exceeds the LDS tile");
  static_assert(R0_ * R1_ * R2_ == H_, "radices must multiply to the length");
  static_assert(H_ <= 256, "compile-time plans are at most 256 long");
};
// The radices in make_fft_plan's order; CartGeom::cplan names one.
constexpr int kColPlanMaxLen = 256;  // longest compile-time plan
using ColPlan256 = ColPlanCt<256, 16, 16, 16>;
using ColPlan128 = ColPlanCt<128, 32, 16, 8>;
using ColPlan160 = ColPlanCt<160, 16, 16, 2, 5>;
using ColPlan192 = ColPlanCt<192, 16, 16, 4, 3>;
using ColPlan224 = ColPlanCt<224, 16, 16, 2, 7>;

// Column transform of one tile, first stage from `ld`, last stage to `st`.
template <class P, class LD, class ST>
__device__ __forceinline__ void col_fft(float2 *tile, const CartGeom &g, float sign, int tid,
                                        const float2 *tw, LD &ld, ST &st) {
  if constexpr (P::H == 0) {
    fft_io<NTC, kColTileElems, true, true, true, P::WIDE>(tile, g.ph, g.XBF, 1, g.XBF, sign, tid, tw,
                                                          ld, st);
  } else {
    constexpr int L = P::H, B = P::XB, E = L * B;
    NoStore ns;
    NoLoad nl;
    if constexpr (P::R2 == 1) {
      fft_stage_io<P::R0, NTC, E, true, true, false>(tile, L, B, 1, B, 1, sign, tw, tid, ld, ns);
      fft_stage_io<P::R1, NTC, E, true, false, true>(tile, L, B, 1, B, P::R0, sign, tw, tid, nl, st);
    } else {
      fft_stage_io<P::R0, NTC, E, true, true, false>(tile, L, B, 1, B, 1, sign, tw, tid, ld, ns);
      fft_stage_io<P::R1, NTC, E, true, false, false>(tile, L, B, 1, B, P::R0, sign, tw, tid, nl,
                                                      ns);
      fft_stage_io<P::R2, NTC, E, true, false, true>(tile, L, B, 1, B, P::R0 * P::R1, sign, tw,
                                                     tid, nl, st);
    }
  }
}

// Compile-time plans.  Every load of a lane's first stage is issued before
// any of them is used (and the adjoint's coil-map loads before the last
// stage's arithmetic): without the scheduling barrier the compiler places
// each complex product right after its two loads, and the wait before it
// leaves one pair of loads in flight per lane.
template <int R_, int L, int B, int Ns>
struct CtStage {
  static constexpr int R = R_, NBF = L / R_, TOTAL = NBF * B, MAXB = (TOTAL + NTC - 1) / NTC;
  // c-th butterfly of lane tid: column b, butterfly j; false past the end
  __device__ static bool lane(int tid, int c, int &b, int &j) {
    const int idx = tid + c * NTC;
    b = idx % B;
    j = idx / B;
    return idx < TOTAL;
  }
  __device__ static int in_index(int j, int r) { return j + r * NBF; }
  __device__ static int out_index(int j, int r) { return (j / Ns) * Ns * R + j % Ns + r * Ns; }
  __device__ static void butterfly(float2 *v, int j, const float2 *tw, float sign) {
    if (Ns > 1) {
      const int k = j % Ns;
#pragma unroll
      for (int r = 1; r < R; ++r) {
        float2 w = tw[r * k * (L / (Ns * R))];
        w.y = -sign * w.y;  // tw holds exp(-i theta)
        v[r] = cmul(v[r], w);
      }
    }
    Dft<R>::run(v, sign);
  }
};

// First-stage elements per lane of a compile-time plan (0: runtime plan).
template <class P>
constexpr int col_first_slots() {
  if constexpr (P::H == 0)
    return 0;
  else
    return CtStage<P::R0, P::H, P::XB, 1>::MAXB * P::R0;
}

// Outputs v of stage S0 become inputs w of stage S1 through the LDS tile
// (element (b, i) at b + i B).  Callers sync before writing the tile again.
template <int B, class S0, class S1>
__device__ __forceinline__ void ct_exchange(float2 *t, float2 (&v)[S0::MAXB][S0::R],
                                            float2 (&w)[S1::MAXB][S1::R], int tid) {
  int b, j;
#pragma unroll
  for (int c = 0; c < S0::MAXB; ++c)
    if (S0::lane(tid, c, b, j))
#pragma unroll
      for (int r = 0; r < S0::R; ++r) t[b + S0::out_index(j, r) * B] = v[c][r];
  __syncthreads();
#pragma unroll
  for (int c = 0; c < S1::MAXB; ++c)
    if (S1::lane(tid, c, b, j))
#pragma unroll
      for (int r = 0; r < S1::R; ++r) w[c][r] = t[b + S1::in_index(j, r) * B];
}

// Last stage S: prefetch what the store functor reads, transform, store.
template <class S, class ST>
__device__ __forceinline__ void ct_last(float2 (&w)[S::MAXB][S::R], int tid, const float2 *tw,
                                        float sign, ST &st) {
  int b, j;
#pragma unroll
  for (int c = 0; c < S::MAXB; ++c)
    if (S::lane(tid, c, b, j))
#pragma unroll
      for (int r = 0; r < S::R; ++r) st.prefetch(b, S::out_index(j, r), c * S::R + r);
  __builtin_amdgcn_sched_barrier(0);
#pragma unroll
  for (int c = 0; c < S::MAXB; ++c)
    if (S::lane(tid, c, b, j)) {
      S::butterfly(w[c], j, tw, sign);
#pragma unroll
      for (int r = 0; r < S::R; ++r) st(b, S::out_index(j, r), w[c][r], c * S::R + r);
    }
}

// Column transform of one tile on a compile-time plan, first stage from
// `ld` (raw loads, then finish), last stage to `st`.
template <class P, class LD, class ST>
__device__ __forceinline__ void col_fft_ct(float2 *tile, float sign, int tid, const float2 *tw,
                                           LD &ld, ST &st) {
  asm volatile("" : "+v"(tid));  // per-lane index math stays inside the coil loop
  constexpr int L = P::H, B = P::XB;
  using S0 = CtStage<P::R0, L, B, 1>;
  using S1 = CtStage<P::R1, L, B, P::R0>;
  int b, j;
  typename LD::Raw raw[S0::MAXB][S0::R];
#pragma unroll
  for (int c = 0; c < S0::MAXB; ++c)
    if (S0::lane(tid, c, b, j))
#pragma unroll
      for (int r = 0; r < S0::R; ++r) raw[c][r] = ld.raw(b, S0::in_index(j, r), c * S0::R + r);
  __builtin_amdgcn_sched_barrier(0);
  float2 v[S0::MAXB][S0::R];
#pragma unroll
  for (int c = 0; c < S0::MAXB; ++c)
    if (S0::lane(tid, c, b, j)) {
#pragma unroll
      for (int r = 0; r < S0::R; ++r) v[c][r] = ld.finish(raw[c][r], b, c * S0::R + r);
      S0::butterfly(v[c], j, tw, sign);
    }
  float2 w[S1::MAXB][S1::R];
  ct_exchange<B, S0, S1>(tile, v, w, tid);
  if constexpr (P::R2 == 1) {
    ct_last<S1>(w, tid, tw, sign, st);
  } else {
    using S2 = CtStage<P::R2, L, B, P::R0 * P::R1>;
#pragma unroll
    for (int c = 0; c < S1::MAXB; ++c)
      if (S1::lane(tid, c, b, j)) S1::butterfly(w[c], j, tw, sign);
    float2 x[S2::MAXB][S2::R];
    __syncthreads();  // every lane has read the previous exchange
    ct_exchange<B, S1, S2>(tile, w, x, tid);
    ct_last<S2>(x, tid, tw, sign, st);
  }
}

// ------------------------------------------------------------ column forward
// 32-bit offsets from per-(t, c) base pointers (uniform, in SGPRs) keep one
// VGPR per in-flight address.  The image tile is re-read per coil (an L2
// hit): keeping it in registers costs occupancy, and on MI355X 256-lane
// blocks without that cache were fastest (0.91 ms against 1.94 ms for
// 512-lane blocks with it, ICTGV2 256x256x30x16, 8 reconstructions).
// HS: the transform length when compile-time (0: runtime, H).  Lanes of
// columns past W load a valid element and drop it: a load in a divergent
// branch whose value is used after it gets its own wait, which serialises the
// lane's loads.
// Index products use 24-bit multiplies (rows and columns are below 2^12).
// SLOTS (compile-time plans: the first stage's elements per lane): the lane
// keeps its image values (u, or u - usub) in registers from the block's
// first coil (`first`) and loads only the coil map for the others; the lane
// -> element mapping of a stage is the same for every coil, and the product
// is formed the same way, so the values are those of the uncached form.
template <bool DELTA, int HS, int SLOTS = 0>
struct ColFwdLoad {
  const float2 *ut, *st, *bc;
  int H, W, xb0, centered;
  bool first = true;
  float2 img[SLOTS > 0 ? SLOTS : 1];
  struct Raw {
    float2 b, u, s;
  };
  __device__ Raw raw(int b, int i, int) const {
    const int Hh = HS ? HS : H;
    const bool cen = centered != 0;
    const int x = xb0 + b;
    const int xc = x < W ? x : W - 1;  // every lane loads in bounds: no branch
    const int sy = cen ? wrap_half(i, Hh) : i;
    const int sx = cen ? wrap_half(xc, W) : xc;
    const unsigned si = __umul24((unsigned)sy, (unsigned)W) + (unsigned)sx;
    if (SLOTS > 0 && !first) return Raw{gld(bc, si), czero(), czero()};
    return Raw{gld(bc, si), gld(ut, si), DELTA ? gld(st, si) : czero()};
  }
  __device__ float2 finish(const Raw &r, int b, int slot) {
    if constexpr (SLOTS > 0) {
      if (first) img[slot] = DELTA ? csub(r.u, r.s) : r.u;
      return xb0 + b < W ? cmul(r.b, img[slot]) : czero();
    } else {
      return xb0 + b < W ? cmul(r.b, DELTA ? csub(r.u, r.s) : r.u) : czero();
    }
  }
  __device__ float2 operator()(int b, int i, int slot) { return finish(raw(b, i, slot), b, slot); }
};
template <int HS>
struct ColFwdStore {
  float2 *kc;       // kcol + (C r0 + c S) W : rows of (t, c)
  const int *rmap;  // LDS: compact row of each k-space row of frame t
  int H, W, xb0, centered;
  __device__ void operator()(int b, int i, float2 v, int) {
    const int Hh = HS ? HS : H;
    const bool cen = centered != 0;
    const int x = xb0 + b;
    const int ky = cen ? wrap_half(i, Hh) : i;
    const int k = rmap[ky];
    if (x < W && k >= 0) gst(kc, __umul24((unsigned)k, (unsigned)W) + (unsigned)x, v);
  }
  __device__ void prefetch(int, int, int) {}
};

// The frame's rowmap (compact row of each k-space row, -1 if not sampled)
// staged in LDS: the store / load functors then need no dependent global
// read before each k-space access.
__device__ inline const int *stage_rowmap(const CartGeom &g, int t, int *lds, int tid, int nt) {
  const int *rm = g.rowmap + (long)t * g.H;
  for (int i = tid; i < g.H; i += nt) lds[i] = rm[i];
  return lds;
}

template <bool DELTA, class P>
__device__ inline void col_forward_fused(const CartGeom &g, int xbi, int t, const float2 *u,
                                         const float2 *usub, const float2 *b1, float2 *kcol,
                                         float2 *tile, float2 *twl, int *rml) {
  const int tid = threadIdx.x;
  const long HW = (long)g.H * g.W;
  const int XB = P::XB ? P::XB : g.XBF;
  const float2 *tw = stage_twiddles(g.ph, twl, tid, NTC);
  const int *rmap = stage_rowmap(g, t, rml, tid, NTC);
  __syncthreads();
  const int r0 = g.row_off[t], S = g.row_off[t + 1] - r0;
  ColFwdLoad<DELTA, P::H, col_first_slots<P>()> ld;
  ld.ut = u + t * HW;
  ld.st = DELTA ? usub + t * HW : nullptr;
  ld.H = g.H;
  ld.W = g.W;
  ld.xb0 = xbi * XB;
  ld.centered = g.centered;
  ColFwdStore<P::H> st{nullptr, rmap, g.H, g.W, xbi * XB, g.centered};
  for (int c = 0; c < g.C; ++c) {
    ld.first = c == 0;
    ld.bc = b1 + c * HW;
    st.kc = kcol + ((long)g.C * r0 + (long)c * S) * g.W;
    if constexpr (P::H == 0)
      col_fft<P>(tile, g, +1.f, tid, tw, ld, st);
    else
      col_fft_ct<P>(tile, +1.f, tid, tw, ld, st);
    __syncthreads();
  }
}

// ------------------------------------------------------------ column adjoint
// Lanes of unsampled rows (and columns past W) load element 0 of zh, a
// cached dummy, so that the loads of a lane's outputs are not in divergent
// branches: a load in a branch whose value is used after it gets its own
// wait, which serialises the lane's loads.
template <int HS>
struct ColAdjLoad {
  const float2 *zh;  // zhat rows of (t, c); when the frame has none, any valid element
  const int *rmap;
  int H, W, xb0, centered;
  struct Raw {
    float2 z;
    bool use;
  };
  __device__ Raw raw(int b, int i, int) const {
    const int Hh = HS ? HS : H;
    const bool cen = centered != 0;
    const int x = xb0 + b;
    const int ky = cen ? wrap_half(i, Hh) : i;
    const int k = rmap[ky];
    const bool use = x < W && k >= 0;
    return Raw{gld(zh, use ? __umul24((unsigned)k, (unsigned)W) + (unsigned)x : 0u), use};
  }
  __device__ float2 finish(const Raw &r, int, int) const { return r.use ? r.z : czero(); }
  __device__ float2 operator()(int b, int i, int) const { return finish(raw(b, i, 0), 0, 0); }
};
// The coil sum stays in registers, one accumulator per output of the lane.
// PRE: the coil map values are loaded by prefetch() ahead of the last stage.
template <int HS, bool PRE, int SLOTS>
struct ColAdjStore {
  float2 acc[SLOTS];
  float2 bpre[PRE ? SLOTS : 1];
  const float2 *bc;
  float2 *out;  // image frame t
  int H, W, xb0, centered;
  bool last;
  float scale;
  __device__ unsigned index(int b, int i) const {
    const int Hh = HS ? HS : H;
    const bool cen = centered != 0;
    const int x = xb0 + b;
    const int xc = x < W ? x : W - 1;  // b1 load in bounds for every lane: no branch
    const int y = cen ? wrap_half(i, Hh) : i;
    return __umul24((unsigned)y, (unsigned)W) + (unsigned)xc;
  }
  __device__ void prefetch(int b, int i, int slot) {
    if (PRE) bpre[slot] = gld(bc, index(b, i));
  }
  __device__ void operator()(int b, int i, float2 v, int slot) {
    const unsigned o = index(b, i);
    const float2 a = cadd(acc[slot], cmulc(PRE ? bpre[slot] : gld(bc, o), v));
    if (!last)
      acc[slot] = a;
    else if (xb0 + b < W)
      gst(out, o, cscale(a, scale));
  }
};

template <class P>
__device__ inline void col_adjoint_fused(const CartGeom &g, int xbi, int t, const float2 *zhat,
                                         const float2 *b1, float2 *out, float2 *tile,
                                         float2 *twl, int *rml) {
  const int tid = threadIdx.x;
  const long HW = (long)g.H * g.W;
  const int XB = P::XB ? P::XB : g.XBF;
  const float2 *tw = stage_twiddles(g.ph, twl, tid, NTC);
  const int *rmap = stage_rowmap(g, t, rml, tid, NTC);
  __syncthreads();
  const int r0 = g.row_off[t], S = g.row_off[t + 1] - r0;
  ColAdjLoad<P::H> ld;
  ld.rmap = rmap;
  ld.H = g.H;
  ld.W = g.W;
  ld.xb0 = xbi * XB;
  ld.centered = g.centered;
  constexpr int SLOTS = col_slots(P::WIDE);
  ColAdjStore<P::H, P::H != 0, SLOTS> st;
#pragma unroll
  for (int k = 0; k < SLOTS; ++k) st.acc[k] = czero();
  st.out = out + t * HW;
  st.H = g.H;
  st.W = g.W;
  st.xb0 = xbi * XB;
  st.centered = g.centered;
  st.scale = g.scale;
  for (int c = 0; c < g.C; ++c) {
    ld.zh = S > 0 ? zhat + ((long)g.C * r0 + (long)c * S) * g.W : zhat;
    st.bc = b1 + c * HW;
    st.last = (c == g.C - 1);
    if constexpr (P::H == 0)
      col_fft<P>(tile, g, -1.f, tid, tw, ld, st);
    else
      col_fft_ct<P>(tile, -1.f, tid, tw, ld, st);
    __syncthreads();
  }
}

// ----------------------------------------------------------------- row pass
struct RowLoad {
  const float2 *src;  // kcol + g0 W
  int W;
  __device__ float2 operator()(int b, int i, int) const { return gld(src, (unsigned)(b * W + i)); }
};
// Mask row (index into mrows) of each row of the block, staged in LDS so
// that the mask load does not wait on a global rowinfo load.
__device__ inline const int *stage_mask_rows(const CartGeom &g, long g0, int nr, int *lds, int tid,
                                             int nt) {
  for (int b = tid; b < nr; b += nt) lds[b] = g.rowinfo[g0 + b].z;
  return lds;
}

struct RowDualStore {
  float2 *z;        // z + g0 W: rows of the block
  const float2 *d;  // d + g0 W
  const float *mrows;
  const int *mrow;  // LDS: mask row of local row b
  float2 *tile2;
  int W, wc, centered;
  float scale, sigma, inv;
  __device__ void operator()(int b, int i, float2 v, int) {
    const int kx = centered ? wrap_half(i, W) : i;
    const float mk = gld(mrows, (unsigned)(mrow[b] * W + kx));
    const float2 val = cscale(v, scale * mk);
    const unsigned gi = (unsigned)(b * W + kx);
    float2 zz = gld(z, gi);
    zz = cadd(zz, cscale(val, sigma));
    zz = cscale(csub(zz, cscale(gld(d, gi), sigma)), inv);
    gst_stream(z, gi, zz);
    const int xp = centered ? wrap_add(kx, wc, W) : kx;
    tile2[b * W + xp] = cscale(zz, mk);
  }
};
struct RowAdjStore {
  float2 *zhat;  // zhat + g0 W
  int W, centered;
  __device__ void operator()(int b, int i, float2 v, int) {
    const int kxp = centered ? wrap_half(i, W) : i;
    gst(zhat, (unsigned)(b * W + kxp), v);
  }
};
struct RowNorm2Store {
  const float *mrows;
  const int *mrow;  // LDS: mask row of local row b
  int W, centered;
  float scale;
  double acc;
  __device__ void operator()(int b, int i, float2 v, int) {
    const int kx = centered ? wrap_half(i, W) : i;
    const float2 val = cscale(v, scale * gld(mrows, (unsigned)(mrow[b] * W + kx)));
    acc += (double)cabs2(val);
  }
};

template <bool WIDE>
__device__ inline void row_dual_fused(const CartGeom &g, int blk, const float2 *kcol, float2 *z,
                                      const float2 *d, float2 *zhat, const DevScalars *sc,
                                      float2 *tile, float2 *tile2, float2 *twl, int *mrl) {
  const int tid = threadIdx.x;
  const int W = g.W;
  const long g0 = (long)blk * g.RB;
  const int nr = (int)min((long)g.RB, g.nrows_c - g0);
  const float2 *tw = stage_twiddles(g.pw, twl, tid, NT);
  const int *mrow = stage_mask_rows(g, g0, nr, mrl, tid, NT);
  __syncthreads();
  RowLoad ld{kcol + g0 * W, W};
  RowDualStore st;
  st.z = z + g0 * W;
  st.d = d + g0 * W;
  st.mrows = g.mrows;
  st.mrow = mrow;
  st.tile2 = tile2;
  st.W = W;
  st.wc = (W + 1) / 2;
  st.centered = g.centered;
  st.scale = g.scale;
  st.sigma = (float)sc->sigma;
  st.inv = (float)(1.0 / (1.0 + sc->sigma / sc->lambda));
  fft_io<NT, RTILE, false, true, true, WIDE>(tile, g.pw, nr, W, 1, +1.f, tid, tw, ld, st);
  __syncthreads();  // tile2 holds M z for every row of the block
  RowAdjStore st2{zhat + g0 * W, W, g.centered};
  NoLoad nl;
  fft_io<NT, RTILE, false, false, true, WIDE>(tile2, g.pw, nr, W, 1, -1.f, tid, tw, nl, st2);
}

// ------------------------------------------------- row pass, compile-time
// Rows of length L = R0 * R1 (256: 16 x 16 on 16 rows per block, 128: 16 x 8
// on 32 rows), every lane one butterfly of every stage.  The lane mapping is
// rows-major (butterfly j fastest): the loads of a row are contiguous.  Per
// block:
//   forward stage 0 from kcol (all loads issued first) -> exchange ->
//   z, d and mask values of the lane's outputs prefetched -> forward stage 1
//   -> z ascent and prox per output (store z) -> M z stays in registers: the
//   outputs i = j + R0 r of a lane's last forward butterfly are the inputs of
//   its first inverse butterfly, and the centring shifts of K and K^H cancel
//   ((i + W/2) + (W+1)/2 = i + W) -> inverse stage 0 -> exchange -> inverse
//   stage 1 -> zhat.
// The tile holds element (b, i) at b L + (i ^ ((i >> 4) & 15)): the stage-0
// writes (i = 16 j + r over lanes j) and stage-1 reads (i = j + 16 r) then hit
// distinct banks.
template <int L_, int R0_, int R1_>
struct RowPlanCt {
  static constexpr int L = L_, R0 = R0_, R1 = R1_, ROWS = NT * R0_ / L_;
  static_assert(R0_ * R1_ == L_ && L_ % 16 == 0, "two-stage plan on 16-element groups");
  static_assert(ROWS * L_ <= kFftTileElems, "row tile exceeds the LDS tile");
};
using RowPlan256 = RowPlanCt<256, 16, 16>;
using RowPlan128 = RowPlanCt<128, 16, 8>;
constexpr int kRowPlanMaxRows = 32;

template <int R_, int L, int ROWS, int Ns>
struct RtStage {
  static constexpr int R = R_, NBF = L / R_, TOTAL = NBF * ROWS, MAXB = (TOTAL + NT - 1) / NT;
  static_assert(TOTAL % NT == 0, "every lane busy");
  __device__ static void lane(int tid, int c, int &b, int &j) {
    const int idx = tid + c * NT;
    j = idx % NBF;
    b = idx / NBF;
  }
  __device__ static int in_index(int j, int r) { return j + r * NBF; }
  __device__ static int out_index(int j, int r) { return (j / Ns) * Ns * R + j % Ns + r * Ns; }
  __device__ static void butterfly(float2 *v, int j, const float2 *tw, float sign) {
    CtStage<R, L, 1, Ns>::butterfly(v, j, tw, sign);
  }
};
__device__ __forceinline__ int row_swz(int b, int i, int L) { return b * L + (i ^ ((i >> 4) & 15)); }

template <class S0, class S1, int L>
__device__ __forceinline__ void rt_exchange(float2 *t, float2 (&v)[S0::MAXB][S0::R],
                                            float2 (&w)[S1::MAXB][S1::R], int tid) {
  int b, j;
#pragma unroll
  for (int c = 0; c < S0::MAXB; ++c) {
    S0::lane(tid, c, b, j);
#pragma unroll
    for (int r = 0; r < S0::R; ++r) t[row_swz(b, S0::out_index(j, r), L)] = v[c][r];
  }
  __syncthreads();
#pragma unroll
  for (int c = 0; c < S1::MAXB; ++c) {
    S1::lane(tid, c, b, j);
#pragma unroll
    for (int r = 0; r < S1::R; ++r) w[c][r] = t[row_swz(b, S1::in_index(j, r), L)];
  }
}

// MODE kRowDualUpdate: z update, zhat = F_row^-1(M z); kRowNorm2: returns the
// lane's sum of |M F_row kcol|^2 (scaled).
template <int MODE, class P>
__device__ inline double row_fused_ct(const CartGeom &g, int blk, const float2 *kcol, float2 *z,
                                      const float2 *d, float2 *zhat, const DevScalars *sc,
                                      float2 *tile, float2 *twl, int *mrl) {
  constexpr int L = P::L, ROWS = P::ROWS;
  using F0 = RtStage<P::R0, L, ROWS, 1>;
  using F1 = RtStage<P::R1, L, ROWS, P::R0>;
  int tid = threadIdx.x;
  const long g0 = (long)blk * ROWS;
  const int nr = (int)min((long)ROWS, g.nrows_c - g0);
  const float2 *tw = stage_twiddles(g.pw, twl, tid, NT);
  const int *mrow = stage_mask_rows(g, g0, nr, mrl, tid, NT);
  __syncthreads();
  asm volatile("" : "+v"(tid));
  const float2 *src = kcol + g0 * L;
  int b, j;
  // forward stage 0: rows past the block's last clamp to its last row
  float2 v[F0::MAXB][F0::R];
#pragma unroll
  for (int c = 0; c < F0::MAXB; ++c) {
    F0::lane(tid, c, b, j);
    const int bc = min(b, nr - 1);
#pragma unroll
    for (int r = 0; r < F0::R; ++r) v[c][r] = gld(src, (unsigned)(bc * L + F0::in_index(j, r)));
  }
  __builtin_amdgcn_sched_barrier(0);
#pragma unroll
  for (int c = 0; c < F0::MAXB; ++c) {
    F0::lane(tid, c, b, j);
    F0::butterfly(v[c], j, tw, +1.f);
  }
  float2 w[F1::MAXB][F1::R];
  rt_exchange<F0, F1, L>(tile, v, w, tid);
  // prefetch what the outputs need
  const int centered = g.centered;
  float mk[F1::MAXB][F1::R];
  float2 zz[F1::MAXB][F1::R], dd[F1::MAXB][F1::R];
#pragma unroll
  for (int c = 0; c < F1::MAXB; ++c) {
    F1::lane(tid, c, b, j);
    const int bc = min(b, nr - 1);
    const unsigned mbase = (unsigned)(mrow[bc] * L);
#pragma unroll
    for (int r = 0; r < F1::R; ++r) {
      const int i = F1::out_index(j, r);
      const int kx = centered ? wrap_half(i, L) : i;
      mk[c][r] = gld(g.mrows, mbase + kx);
      if (MODE == kRowDualUpdate) {
        const unsigned gi = (unsigned)(bc * L + kx);
        zz[c][r] = gld(z + g0 * L, gi);
        dd[c][r] = gld(d + g0 * L, gi);
      }
    }
  }
  __builtin_amdgcn_sched_barrier(0);
  const float scale = g.scale;
  if (MODE == kRowNorm2) {
    double acc = 0.0;
#pragma unroll
    for (int c = 0; c < F1::MAXB; ++c) {
      F1::lane(tid, c, b, j);
      F1::butterfly(w[c], j, tw, +1.f);
      if (b < nr)
#pragma unroll
        for (int r = 0; r < F1::R; ++r) acc += (double)cabs2(cscale(w[c][r], scale * mk[c][r]));
    }
    return acc;
  }
  const float sigma = (float)sc->sigma;
  const float inv = (float)(1.0 / (1.0 + sc->sigma / sc->lambda));
  float2 *zb = z + g0 * L;
  // forward stage 1, z update per output: M z at natural index i
#pragma unroll
  for (int c = 0; c < F1::MAXB; ++c) {
    F1::lane(tid, c, b, j);
    F1::butterfly(w[c], j, tw, +1.f);
#pragma unroll
    for (int r = 0; r < F1::R; ++r) {
      const int i = F1::out_index(j, r);
      const int kx = centered ? wrap_half(i, L) : i;
      const float2 val = cscale(w[c][r], scale * mk[c][r]);
      float2 zn = cadd(zz[c][r], cscale(val, sigma));
      zn = cscale(csub(zn, cscale(dd[c][r], sigma)), inv);
      if (b < nr) gst_stream(zb, (unsigned)(b * L + kx), zn);
      w[c][r] = cscale(zn, mk[c][r]);
    }
  }
  // inverse: stage 0 takes M z at natural indices j + R0 r of its row
  using I0 = F0;
  using I1 = F1;
  float2 u[I0::MAXB][I0::R];
  if constexpr (P::R0 == P::R1) {
    // same lane mapping: a lane's forward outputs are its inverse inputs
#pragma unroll
    for (int c = 0; c < I0::MAXB; ++c)
#pragma unroll
      for (int r = 0; r < I0::R; ++r) u[c][r] = w[c][r];
    __syncthreads();  // every lane has read the forward exchange
  } else {
    __syncthreads();  // every lane has read the forward exchange
#pragma unroll
    for (int c = 0; c < F1::MAXB; ++c) {
      F1::lane(tid, c, b, j);
#pragma unroll
      for (int r = 0; r < F1::R; ++r) tile[row_swz(b, F1::out_index(j, r), L)] = w[c][r];
    }
    __syncthreads();
#pragma unroll
    for (int c = 0; c < I0::MAXB; ++c) {
      I0::lane(tid, c, b, j);
#pragma unroll
      for (int r = 0; r < I0::R; ++r) u[c][r] = tile[row_swz(b, I0::in_index(j, r), L)];
    }
    __syncthreads();
  }
#pragma unroll
  for (int c = 0; c < I0::MAXB; ++c) {
    I0::lane(tid, c, b, j);
    I0::butterfly(u[c], j, tw, -1.f);
  }
  float2 x[I1::MAXB][I1::R];
  rt_exchange<I0, I1, L>(tile, u, x, tid);
  float2 *hb = zhat + g0 * L;
#pragma unroll
  for (int c = 0; c < I1::MAXB; ++c) {
    I1::lane(tid, c, b, j);
    I1::butterfly(x[c], j, tw, -1.f);
    if (b < nr)
#pragma unroll
      for (int r = 0; r < I1::R; ++r) {
        const int i = I1::out_index(j, r);
        gst(hb, (unsigned)(b * L + (centered ? wrap_half(i, L) : i)), x[c][r]);
      }
  }
  return 0.0;
}

template <bool WIDE>
__device__ inline double row_norm2_fused(const CartGeom &g, int blk, const float2 *kcol,
                                         float2 *tile, float2 *twl, int *mrl) {
  const int tid = threadIdx.x;
  const int W = g.W;
  const long g0 = (long)blk * g.RB;
  const int nr = (int)min((long)g.RB, g.nrows_c - g0);
  const float2 *tw = stage_twiddles(g.pw, twl, tid, NT);
  const int *mrow = stage_mask_rows(g, g0, nr, mrl, tid, NT);
  __syncthreads();
  RowLoad ld{kcol + g0 * W, W};
  RowNorm2Store st{g.mrows, mrow, W, g.centered, g.scale, 0.0};
  fft_io<NT, RTILE, false, true, true, WIDE>(tile, g.pw, nr, W, 1, +1.f, tid, tw, ld, st);
  return st.acc;
}

}  // namespace avionic
