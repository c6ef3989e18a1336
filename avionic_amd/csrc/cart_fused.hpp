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
#ifndef AVIONIC_COL_UCACHE
#define AVIONIC_COL_UCACHE 0
#endif

constexpr int RTILE = kRowTileElems;
constexpr int NTC = AVIONIC_COL_THREADS;  // threads of a column-pass block
constexpr int kColSlots = max_slots(kColTileElems, NTC);
constexpr bool kColUcache = AVIONIC_COL_UCACHE != 0;  // keep the image tile in registers
#ifndef AVIONIC_COL_ACC
#define AVIONIC_COL_ACC 1
#endif
// adjoint coil sum in registers (else read-modify-write of the output per coil)
constexpr bool kColAcc = AVIONIC_COL_ACC != 0;
#ifndef AVIONIC_COL_PREFETCH
#define AVIONIC_COL_PREFETCH 0
#endif
// load the next coil's b1 (forward) / zhat and b1 (adjoint) while this coil
// is transformed
constexpr bool kColPrefetch = AVIONIC_COL_PREFETCH != 0;
constexpr int kPre = kColPrefetch ? kColSlots : 1;

// ------------------------------------------------------------ column forward
// 32-bit offsets from per-(t, c) base pointers (uniform, in SGPRs) keep one
// VGPR per in-flight address.  The image tile is re-read per coil (an L2 hit)
// by default: keeping it in registers costs occupancy, and on MI355X 256-lane
// blocks without that cache were fastest (0.91 ms against 1.94 ms for
// 512-lane blocks with it, ICTGV2 256x256x30x16, 8 reconstructions).
template <bool DELTA>
struct ColFwdLoad {
  float2 ucache[kColUcache ? kColSlots : 1];
  float2 bpre[kPre];
  const float2 *ut, *st, *bc, *bn;  // bn: next coil's map (prefetch), or null
  int H, W, xb0, centered;
  bool first;
  __device__ float2 operator()(int b, int i, int slot) {
    const int x = xb0 + b;
    if (x >= W) return czero();
    const int sy = centered ? (i + H / 2) % H : i;
    const int sx = centered ? (x + W / 2) % W : x;
    const unsigned si = (unsigned)(sy * W + sx);
    float2 bv;
    if (kColPrefetch) {
      bv = first ? bc[si] : bpre[slot];
      if (bn) bpre[slot] = bn[si];
    } else {
      bv = bc[si];
    }
    if (!kColUcache) return cmul(bv, DELTA ? csub(ut[si], st[si]) : ut[si]);
    if (first) ucache[slot] = DELTA ? csub(ut[si], st[si]) : ut[si];
    return cmul(bv, ucache[slot]);
  }
};
struct ColFwdStore {
  float2 *kc;       // kcol + (C r0 + c S) W : rows of (t, c)
  const int *rmap;  // rowmap + t H
  int H, W, xb0, centered;
  __device__ void operator()(int b, int i, float2 v, int) {
    const int x = xb0 + b;
    if (x >= W) return;
    const int ky = centered ? (i + H / 2) % H : i;
    const int k = rmap[ky];
    if (k >= 0) kc[(unsigned)(k * W + x)] = v;
  }
};

// The frame's rowmap (compact row of each k-space row, -1 if not sampled)
// staged in LDS: the store / load functors then need no dependent global
// read before each k-space access.
__device__ inline const int *stage_rowmap(const CartGeom &g, int t, int *lds, int tid, int nt) {
  const int *rm = g.rowmap + (long)t * g.H;
  for (int i = tid; i < g.H; i += nt) lds[i] = rm[i];
  return lds;
}

template <bool DELTA>
__device__ inline void col_forward_fused(const CartGeom &g, int xbi, int t, const float2 *u,
                                         const float2 *usub, const float2 *b1, float2 *kcol,
                                         float2 *tile, float2 *twl, int *rml) {
  const int tid = threadIdx.x;
  const long HW = (long)g.H * g.W;
  const float2 *tw = stage_twiddles(g.ph, twl, tid, NTC);
  const int *rmap = stage_rowmap(g, t, rml, tid, NTC);
  __syncthreads();
  const int r0 = g.row_off[t], S = g.row_off[t + 1] - r0;
  ColFwdLoad<DELTA> ld;
  ld.ut = u + t * HW;
  ld.st = DELTA ? usub + t * HW : nullptr;
  ld.H = g.H;
  ld.W = g.W;
  ld.xb0 = xbi * g.XBF;
  ld.centered = g.centered;
  ColFwdStore st{nullptr, rmap, g.H, g.W, xbi * g.XBF, g.centered};
  for (int c = 0; c < g.C; ++c) {
    ld.bc = b1 + c * HW;
    ld.bn = (c + 1 < g.C) ? b1 + (c + 1) * HW : nullptr;
    ld.first = (c == 0);
    st.kc = kcol + ((long)g.C * r0 + (long)c * S) * g.W;
    fft_io<NTC, kColTileElems, true, true, true>(tile, g.ph, g.XBF, 1, g.XBF, +1.f, tid, tw, ld, st);
    __syncthreads();
  }
}

// ------------------------------------------------------------ column adjoint
struct ColAdjLoad {
  float2 zpre[kPre];
  const float2 *zh, *zn;  // zhat rows of (t, c); of (t, c + 1) or null
  const int *rmap;
  int H, W, xb0, centered;
  bool first;
  __device__ float2 operator()(int b, int i, int slot) {
    const int x = xb0 + b;
    if (x >= W) return czero();
    const int ky = centered ? (i + H / 2) % H : i;
    const int k = rmap[ky];
    if (k < 0) return czero();
    const unsigned o = (unsigned)(k * W + x);
    if (!kColPrefetch) return zh[o];
    const float2 v = first ? zh[o] : zpre[slot];
    if (zn) zpre[slot] = zn[o];
    return v;
  }
};
struct ColAdjStore {
  float2 acc[kColAcc ? kColSlots : 1];
  float2 bpre[kPre];
  const float2 *bc, *bn;
  float2 *out;  // image frame t
  int H, W, xb0, centered;
  bool first, last;
  float scale;
  __device__ void operator()(int b, int i, float2 v, int slot) {
    const int x = xb0 + b;
    if (x >= W) return;
    const int y = centered ? (i + H / 2) % H : i;
    const unsigned o = (unsigned)(y * W + x);
    float2 bv;
    if (kColPrefetch) {
      bv = first ? bc[o] : bpre[slot];
      if (bn) bpre[slot] = bn[o];
    } else {
      bv = bc[o];
    }
    if (!kColAcc) {
      const float2 a = cadd(first ? czero() : out[o], cmulc(bv, v));
      out[o] = last ? cscale(a, scale) : a;
      return;
    }
    const float2 a = cadd(acc[slot], cmulc(bv, v));
    if (last)
      out[o] = cscale(a, scale);
    else
      acc[slot] = a;
  }
};

__device__ inline void col_adjoint_fused(const CartGeom &g, int xbi, int t, const float2 *zhat,
                                         const float2 *b1, float2 *out, float2 *tile,
                                         float2 *twl, int *rml) {
  const int tid = threadIdx.x;
  const long HW = (long)g.H * g.W;
  const float2 *tw = stage_twiddles(g.ph, twl, tid, NTC);
  const int *rmap = stage_rowmap(g, t, rml, tid, NTC);
  __syncthreads();
  const int r0 = g.row_off[t], S = g.row_off[t + 1] - r0;
  ColAdjLoad ld;
  ld.rmap = rmap;
  ld.H = g.H;
  ld.W = g.W;
  ld.xb0 = xbi * g.XBF;
  ld.centered = g.centered;
  ColAdjStore st;
#pragma unroll
  for (int k = 0; k < (kColAcc ? kColSlots : 1); ++k) st.acc[k] = czero();
  st.out = out + t * HW;
  st.H = g.H;
  st.W = g.W;
  st.xb0 = xbi * g.XBF;
  st.centered = g.centered;
  st.scale = g.scale;
  for (int c = 0; c < g.C; ++c) {
    ld.zh = zhat + ((long)g.C * r0 + (long)c * S) * g.W;
    ld.zn = (c + 1 < g.C) ? ld.zh + (long)S * g.W : nullptr;
    ld.first = (c == 0);
    st.bc = b1 + c * HW;
    st.bn = (c + 1 < g.C) ? b1 + (c + 1) * HW : nullptr;
    st.first = (c == 0);
    st.last = (c == g.C - 1);
    fft_io<NTC, kColTileElems, true, true, true>(tile, g.ph, g.XBF, 1, g.XBF, -1.f, tid, tw, ld, st);
    __syncthreads();
  }
}

// ----------------------------------------------------------------- row pass
struct RowLoad {
  const float2 *src;  // kcol + g0 W
  int W;
  __device__ float2 operator()(int b, int i, int) const { return src[(unsigned)(b * W + i)]; }
};
// Mask row (index into mrows) of each row of the block, staged in LDS so
// that the mask load does not wait on a global rowinfo load.
__device__ inline const int *stage_mask_rows(const CartGeom &g, long g0, int nr, int *lds, int tid,
                                             int nt) {
  for (int b = tid; b < nr; b += nt) lds[b] = g.rowinfo[g0 + b].z;
  return lds;
}

struct RowDualStore {
  float2 *z;
  const float2 *d;
  const float *mrows;
  const int *mrow;  // LDS: mask row of local row b
  float2 *tile2;
  long g0;
  int W, wc, centered;
  float scale, sigma, inv;
  __device__ void operator()(int b, int i, float2 v, int) {
    const int kx = centered ? (i + W / 2) % W : i;
    const float mk = mrows[(long)mrow[b] * W + kx];
    const float2 val = cscale(v, scale * mk);
    const long gi = (g0 + b) * W + kx;
    float2 zz = z[gi];
    zz = cadd(zz, cscale(val, sigma));
    zz = cscale(csub(zz, cscale(d[gi], sigma)), inv);
    z[gi] = zz;
    const int xp = centered ? (kx + wc) % W : kx;
    tile2[b * W + xp] = cscale(zz, mk);
  }
};
struct RowAdjStore {
  float2 *zhat;
  long g0;
  int W, centered;
  __device__ void operator()(int b, int i, float2 v, int) {
    const int kxp = centered ? (i + W / 2) % W : i;
    zhat[(g0 + b) * W + kxp] = v;
  }
};
struct RowNorm2Store {
  const float *mrows;
  const int *mrow;  // LDS: mask row of local row b
  long g0;
  int W, centered;
  float scale;
  double acc;
  __device__ void operator()(int b, int i, float2 v, int) {
    const int kx = centered ? (i + W / 2) % W : i;
    const float2 val = cscale(v, scale * mrows[(long)mrow[b] * W + kx]);
    acc += (double)cabs2(val);
  }
};

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
  st.z = z;
  st.d = d;
  st.mrows = g.mrows;
  st.mrow = mrow;
  st.tile2 = tile2;
  st.g0 = g0;
  st.W = W;
  st.wc = (W + 1) / 2;
  st.centered = g.centered;
  st.scale = g.scale;
  st.sigma = (float)sc->sigma;
  st.inv = (float)(1.0 / (1.0 + sc->sigma / sc->lambda));
  fft_io<NT, RTILE, false, true, true>(tile, g.pw, nr, W, 1, +1.f, tid, tw, ld, st);
  __syncthreads();  // tile2 holds M z for every row of the block
  RowAdjStore st2{zhat, g0, W, g.centered};
  NoLoad nl;
  fft_io<NT, RTILE, false, false, true>(tile2, g.pw, nr, W, 1, -1.f, tid, tw, nl, st2);
}

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
  RowNorm2Store st{g.mrows, mrow, g0, W, g.centered, g.scale, 0.0};
  fft_io<NT, RTILE, false, true, true>(tile, g.pw, nr, W, 1, +1.f, tid, tw, ld, st);
  return st.acc;
}

}  // namespace avionic
