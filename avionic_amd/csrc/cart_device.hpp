// Device-side block bodies of the Cartesian operator passes (shared by the
// per-operator kernels in k_cartesian.hip and the task-list kernels in
// k_tasks.hip).  See cartesian.hpp for the algorithm.
#pragma once

#include "cartesian.hpp"
#include "fft_device.hpp"

namespace avionic {

constexpr int NT = kFftThreads;
constexpr int EMAX = kFftTileElems;

__device__ inline long kidx(const CartGeom &g, int t, int c, int k, int x) {
  const int r0 = g.row_off[t];
  const int S = g.row_off[t + 1] - r0;
  return ((long)g.C * r0 + (long)c * S + k) * g.W + x;
}

// ---------------------------------------------------------------------------
// Column pass of K.  grid (nxb, T); each block holds an H x XB tile and loops
// over the coils: tile = b1_c . u (ifftshifted in y and x when centred),
// inverse FFT along y, keep the sampled rows (fftshifted in y).  The image
// tile is re-read per coil from L2; HBM sees it once.
__device__ inline void col_forward_block(const CartGeom &g, int xbi, int t,
                                         const float2 *__restrict__ u,
                                         const float2 *__restrict__ usub,
                                         const float2 *__restrict__ b1, float2 *__restrict__ kcol,
                                         float2 *tile) {
  const int tid = threadIdx.x;
  const int xb0 = xbi * g.XB;
  const int H = g.H, W = g.W, XB = g.XB;
  const int HX = H * XB;
  const long HW = (long)H * W;
  const float2 *ut = u + (long)t * HW;
  const float2 *st = usub ? usub + (long)t * HW : nullptr;
  const int r0 = g.row_off[t];
  const int S = g.row_off[t + 1] - r0;
  const int hc = (H + 1) / 2;
  for (int c = 0; c < g.C; ++c) {
    const float2 *bc = b1 + (long)c * HW;
    for (int e = tid; e < HX; e += NT) {
      const int y = e / XB, x = xb0 + e % XB;
      float2 v = czero();
      if (x < W) {
        const int sy = g.centered ? (y + H / 2) % H : y;
        const int sx = g.centered ? (x + W / 2) % W : x;
        const long si = (long)sy * W + sx;
        const float2 uu = st ? csub(ut[si], st[si]) : ut[si];
        v = cmul(bc[si], uu);
      }
      tile[e] = v;
    }
    __syncthreads();
    fft_lds<NT, EMAX, true>(tile, g.ph, XB, 1, XB, +1.f, tid);
    for (int e = tid; e < S * XB; e += NT) {
      const int k = e / XB, xb = e % XB, x = xb0 + xb;
      if (x < W) {
        const int ky = g.rows[r0 + k];
        const int src = g.centered ? (ky + hc) % H : ky;
        // x stays in ifftshifted order; the row pass fftshifts its output.
        kcol[kidx(g, t, c, k, x)] = tile[src * XB + xb];
      }
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Row pass.  Each block transforms RB compact rows of length W along x.  The
// adjoint's row transform (modes DualUpdate / AdjDense / AdjCompact) uses a
// second LDS tile.
template <int MODE>
__device__ inline void row_pass_block(const CartGeom &g, int blk, const float2 *__restrict__ kcol,
                                      float2 *__restrict__ z, const float2 *__restrict__ d,
                                      float2 *__restrict__ zhat, float2 *__restrict__ out,
                                      const float2 *__restrict__ zin,
                                      const DevScalars *__restrict__ sc,
                                      double *__restrict__ partial, float2 *tile, float2 *tile2,
                                      double *red) {
  const int tid = threadIdx.x;
  const int W = g.W;
  const long g0 = (long)blk * g.RB;
  const int nr = (int)min((long)g.RB, g.nrows_c - g0);
  const int n = nr * W;
  const int wc = (W + 1) / 2;
  double acc = 0.0;
  float2 *adj = (MODE == kRowDualUpdate) ? tile2 : tile;

  if (MODE == kRowAdjDense || MODE == kRowAdjCompact) {
    // a[x'] = (M z)[(x' + W/2) % W] when centred.
    for (int e = tid; e < n; e += NT) {
      const int r = e / W, xp = e % W;
      const int4 ri = g.rowinfo[g0 + r];
      const int kx = g.centered ? (xp + W / 2) % W : xp;
      float2 v;
      if (MODE == kRowAdjDense)
        v = zin[(((long)ri.x * g.C + ri.y) * g.H + ri.w) * W + kx];
      else
        v = zin[(g0 + r) * W + kx];
      tile[e] = cscale(v, g.mrows[(long)ri.z * W + kx]);
    }
    __syncthreads();
  } else {
    for (int e = tid; e < n; e += NT) tile[e] = kcol[g0 * W + e];
    __syncthreads();
    fft_lds<NT, EMAX, false>(tile, g.pw, nr, W, 1, +1.f, tid);
    float sigma = 0.f, inv = 0.f;
    if (MODE == kRowDualUpdate) {
      sigma = (float)sc->sigma;
      inv = (float)(1.0 / (1.0 + sc->sigma / sc->lambda));
    }
    for (int e = tid; e < n; e += NT) {
      const int r = e / W, kx = e % W;
      const int src = g.centered ? (kx + wc) % W : kx;
      const int4 ri = g.rowinfo[g0 + r];
      const float mk = g.mrows[(long)ri.z * W + kx];
      const float2 val = cscale(tile[r * W + src], g.scale * mk);
      const long gi = (g0 + r) * W + kx;
      if (MODE == kRowApplyK) {
        if (out)
          out[(((long)ri.x * g.C + ri.y) * g.H + ri.w) * W + kx] = val;
        else
          zhat[gi] = val;
      } else if (MODE == kRowNorm2) {
        acc += (double)cabs2(val);
      } else if (MODE == kRowResidual) {
        acc += (double)cabs2(csub(val, d[gi]));
      } else if (MODE == kRowDualUpdate) {
        float2 zz = z[gi];
        zz = cadd(zz, cscale(val, sigma));
        zz = cscale(csub(zz, cscale(d[gi], sigma)), inv);
        z[gi] = zz;
        const int xp = g.centered ? (kx + wc) % W : kx;
        tile2[r * W + xp] = cscale(zz, mk);
      }
    }
    if (MODE == kRowDualUpdate) __syncthreads();
  }

  if (MODE == kRowDualUpdate || MODE == kRowAdjDense || MODE == kRowAdjCompact) {
    fft_lds<NT, EMAX, false>(adj, g.pw, nr, W, 1, -1.f, tid);
    for (int e = tid; e < n; e += NT) {
      const int r = e / W, kxp = e % W;
      const int src = g.centered ? (kxp + wc) % W : kxp;
      zhat[g0 * W + e] = adj[r * W + src];
    }
  }

  if (MODE == kRowNorm2 || MODE == kRowResidual) {
    red[tid] = acc;
    __syncthreads();
    for (int s = NT / 2; s > 0; s >>= 1) {
      if (tid < s) red[tid] += red[tid + s];
      __syncthreads();
    }
    if (tid == 0) partial[blk] = red[0];
  }
}

// ---------------------------------------------------------------------------
// Column pass of K^H with the coil reduction.  grid (nxb, T).  The coil sum
// accumulates in a second LDS tile.
__device__ inline void col_adjoint_block(const CartGeom &g, int xbi, int t,
                                         const float2 *__restrict__ zhat,
                                         const float2 *__restrict__ b1, float2 *__restrict__ out,
                                         float2 *tile, float2 *acc) {
  const int tid = threadIdx.x;
  const int xb0 = xbi * g.XB;
  const int H = g.H, W = g.W, XB = g.XB;
  const int HX = H * XB;
  const long HW = (long)H * W;
  const int r0 = g.row_off[t];
  const int S = g.row_off[t + 1] - r0;
  const int hc = (H + 1) / 2;
  for (int e = tid; e < HX; e += NT) acc[e] = czero();
  for (int c = 0; c < g.C; ++c) {
    for (int e = tid; e < HX; e += NT) tile[e] = czero();
    __syncthreads();
    for (int e = tid; e < S * XB; e += NT) {
      const int k = e / XB, xb = e % XB, x = xb0 + xb;
      if (x < W) {
        const int ky = g.rows[r0 + k];
        const int py = g.centered ? (ky + hc) % H : ky;
        tile[py * XB + xb] = zhat[kidx(g, t, c, k, x)];
      }
    }
    __syncthreads();
    fft_lds<NT, EMAX, true>(tile, g.ph, XB, 1, XB, -1.f, tid);
    const float2 *bc = b1 + (long)c * HW;
    for (int e = tid; e < HX; e += NT) {
      const int y = e / XB, x = xb0 + e % XB;
      if (x < W) {
        const int src = g.centered ? (y + hc) % H : y;
        acc[e] = cadd(acc[e], cmulc(bc[(long)y * W + x], tile[src * XB + e % XB]));
      }
    }
    __syncthreads();
  }
  for (int e = tid; e < HX; e += NT) {
    const int y = e / XB, x = xb0 + e % XB;
    if (x < W) out[(long)t * HW + (long)y * W + x] = cscale(acc[e], g.scale);
  }
}

}  // namespace avionic
