// Cartesian operator kernels for the per-operator API (K, K^H, norms); the
// block bodies live in cart_device.hpp.
#include <algorithm>
#include <cmath>
#include <vector>

#include "cart_device.hpp"

namespace avionic {

__global__ __launch_bounds__(NT) void k_col_forward(CartGeom g, const float2 *__restrict__ u,
                                                    const float2 *__restrict__ usub,
                                                    const float2 *__restrict__ b1,
                                                    float2 *__restrict__ kcol) {
  __shared__ float2 tile[EMAX];
  col_forward_block(g, blockIdx.x, blockIdx.y, u, usub, b1, kcol, tile);
}

template <int MODE>
__global__ __launch_bounds__(NT) void k_row_pass(CartGeom g, const float2 *__restrict__ kcol,
                                                 float2 *__restrict__ z,
                                                 const float2 *__restrict__ d,
                                                 float2 *__restrict__ zhat,
                                                 float2 *__restrict__ out,
                                                 const float2 *__restrict__ zin,
                                                 const DevScalars *__restrict__ sc,
                                                 double *__restrict__ partial) {
  __shared__ float2 tile[EMAX];
  __shared__ float2 tile2[(MODE == kRowDualUpdate) ? EMAX : 1];
  __shared__ double red[NT];
  row_pass_block<MODE>(g, blockIdx.x, kcol, z, d, zhat, out, zin, sc, partial, tile, tile2, red);
}

__global__ __launch_bounds__(NT) void k_col_adjoint(CartGeom g, const float2 *__restrict__ zhat,
                                                    const float2 *__restrict__ b1,
                                                    float2 *__restrict__ out) {
  __shared__ float2 tile[EMAX];
  __shared__ float2 acc[EMAX];
  col_adjoint_block(g, blockIdx.x, blockIdx.y, zhat, b1, out, tile, acc);
}

// Dense [t][c][H][W] -> compact rows, times the mask.
__global__ void k_compact_gather(CartGeom g, const float2 *__restrict__ dense,
                                 float2 *__restrict__ compact) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long n = g.nrows_c * g.W;
  if (i >= n) return;
  const long row = i / g.W;
  const int x = (int)(i % g.W);
  const int4 ri = g.rowinfo[row];
  compact[i] = cscale(dense[(((long)ri.x * g.C + ri.y) * g.H + ri.w) * g.W + x],
                      g.mrows[(long)ri.z * g.W + x]);
}

// Standalone batched 2D FFT (used by avionic_fft2, the AGILE FFT
// replacement): rows out of place (the centred y shift permutes rows between
// blocks), then columns in place.
__global__ __launch_bounds__(NT) void k_fft_rows_oop(const float2 *in, float2 *outp, FftPlan pw,
                                                     int H, int W, int RB, int sign,
                                                     int centered) {
  __shared__ float2 tile[EMAX];
  const int tid = threadIdx.x;
  const long img = blockIdx.y;
  const int y0 = blockIdx.x * RB;
  const int nr = min(RB, H - y0);
  const float2 *base = in + img * (long)H * W;
  float2 *ob = outp + img * (long)H * W;
  for (int e = tid; e < nr * W; e += NT) {
    const int r = e / W, x = e % W;
    const int y = y0 + r;
    const int sy = centered ? (y + H / 2) % H : y;
    const int sx = centered ? (x + W / 2) % W : x;
    tile[e] = base[(long)sy * W + sx];
  }
  __syncthreads();
  fft_lds<NT, EMAX, false>(tile, pw, nr, W, 1, (float)sign, tid);
  const int wc = (W + 1) / 2;
  for (int e = tid; e < nr * W; e += NT) {
    const int r = e / W, kx = e % W;
    const int src = centered ? (kx + wc) % W : kx;
    ob[(long)(y0 + r) * W + kx] = tile[r * W + src];
  }
}

__global__ __launch_bounds__(NT) void k_fft_cols(float2 *data, FftPlan ph, int H, int W, int XB,
                                                 int sign, int centered, float scale) {
  __shared__ float2 tile[EMAX];
  const int tid = threadIdx.x;
  const long img = blockIdx.y;
  const int xb0 = blockIdx.x * XB;
  float2 *base = data + img * (long)H * W;
  for (int e = tid; e < H * XB; e += NT) {
    const int y = e / XB, x = xb0 + e % XB;
    tile[e] = (x < W) ? base[(long)y * W + x] : czero();
  }
  __syncthreads();
  fft_lds<NT, EMAX, true>(tile, ph, XB, 1, XB, (float)sign, tid);
  const int hc = (H + 1) / 2;
  for (int e = tid; e < H * XB; e += NT) {
    const int y = e / XB, x = xb0 + e % XB;
    const int src = centered ? (y + hc) % H : y;
    if (x < W) base[(long)y * W + x] = cscale(tile[src * XB + e % XB], scale);
  }
}

// ---------------------------------------------------------------------------
// host side

#ifndef AVIONIC_COL_MAXXB
#define AVIONIC_COL_MAXXB 64
#endif
static int pick_xb(int H, int W, int emax = EMAX) {
  int xb = 1;
  while (xb * 2 * H <= emax && xb * 2 <= AVIONIC_COL_MAXXB && xb < W) xb *= 2;
  return xb;
}

static int pick_col_plan(int H, int XBF) {
  if (H == 256 && XBF == 16) return kColPlan256;
  if (H == 128 && XBF == 32) return kColPlan128;
  if (H == 160 && XBF == 16) return kColPlan160;
  if (H == 192 && XBF == 16) return kColPlan192;
  if (H == 224 && XBF == 16) return kColPlan224;
  return kColPlanRt;
}

void make_fft_plan(int L, FftPlan &p, void **d_tw) {
  AV_CHECK(L >= 1 && L <= kMaxFftLen, "FFT length out of range (1..1024)");
  p = FftPlan{};
  p.L = L;
  int m = L, ns = 1;
  const int order[] = {16, 8, 4, 2, 3, 5, 7, 11, 13};
  for (int r : order) {
    while (m % r == 0) {
      AV_CHECK(p.nst < kMaxStages, "FFT plan too long");
      p.radix[p.nst] = r;
      p.ns[p.nst] = ns;
      ns *= r;
      m /= r;
      p.nst++;
    }
  }
  AV_CHECK(m == 1, "FFT length must factor into 2, 3, 5, 7, 11 and 13");
  std::vector<float2> tw(L);
  for (int k = 0; k < L; ++k) {
    const double a = -2.0 * M_PI * (double)k / (double)L;
    tw[k] = make_float2((float)std::cos(a), (float)std::sin(a));
  }
  AV_HIP(hipMalloc(d_tw, sizeof(float2) * L));
  AV_HIP(hipMemcpy(*d_tw, tw.data(), sizeof(float2) * L, hipMemcpyHostToDevice));
  AV_HIP(hipStreamSynchronize(nullptr));  // landed before any stream reads it
  p.tw = (const float2 *)*d_tw;
}

CartPlan::~CartPlan() {
  for (void *p : {d_row_off, d_rows, d_mrows, d_rowinfo, d_rowmap, d_twx, d_twy})
    if (p) (void)hipFree(p);
}

void CartPlan::init(int W, int H, int C, int T, const float *mask, int centered, hipStream_t s) {
  AV_CHECK(W >= 1 && H >= 1 && C >= 1 && T >= 1, "invalid operator dimensions");
  AV_CHECK(H <= kMaxFftLen && W <= kMaxFftLen, "image dimensions exceed the longest FFT (1024)");
  g = CartGeom{};
  g.W = W; g.H = H; g.C = C; g.T = T;
  g.centered = centered ? 1 : 0;
  g.XB = pick_xb(H, W);
  g.RB = std::max(1, std::min(kRowMaxRows, kRowTileElems / W));
  g.nxb = (W + g.XB - 1) / g.XB;
  g.XBF = pick_xb(H, W, kColTileElems);
  g.nxbf = (W + g.XBF - 1) / g.XBF;
  g.cplan = pick_col_plan(H, g.XBF);
  g.rplan = W == 256 ? kRowPlan256 : (W == 128 ? kRowPlan128 : kRowPlanRt);
  g.RBF = g.rplan == kRowPlan256 ? 16 : (g.rplan == kRowPlan128 ? 32 : g.RB);
  g.scale = (float)(1.0 / std::sqrt((double)H * W));
  const long HW = (long)H * W;
  std::vector<float> hm;
  has_mask = mask != nullptr;
  if (mask) {
    hm.resize((size_t)HW * T);
    AV_HIP(hipMemcpyAsync(hm.data(), mask, sizeof(float) * hm.size(), hipMemcpyDeviceToHost, s));
    AV_HIP(hipStreamSynchronize(s));
  }
  h_row_off.assign(T + 1, 0);
  h_rows.clear();
  std::vector<float> mrows;
  mask_norm2 = 0.0;
  for (int t = 0; t < T; ++t) {
    for (int y = 0; y < H; ++y) {
      bool any = !mask;
      if (mask)
        for (int x = 0; x < W; ++x) {
          const float v = hm[(size_t)t * HW + (size_t)y * W + x];
          mask_norm2 += (double)v * v;
          if (v != 0.f) any = true;
        }
      if (any) {
        h_rows.push_back(y);
        for (int x = 0; x < W; ++x) mrows.push_back(mask ? hm[(size_t)t * HW + (size_t)y * W + x] : 1.f);
      }
    }
    h_row_off[t + 1] = (int)h_rows.size();
  }
  g.nrows = (long)h_rows.size();
  g.nrows_c = g.nrows * C;
  std::vector<int> rowmap((size_t)T * H, -1);
  for (int t = 0; t < T; ++t)
    for (int k = h_row_off[t]; k < h_row_off[t + 1]; ++k) rowmap[(size_t)t * H + h_rows[k]] = k - h_row_off[t];
  std::vector<int4> info;
  info.reserve(g.nrows_c);
  for (int t = 0; t < T; ++t) {
    const int r0 = h_row_off[t], S = h_row_off[t + 1] - r0;
    for (int c = 0; c < C; ++c)
      for (int k = 0; k < S; ++k) info.push_back(make_int4(t, c, r0 + k, h_rows[r0 + k]));
  }
  auto up = [&](void **dst, const void *src, size_t bytes) {
    AV_HIP(hipMalloc(dst, bytes > 0 ? bytes : 16));
    if (bytes) AV_HIP(hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice));
  };
  up(&d_row_off, h_row_off.data(), sizeof(int) * h_row_off.size());
  up(&d_rows, h_rows.data(), sizeof(int) * h_rows.size());
  up(&d_mrows, mrows.data(), sizeof(float) * mrows.size());
  up(&d_rowinfo, info.data(), sizeof(int4) * info.size());
  up(&d_rowmap, rowmap.data(), sizeof(int) * rowmap.size());
  g.row_off = (const int *)d_row_off;
  g.rows = (const int *)d_rows;
  g.mrows = (const float *)d_mrows;
  g.rowinfo = (const int4 *)d_rowinfo;
  g.rowmap = (const int *)d_rowmap;
  make_fft_plan(W, g.pw, &d_twx);
  make_fft_plan(H, g.ph, &d_twy);
  g.wide_col = plan_is_wide(g.ph) ? 1 : 0;
  g.wide_row = plan_is_wide(g.pw) ? 1 : 0;
  // The tables above come from pageable host memory: such a copy (on the
  // null stream) may return before its data has landed, and the kernels that
  // index with them may run on non-blocking streams, which do not wait for
  // the null stream.  Waiting for the null stream alone leaves those streams'
  // own work running.
  AV_HIP(hipStreamSynchronize(nullptr));
}

void launch_col_forward(const CartGeom &g, const float2 *u, const float2 *usub, const float2 *b1,
                        float2 *kcol, hipStream_t s) {
  hipLaunchKernelGGL(k_col_forward, dim3(g.nxb, g.T), dim3(NT), 0, s, g, u, usub, b1, kcol);
  AV_LAUNCH_CHECK();
}

void launch_row_pass(int mode, const CartGeom &g, const float2 *kcol, float2 *z, const float2 *d,
                     float2 *zhat, float2 *out, const float2 *zin, const DevScalars *sc,
                     double *partial, hipStream_t s) {
  const int nb = (int)((g.nrows_c + g.RB - 1) / g.RB);
  if (nb == 0) return;
  dim3 grid(nb), block(NT);
  switch (mode) {
    case kRowApplyK: hipLaunchKernelGGL(k_row_pass<kRowApplyK>, grid, block, 0, s, g, kcol, z, d, zhat, out, zin, sc, partial); break;
    case kRowDualUpdate: hipLaunchKernelGGL(k_row_pass<kRowDualUpdate>, grid, block, 0, s, g, kcol, z, d, zhat, out, zin, sc, partial); break;
    case kRowNorm2: hipLaunchKernelGGL(k_row_pass<kRowNorm2>, grid, block, 0, s, g, kcol, z, d, zhat, out, zin, sc, partial); break;
    case kRowAdjDense: hipLaunchKernelGGL(k_row_pass<kRowAdjDense>, grid, block, 0, s, g, kcol, z, d, zhat, out, zin, sc, partial); break;
    case kRowResidual: hipLaunchKernelGGL(k_row_pass<kRowResidual>, grid, block, 0, s, g, kcol, z, d, zhat, out, zin, sc, partial); break;
    case kRowAdjCompact: hipLaunchKernelGGL(k_row_pass<kRowAdjCompact>, grid, block, 0, s, g, kcol, z, d, zhat, out, zin, sc, partial); break;
    default: AV_CHECK(false, "bad row-pass mode");
  }
  AV_LAUNCH_CHECK();
}

void launch_col_adjoint(const CartGeom &g, const float2 *zhat, const float2 *b1, float2 *out,
                        hipStream_t s) {
  hipLaunchKernelGGL(k_col_adjoint, dim3(g.nxb, g.T), dim3(NT), 0, s, g, zhat, b1, out);
  AV_LAUNCH_CHECK();
}

void launch_compact_gather(const CartGeom &g, const float2 *dense, float2 *compact, hipStream_t s) {
  const long n = g.nrows_c * g.W;
  if (n == 0) return;
  hipLaunchKernelGGL(k_compact_gather, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, g, dense,
                     compact);
  AV_LAUNCH_CHECK();
}

void launch_fft2(float2 *data, const FftPlan &pw, const FftPlan &ph, int H, int W, int batch,
                 int sign, int centered, float scale, hipStream_t s) {
  const int RB = EMAX / W;
  float2 *tmp = nullptr;
  AV_HIP(hipMallocAsync((void **)&tmp, sizeof(float2) * (size_t)H * W * batch, s));
  hipLaunchKernelGGL(k_fft_rows_oop, dim3((H + RB - 1) / RB, batch), dim3(NT), 0, s, data, tmp, pw, H,
                     W, RB, sign, centered);
  AV_LAUNCH_CHECK();
  const int XB = pick_xb(H, W);
  hipLaunchKernelGGL(k_fft_cols, dim3((W + XB - 1) / XB, batch), dim3(NT), 0, s, tmp, ph, H, W, XB,
                     sign, centered, scale);
  AV_LAUNCH_CHECK();
  AV_HIP(hipMemcpyAsync(data, tmp, sizeof(float2) * (size_t)H * W * batch, hipMemcpyDeviceToDevice, s));
  AV_HIP(hipFreeAsync(tmp, s));
}

// In-place transform along the slowest axis of `batch` arrays [L][plane]:
// one length-L transform per column of the plane (the depth axis of a volume).
void launch_fft_axis0(float2 *data, const FftPlan &p, int L, int plane, int batch, int sign,
                      float scale, hipStream_t s) {
  const int XB = pick_xb(L, plane);
  hipLaunchKernelGGL(k_fft_cols, dim3((plane + XB - 1) / XB, batch), dim3(NT), 0, s, data, p, L,
                     plane, XB, sign, 0, scale);
  AV_LAUNCH_CHECK();
}

}  // namespace avionic
