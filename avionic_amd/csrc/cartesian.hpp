// Coil-fused masked Cartesian operator pair (K, K^H) on compact k-space.
//
// K  (image -> k-space): z[t,c] = M_t . F^{-1}_ortho(b1_c . u_t)
// K^H (k-space -> image): u_t   = sum_c conj(b1_c) . F_ortho(M_t . z[t,c])
// (CUDA/src/cartesian_operator.cpp:65-160; matlab/backward_opt.m,
// forward_opt.m).
//
// k-space is stored compactly: only the phase-encode rows that carry a nonzero
// mask value are kept, as [t][c][k][x] with k over the sampled rows of frame t.
// A 2D transform is split into a column pass (along y, LDS tile of XB
// columns, coils looped inside the workgroup so the image is read once) and a
// row pass (along x, only sampled rows).  The dual update of z, its prox and
// the adjoint's row transform are fused into the forward row pass.
#pragma once

#include <vector>

#include "common.hpp"

namespace avionic {

struct CartGeom {
  int W, H, C, T;
  int centered;
  int XB;  // columns per column-pass tile (API kernels)
  int RB;  // rows per row-pass tile
  int nxb;
  int XBF, nxbf;  // columns per tile / tiles of the hot loop's fused column passes
  int cplan;      // compile-time plan of the fused column passes (ColPlanId)
  int rplan;      // compile-time plan of the fused row pass (RowPlanId)
  int RBF;        // rows per block of the fused row pass
  int wide_col, wide_row;  // column / row plan is wide (plan_is_wide)
  long nrows;          // sampled rows over all frames (row_off[T])
  long nrows_c;        // C * nrows
  const int *row_off;  // [T+1]
  const int *rows;     // [nrows] ky of each sampled row, frame-major
  const float *mrows;  // [nrows][W] mask values of the sampled rows
  const int4 *rowinfo; // [nrows_c] (t, c, r = row_off[t] + k, ky) per compact row
  const int *rowmap;   // [T][H] compact row index k of (t, ky), -1 if not sampled
  float scale;         // 1 / sqrt(H W)
  FftPlan pw, ph;
};

// Compile-time column plans (transform length H, XBF columns per tile); any
// other geometry runs the runtime plan.
enum ColPlanId { kColPlanRt = 0, kColPlan256, kColPlan128, kColPlan160, kColPlan192, kColPlan224 };
// Kind of a geometry's fused column passes: 0 runtime plan, 1 compile-time
// plan, 2 wide runtime plan (one kernel per kind).
constexpr int kColKinds = 3;
__host__ __device__ inline int col_kind(const CartGeom &g) {
  return g.cplan != kColPlanRt ? 1 : (g.wide_col ? 2 : 0);
}
// Compile-time row plans (row length W); rows per block 16 and 32.
enum RowPlanId { kRowPlanRt = 0, kRowPlan256, kRowPlan128 };
// Kind of the fused row pass: 0 compile-time plan, 1 runtime plan, 2 wide
// runtime plan (one kernel per kind).
constexpr int kRowKinds = 3;
__host__ __device__ inline int row_kind(const CartGeom &g) {
  return g.rplan != kRowPlanRt ? 0 : (g.wide_row ? 2 : 1);
}

// Row-pass modes.
enum RowMode {
  kRowApplyK = 0,     // out = K(u) rows (compact or dense)
  kRowDualUpdate = 1, // z = (z + s K u - s d)/(1 + s/l); zhat = F_row(M z)
  kRowNorm2 = 2,      // partial sums of |K u|^2
  kRowAdjDense = 3,   // zhat = F_row(M z_dense)
  kRowResidual = 4,   // partial sums of |K u - d|^2
  kRowAdjCompact = 5, // zhat = F_row(M z_compact)
};

class FftPlanCache;

// Host-side owner of the mask compaction and FFT plans of one geometry.
struct CartPlan {
  CartGeom g{};
  std::vector<int> h_row_off, h_rows;
  void *d_row_off = nullptr, *d_rows = nullptr, *d_mrows = nullptr, *d_rowinfo = nullptr,
       *d_rowmap = nullptr;
  void *d_twx = nullptr, *d_twy = nullptr;
  double mask_norm2 = 0.0;  // ||mask||^2 (AdaptLambda)
  bool has_mask = false;

  CartPlan() = default;
  CartPlan(const CartPlan &) = delete;
  CartPlan &operator=(const CartPlan &) = delete;
  ~CartPlan();
  // mask: device pointer [T][H][W] or nullptr.
  void init(int W, int H, int C, int T, const float *mask, int centered, hipStream_t s);
  long kspace_elems() const { return g.nrows_c * (long)g.W; }
  int row_blocks() const { return (int)((g.nrows_c + g.RB - 1) / g.RB); }
  int fused_row_blocks() const { return (int)((g.nrows_c + g.RBF - 1) / g.RBF); }
};

void make_fft_plan(int L, FftPlan &p, void **d_tw);

// ---- launchers (all asynchronous on `s`) ----
// Column pass of K: kcol = F^{-1}_y(b1 . (u - usub)) on the sampled rows.
void launch_col_forward(const CartGeom &g, const float2 *u, const float2 *usub, const float2 *b1,
                        float2 *kcol, hipStream_t s);
// Row pass (mode-dependent arguments may be null).
void launch_row_pass(int mode, const CartGeom &g, const float2 *kcol, float2 *z, const float2 *d,
                     float2 *zhat, float2 *out, const float2 *zin, const DevScalars *sc,
                     double *partial, hipStream_t s);
// Column pass of K^H with the coil sum: out = scale * sum_c conj(b1_c) F_y(zhat).
void launch_col_adjoint(const CartGeom &g, const float2 *zhat, const float2 *b1, float2 *out,
                        hipStream_t s);
// Gather dense [t][c][H][W] k-space rows into compact storage (times mask).
void launch_compact_gather(const CartGeom &g, const float2 *dense, float2 *compact, hipStream_t s);
// Standalone batched 2D FFT (in place).
void launch_fft2(float2 *data, const FftPlan &pw, const FftPlan &ph, int H, int W, int batch,
                 int sign, int centered, float scale, hipStream_t s);
// In-place transform along the slowest axis of `batch` arrays [L][plane].
void launch_fft_axis0(float2 *data, const FftPlan &p, int L, int plane, int batch, int sign,
                      float scale, hipStream_t s);

}  // namespace avionic
