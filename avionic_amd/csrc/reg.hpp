// Regulariser kernels of the PD loop: the fused dual update (stencils +
// ascent + norm_3 / norm_6 projections), the fused primal update (divergences
// + descent + extragradient), the step-size reductions and the functional /
// gap reductions, for TV, TGV2 and ICTGV2.
#pragma once

#include "common.hpp"

namespace avionic {

// Kernel methods.  ICTV (ictv.cpp) keeps u, w and the duals y1(3), y3(3).
enum Method { kTV = 0, kTGV2 = 1, kICTGV2 = 2, kICTV = 3 };
// Method numbers of the C-ABI (avionic_method) that differ from the kernel
// method: temporal TV (tv_temp.cpp) runs on the TV kernels with zero spatial
// weights (1/dx = 1/dy = 0).  Its spatial dual components then stay exactly
// zero, so the norm_3 projection, the divergence and the step-rule sums equal
// the one-component temporal ones of the reference bit for bit.
// TGV2_3D (tgv2_3d.cpp) is TGV2 on a (W, H, depth) grid with (dx, dy, dz).
constexpr int kApiTVTEMP = 3, kApiICTV = 4, kApiTGV2_3D = 5;
inline int kernel_method(int api) {
  return api == kApiTVTEMP ? kTV : api == kApiICTV ? kICTV : api == kApiTGV2_3D ? kTGV2 : api;
}

inline int primal_components(int m) {
  return m == kTV ? 1 : m == kTGV2 ? 4 : m == kICTV ? 2 : 8;
}
inline int dual_components(int m) {
  return m == kTV ? 3 : m == kTGV2 ? 9 : m == kICTV ? 6 : 18;
}

struct Grid3 {
  int W, H, T;
  long N;
};

// Voxel kernels index inside one component with 32 bits of byte offset.
constexpr long kMaxVoxels = 1L << 29;
inline void check_grid(const Grid3 &g) {
  AV_CHECK(g.W >= 1 && g.H >= 1 && g.T >= 1 && g.N < kMaxVoxels,
           "width * height * frames must be below 2^29 voxels");
}

// Inverse step sizes of the two derivative sets and the projection radii
// (radius = 0 means "no projection").
struct RegParams {
  float idx, idy, idt;     // 1/dx, 1/dy, 1/dt
  float idx2, idy2, idt2;  // second TGV (ICTGV2)
  float r1, r2, r3, r4;    // radii of y1 (|.|_3), y2 (|.|_6), y3, y4
};

RegParams make_reg_params(int method, const double d6[6], double alpha0, double alpha1,
                          double alpha);

// y += sigma K_reg(ext); projections.  ds: dual storage (DualStorage).
void launch_dual(int method, int ds, const Grid3 &g, const RegParams &rp,
                 const float2 *ext, void *y, const DevScalars *sc, hipStream_t s);
// x+ = x - tau (K_reg^T y + [khz, 0..]); ext = 2 x+ - x; x = x+.
void launch_primal(int method, int ds, const Grid3 &g, const RegParams &rp, float2 *x,
                   float2 *ext, const void *y, const float2 *khz, const DevScalars *sc,
                   hipStream_t s);
// Step-rule partial sums over the primal difference ext - x:
// partial[2b] = weighted sum |K_reg delta|^2, partial[2b+1] = sum |delta|^2.
int step_blocks(const Grid3 &g);
// Blocks of the two-voxels-per-lane dual / primal updates (2 RT voxels each).
int pair_blocks(const Grid3 &g);
// Partial sums written by launch_reg_norm / launch_gstar_terms (one per
// RT consecutive voxels).
int norm_blocks(const Grid3 &g);
void launch_step_reg(int method, const Grid3 &g, const RegParams &rp, const float2 *ext,
                     const float2 *x, double *partial, hipStream_t s);
// Final deterministic reduction + sigma/tau rule (pd_recon.cpp:40-61).
void launch_step_finalize(const double *reg_partial, int nreg, const double *k_partial, int nk,
                          DevScalars *sc, double *sig_log, double *tau_log, int log_index,
                          hipStream_t s);
// Bump the iteration counter.
void launch_iter_bump(DevScalars *sc, hipStream_t s);

// Functional value R(x) (TV / TGV2 / ICTGV2 norms), per-block partials.
void launch_reg_norm(int method, const Grid3 &g, const RegParams &rp, double alpha0, double alpha1,
                     double alpha, const float2 *x, double *partial, hipStream_t s);
// G* dual residual terms: l1 norms of (khz - div1 y1), (y1 + divsym1 y2), ...
// y in dual storage ds (DualStorage), decoded on the device.
void launch_gstar_terms(int method, int ds, const Grid3 &g, const RegParams &rp,
                        const float2 *khz, const void *y, double *partial, hipStream_t s);
// sum over compact k-space of Re(conj(d) z) and |z|^2 (two partial arrays).
void launch_kspace_dots(const float2 *d, const float2 *z, long n, double *partial, hipStream_t s);
int reduce_blocks(long n);
double reduce_host(const double *partial, int n, int stride, int offset, hipStream_t s);

// Standalone utilities (utils::Gradient etc.).
void launch_gradient(const Grid3 &g, float idx, float idy, float idt, const float2 *u, float2 *out,
                     hipStream_t s);
void launch_symgrad(const Grid3 &g, float idx, float idy, float idt, const float2 *v, float2 *out,
                    hipStream_t s);
void launch_divergence(const Grid3 &g, float idx, float idy, float idt, const float2 *p,
                       float2 *out, hipStream_t s);
void launch_symdiv(const Grid3 &g, float idx, float idy, float idt, const float2 *q, float2 *out,
                   hipStream_t s);
void launch_prox(int ncomp, float2 *y, long n, float scale, hipStream_t s);

}  // namespace avionic
