// Operator norm of the full primal-dual operator K_full = [K_reg; K] in FP64
// by power iteration (the step-size estimate of matlab/getopnorm.m:1-16, on
// the stacked regulariser + data operator), for the fixed-step mode
// sigma = tau = 0.99 / ||K_full||.
//
// The iteration runs on M = K_full^* K_full:
//   M x = K_reg^* K_reg x + [K^H K u, 0, ...]
// where K_reg^* is the adjoint for the dual pairing the solver uses (the
// off-diagonal symmetric-gradient entries weighted 2), i.e. the operator of
// the primal descent.  Everything (iterate, stencils, transforms, coil maps
// promoted from float, reductions) is double precision on the device.
#pragma once

#include "reg.hpp"

namespace avionic {

class Operator;

// FP64 work buffers of the data operator (grown on demand).
struct Fp64Scratch {
  void *k = nullptr;   // coil images / k-space, double2
  size_t k_bytes = 0;
  void *red = nullptr;  // reduction partials
  size_t red_bytes = 0;
  ~Fp64Scratch();
  void ensure_k(size_t bytes);
  void ensure_red(size_t bytes);
};

// Deterministic start vector of the power iteration: component j (real and
// imaginary parts of each complex entry in turn) is a splitmix64 hash of j
// mapped to [-1, 1), so the oracle draws the identical vector.
constexpr unsigned long long kOpnormSeed = 0x9e3779b97f4a7c15ull;
// Power-iteration steps of the fixed-step mode when none are given.
constexpr int kOpnormDefaultIterations = 50;

// ||K_full|| after `iterations` steps (Rayleigh quotient sqrt(<M x, x>) with
// x of unit norm).  d6 = {dx, dy, dt, dx2, dy2, dt2} as the solver uses them.
double operator_norm_fp64(int method, int W, int H, int T, const double d6[6],
                          const Operator &op, const float2 *b1, int iterations, hipStream_t s);

// FP64 K^H K of the Cartesian 2D+t operator (out += K^H K u), and of the
// volumetric one; `mask` as the operator stores it.
struct CartGeom;
// Double-precision gridding tables of the non-Cartesian operator: per point
// (frame-major) its first grid column / row (wrapped), kNcTaps column then
// row weights, and sqrt(density) / sqrt(HW); per cell entry of the spreading
// table its weight; the deapodisation per image column / row.
struct Nc64 {
  int W, H, C, T, M, Gx, Gy;
  const int2 *q;        // [T][M]
  const double *w;      // [T][M][2 * kNcTaps]
  const double *sw;     // [T][M]
  const double *dxi, *dyi;
  const int *cptr;      // [T][Gx*Gy+1] (the float operator's table)
  const int2 *ce;       // (point, -) per entry (the float operator's table)
  const double *ce_w;   // weight per entry
};
void nc_khk_fp64(const Nc64 &g, const double2 *u, double2 *out, const float2 *b1,
                 Fp64Scratch &ws, hipStream_t s);
void cart_khk_fp64(const CartGeom &g, const double2 *u, double2 *out, const float2 *b1,
                   Fp64Scratch &ws, hipStream_t s);
void cart3_khk_fp64(int W, int H, int D, int C, const float *mask, const double2 *u, double2 *out,
                    const float2 *b1, Fp64Scratch &ws, hipStream_t s);

}  // namespace avionic
