// Non-Cartesian operator pair by Kaiser-Bessel gridding, as GPU passes.
//
//   (K u)_{t,c,m} = s_m / sqrt(HW) sum_{y,x} b1_c u_t exp(-2 pi i k_m . (r - N/2))
//
// s_m = sqrt(density), k_m in cycles per pixel.  The sum is evaluated on a
// grid of Gx x Gy = round(osf W) x round(osf H) samples: the image, divided by
// the kernel's Fourier transform (deapodisation) and centred at grid index 0,
// is transformed, and each trajectory point interpolates its J x J
// neighbourhood with separable Kaiser-Bessel weights.
//
// K as three passes, none of which writes the oversampled grid:
//   1. column pass (grid (ceil(W/XB), T, CG), a coil loop over each of the CG
//      coil groups inside): image columns -> zero-padded length-Gy transforms
//      -> tmp [t][c][Gy][W];
//   2. row pass (grid (ceil(Gy/RB), T*C); the kernels compiled for grid length
//      384 take two coils of a frame per block): zero-padded length-Gx
//      transforms of RB grid rows in LDS; every trajectory point touching one
//      of those rows reads its J neighbours there and writes one partial sum
//      per row it touches -> part [t][c][M][4];
//   3. sample pass: the four partials of a point, scaled, feed K u (or the z
//      update, or a norm reduction).
// K^H as two passes:
//   1. column pass (grid (ceil(Gx/XB), T*C); the compiled kernels take three
//      coils of a frame per block): the spreading, by gathers without
//      atomics: the lanes form the products of the column tile's cell ->
//      point entries (tce) into LDS, each grid cell sums its own in entry
//      order; length-Gy inverse transform, keeps the H image rows ->
//      tmp [t][c][H][Gx];
//   2. row pass (grid (ceil(H/RB), T), coil loop inside): length-Gx inverse
//      transforms, crop to W, conj(b1_c) coil sum in registers, deapodise.
// Every output element is written by exactly one lane in a fixed order, so
// both directions are deterministic.
#pragma once

#include "common.hpp"

namespace avionic {

constexpr int kNcTaps = 4;  // grid rows / columns per trajectory point (J <= 3)

struct NcGeom {
  int W, H, C, T, M;  // image, coils, frames, points per frame
  int Gx, Gy;         // oversampled grid
  int XBf;            // image columns per forward column-pass block
  int XBa;            // grid columns per adjoint column-pass block
  int RBf;            // grid rows per forward row-pass block
  int RBa;            // image rows per adjoint row-pass block
  int CG;             // coil groups of the forward column pass (blocks per tile)
  int wide;           // a grid plan is wide (plan_is_wide): the wide kernel variants
  float scale;        // 1 / sqrt(HW)
  const float *dxi, *dyi;  // deapodisation 1 / psi_hat per image column / row
  const float *sw;         // [T][M] sqrt(density) / sqrt(HW)
  const int *rptr;         // [T][Gy+1] row -> (point, tap) list offsets
  const int4 *rhdr;        // per entry: m * kNcTaps + dy, grid row, first grid column
                           // (wrapped), bits of the row weight wy[dy]
  const float4 *rwx;       // per entry: the point's kNcTaps column weights
  const int *cptr;         // [T][Gx*Gy+1] cell -> point list offsets
  const int2 *ce;          // per cell entry: point index m, and the bits of its
                           // weight * sqrt(density) / sqrt(HW)
  // The same entries grouped by adjoint column tile (XBa grid columns): for
  // frame t and tile X (nXa = ceil(Gx / XBa) tiles), the entries of cells
  // (r, X XBa + b), r < Gy, b < XBa, in cell order k = r XBa + b, start at
  // tce[tstart[t nXa + X]]; tcell[(t nXa + X) (XBa Gy + 1) + k] is cell k's
  // first entry relative to that start (the last value: the tile's count).
  const int *tstart;
  const int *tcell;
  const int2 *tce;
  FftPlan px, py;          // lengths Gx, Gy
};

enum NcSampleMode {
  kNcApplyK = 0,    // out = K u
  kNcDualUpdate = 1,  // z = (z + sigma K u - sigma d) / (1 + sigma/lambda)
  kNcNorm2 = 2,     // partial sums of |K u|^2
  kNcResidual = 3,  // partial sums of |K u - d|^2
};

// K (u - usub): column and row passes into `part` ([T][C][M][4]).
void launch_nc_forward_grid(const NcGeom &g, const float2 *u, const float2 *usub,
                            const float2 *b1, float2 *tmp, float2 *part, hipStream_t s);
// Sample pass; returns the number of partial sums written (reduction modes).
int launch_nc_samples(int mode, const NcGeom &g, const float2 *part, float2 *out, float2 *z,
                      const float2 *d, const DevScalars *sc, double *partial, hipStream_t s);
int nc_sample_blocks(const NcGeom &g);
// K^H z.
void launch_nc_adjoint(const NcGeom &g, const float2 *z, const float2 *b1, float2 *tmp,
                       float2 *img, hipStream_t s);

}  // namespace avionic
