// Kaiser-Bessel gridding passes of the non-Cartesian operator (noncart.hpp).
#include <algorithm>

#include "fft_device.hpp"
#include "noncart.hpp"

namespace avionic {

namespace {

#ifndef AVIONIC_NC_COLF_THREADS
#define AVIONIC_NC_COLF_THREADS 256
#endif
#ifndef AVIONIC_NC_COLA_THREADS
#define AVIONIC_NC_COLA_THREADS 512
#endif
constexpr int NCF = AVIONIC_NC_COLF_THREADS;  // threads of a forward column-pass block
constexpr int NCC = AVIONIC_NC_COLA_THREADS;  // threads of an adjoint column-pass block
constexpr int NCR = 256;   // threads of a row-pass / sample-pass block
constexpr int NCE = 4096;  // complex elements of an LDS transform tile
// Twiddle-table entries and per-lane outputs of the narrow / wide kernel
// variants (plan_is_wide: a grid longer than 512 or a radix-11 / -13 stage).
constexpr int nc_table_len(bool wide) { return wide ? kMaxFftLen : kMaxFftLenNarrow; }
constexpr int row_slots_nc(bool wide) { return max_slots(NCE, NCR, wide); }

__device__ inline int wrap_up(int v, int G) { return v >= G ? v - G : v; }

// Grid transforms: on the geometry's runtime plans (NcPlanRt), or compiled
// for a grid length in both directions, with the radices in make_fft_plan's
// order and XB grid columns per column-pass block (pick_cols at that length):
// 384 (a 256 x 256 image at oversampling 1.5; 16, 8, 3), 512 (256 x 256 at
// 2, the reference's default oversampling; 16, 16, 2), 192 and 256 (128 x 128
// at 1.5 and 2; 16, 4, 3 and 16, 16).  nc_ct_len picks the compiled kernels;
// they run the runtime plan's stages with the index arithmetic folded into
// constants.
struct NcPlanRt {
  static constexpr int L = 0, XB = 0;
};
template <class F, int XB_>
struct NcCt : F {
  static constexpr int XB = XB_;
  static constexpr int RL = F::R2 > 1 ? F::R2 : F::R1;  // last radix
};
using NcPlan384 = NcCt<FftCt<384, 16, 8, 3>, 8>;
using NcPlan512 = NcCt<FftCt<512, 16, 16, 2>, 8>;
using NcPlan192 = NcCt<FftCt<192, 16, 4, 3>, 16>;
using NcPlan256 = NcCt<FftCt<256, 16, 16>, 16>;

// ------------------------------------------------------------ K, column pass
// The first stage's input b1_c (u - usub) dxi dyi, zero outside the image.
// With SLOTS (the compiled kernels) a lane keeps (u - usub) dxi dyi of its
// first-stage elements in registers from the block's first coil (`first`)
// and loads only the coil map for the others: the lane -> element mapping of
// a stage is the same for every coil.  The product is formed the same way
// either way.
template <int SLOTS>
struct NcColFwdLoad {
  const float2 *ut, *us, *bc;
  const float *dxi, *dyi;
  int W, H, Gy, xb0;
  bool delta, first;
  float2 img[SLOTS > 0 ? SLOTS : 1];
  __device__ float2 operator()(int b, int i, int slot) {
    const int x = xb0 + b;
    if (x >= W) return czero();
    const int y = wrap_up(i + H / 2, Gy);
    if (y >= H) return czero();
    const unsigned si = (unsigned)(y * W + x);
    if constexpr (SLOTS > 0) {
      if (first) img[slot] = cscale(delta ? csub(ut[si], us[si]) : ut[si], dxi[x] * dyi[y]);
      return cmul(bc[si], img[slot]);
    } else {
      const float2 v = delta ? csub(ut[si], us[si]) : ut[si];
      return cmul(bc[si], cscale(v, dxi[x] * dyi[y]));
    }
  }
};
// First-stage elements per lane of the compiled column forward pass.
template <class PL>
constexpr int nc_colf_slots() {
  if constexpr (PL::L == 0)
    return 0;
  else
    return (PL::XB * PL::L / PL::R0 + AVIONIC_NC_COLF_THREADS - 1) / AVIONIC_NC_COLF_THREADS * PL::R0;
}
struct NcColFwdStore {
  float2 *tc;  // tmp rows of (t, c): [Gy][W]
  int W, xb0;
  __device__ void operator()(int b, int i, float2 v, int) {
    const int x = xb0 + b;
    if (x < W) tc[(unsigned)(i * W + x)] = v;
  }
};

template <bool WIDE, class PL>
__global__ __launch_bounds__(NCF) void k_nc_col_fwd(NcGeom g, const float2 *__restrict__ u,
                                                    const float2 *__restrict__ usub,
                                                    const float2 *__restrict__ b1,
                                                    float2 *__restrict__ tmp) {
  __shared__ float2 tile[NCE];
  __shared__ float2 twl[nc_table_len(WIDE)];
  const int tid = threadIdx.x;
  const int t = blockIdx.y;
  const long HW = (long)g.H * g.W;
  constexpr bool CT = PL::L != 0;
  const int XB = CT ? PL::XB : g.XBf;
  const float2 *tw = stage_twiddles(g.py, twl, tid, NCF);
  __syncthreads();
  NcColFwdLoad<nc_colf_slots<PL>()> ld;
  ld.ut = u + t * HW;
  ld.us = usub ? usub + t * HW : nullptr;
  ld.delta = usub != nullptr;
  ld.dxi = g.dxi;
  ld.dyi = g.dyi;
  ld.W = g.W;
  ld.H = g.H;
  ld.Gy = CT ? PL::L : g.Gy;
  ld.xb0 = blockIdx.x * XB;
  NcColFwdStore st{nullptr, g.W, (int)blockIdx.x * XB};
  // coils [c0, c1) of this block (blockIdx.z picks the coil group)
  const int cpg = (g.C + gridDim.z - 1) / gridDim.z;
  const int c0 = blockIdx.z * cpg, c1 = min(g.C, c0 + cpg);
  for (int c = c0; c < c1; ++c) {
    ld.first = c == c0;
    ld.bc = b1 + c * HW;
    st.tc = tmp + ((long)t * g.C + c) * g.Gy * g.W;
    if constexpr (CT)
      fft_io_ct<NCF, PL, PL::XB, true, true, true>(tile, PL::XB, 1, PL::XB, -1.f, tid, tw,
                                                      ld, st);
    else
      fft_io<NCF, NCE, true, true, true, WIDE>(tile, g.py, XB, 1, XB, -1.f, tid, tw, ld, st);
    __syncthreads();
  }
}

// --------------------------------------------------------------- K, row pass
// Transform b of a block: coil cz + b / RB, grid row gy0 + b % RB of that
// coil's tmp plane (zero past the block's coils or rows).
template <int NZ>
struct NcRowFwdLoad {
  const float2 *tr;  // tmp row gy0 of plane (t, cz): [.][W]; coil planes GyW apart
  long plane;        // Gy W
  int W, Gx, RB, nr, nz;
  __device__ float2 operator()(int b, int i, int) const {
    const int x = wrap_up(i + W / 2, Gx);
    if constexpr (NZ == 1) {
      return x < W ? tr[(unsigned)(b * W + x)] : czero();
    } else {
      const int q = b / RB, r = b - q * RB;
      return x < W && q < nz && r < nr ? tr[q * plane + (unsigned)(r * W + x)] : czero();
    }
  }
};

// NZ coils of a frame per block (2 in the compiled kernels, RB = 5 grid rows
// each): the tile holds NZ x RB transformed rows, and the row table's entries
// of those rows are loaded once for all coils.  Each coil's partials are the
// one-coil partials.
template <bool WIDE, class PL, int NZ>
__global__ __launch_bounds__(NCR) void k_nc_row_fwd(NcGeom g, const float2 *__restrict__ tmp,
                                                    float2 *__restrict__ part) {
  __shared__ float2 tile[NCE];
  __shared__ float2 twl[nc_table_len(WIDE)];
  constexpr bool CT = PL::L != 0;
  static_assert(CT || NZ == 1, "coil groups need the compiled kernels");
  const int tid = threadIdx.x;
  const int RB = CT ? NCE / PL::L / NZ : g.RBf;  // grid rows per coil
  const int ngrp = (g.C + NZ - 1) / NZ;
  const int t = blockIdx.y / ngrp, cz = (blockIdx.y % ngrp) * NZ;
  const int nz = min(NZ, g.C - cz);
  const int gy0 = blockIdx.x * RB;
  const int nr = min(RB, g.Gy - gy0);
  const int Gx = CT ? PL::L : g.Gx;
  const float2 *tw = stage_twiddles(g.px, twl, tid, NCR);
  __syncthreads();
  const long plane = (long)g.Gy * g.W;
  NcRowFwdLoad<NZ> ld{tmp + ((long)t * g.C + cz) * plane + (long)gy0 * g.W, plane, g.W, Gx, RB, nr,
                  nz};
  NoStore ns;
  if constexpr (CT)
    fft_io_ct<NCR, PL, NCE / PL::L, false, true, false>(tile, NZ * RB, PL::L, 1, -1.f, tid, tw, ld,
                                                        ns);
  else
    fft_io<NCR, NCE, false, true, false, WIDE>(tile, g.px, nr, Gx, 1, -1.f, tid, tw, ld, ns);
  // tile[(q RB + r) Gx + i] holds grid row gy0 + r of coil cz + q in natural order.
  const int *rp = g.rptr + (long)t * (g.Gy + 1);
  const int e0 = rp[gy0], e1 = rp[gy0 + nr];
  float2 *pp = part + ((long)t * g.C + cz) * g.M * kNcTaps;  // coil cz + q: M kNcTaps further
  for (int e = e0 + tid; e < e1; e += NCR) {
    // (partial index m*4 + dy, grid row, first grid column, row weight) and
    // the column weights: no load here depends on another
    const int4 h = g.rhdr[e];
    const float4 wx = g.rwx[e];
    const int qx = h.z;
#pragma unroll
    for (int q = 0; q < NZ; ++q) {
      if (q >= nz) break;
      const float2 *row = tile + (q * RB + h.y - gy0) * Gx;
      float2 acc = cscale(row[qx], wx.x);
      acc = cadd(acc, cscale(row[wrap_up(qx + 1, Gx)], wx.y));
      acc = cadd(acc, cscale(row[wrap_up(qx + 2, Gx)], wx.z));
      acc = cadd(acc, cscale(row[wrap_up(qx + 3, Gx)], wx.w));
      pp[(long)q * g.M * kNcTaps + (unsigned)h.x] = cscale(acc, __int_as_float(h.w));
    }
  }
}

// ------------------------------------------------------------- sample pass
__device__ inline double block_sum(double v, double *red) {
  const int tid = threadIdx.x;
  red[tid] = v;
  __syncthreads();
  for (int s = NCR / 2; s > 0; s >>= 1) {
    if (tid < s) red[tid] += red[tid + s];
    __syncthreads();
  }
  return red[0];
}

template <int MODE>
__global__ __launch_bounds__(NCR) void k_nc_samples(NcGeom g, const float2 *__restrict__ part,
                                                    float2 *__restrict__ out, float2 *z,
                                                    const float2 *__restrict__ d,
                                                    const DevScalars *__restrict__ sc,
                                                    double *__restrict__ partial) {
  __shared__ double red[NCR];
  const long n = (long)g.T * g.C * g.M;
  const long CM = (long)g.C * g.M;
  float sigma = 0.f, inv = 0.f;
  if (MODE == kNcDualUpdate) {
    sigma = (float)sc->sigma;
    inv = (float)(1.0 / (1.0 + sc->sigma / sc->lambda));
  }
  double acc = 0.0;
  for (long e = (long)blockIdx.x * NCR + threadIdx.x; e < n; e += (long)gridDim.x * NCR) {
    const long t = e / CM;
    const long m = e % g.M;
    const float4 *p4 = reinterpret_cast<const float4 *>(part + e * kNcTaps);
    const float4 a = p4[0], b = p4[1];
    float2 v = make_float2(a.x + a.z + b.x + b.z, a.y + a.w + b.y + b.w);
    v = cscale(v, g.sw[t * g.M + m]);
    if (MODE == kNcApplyK) {
      out[e] = v;
    } else if (MODE == kNcDualUpdate) {
      float2 zz = cadd(z[e], cscale(v, sigma));
      zz = cscale(csub(zz, cscale(d[e], sigma)), inv);
      z[e] = zz;
    } else if (MODE == kNcNorm2) {
      acc += (double)cabs2(v);
    } else {
      acc += (double)cabs2(csub(v, d[e]));
    }
  }
  if (MODE == kNcNorm2 || MODE == kNcResidual) {
    const double s = block_sum(acc, red);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
  }
}

// --------------------------------------------------------- K^H, column pass
struct NcColAdjStore {
  float2 *tp;  // tmp of the plane: [H][Gx]
  int H, Gx, Gy, gx0;
  __device__ void operator()(int b, int i, float2 v, int) {
    const int gx = gx0 + b;
    if (gx >= Gx) return;
    const int y = wrap_up(i + H / 2, Gy);
    if (y < H) tp[(unsigned)(y * Gx + gx)] = v;
  }
};

// Spreading by gathers, balanced over the lanes: a block takes its column
// tile's entries (NcGeom::tce, in cell order) in chunks; the lanes form the
// products z[m] w of a chunk, each lane NCC apart (their index and z loads
// issued together), into the LDS tile; then every cell sums the products of
// its own entries in entry order from LDS.  Dense cells at the k-space
// centre (a hundred entries) then cost LDS reads, not a chain of dependent
// global loads on one lane.  The sums keep each cell's entry order, so every
// application gives the same bits.
// NZ coils of a frame per block (3 in the compiled kernels; 2 and 4 measured
// slower at the 256x256x30x12 radial geometry): the tile's offsets are
// staged and its entries loaded once for all of them, each entry's products
// of the NZ coils sit in NZ parts of the tile, and the transforms run one
// after the other.  Every coil's sums are the one-coil
// sums.  (Running a block over all coils of a frame, one after the other,
// measured slower: 0.43 against 0.37 ms at the 256x256x30x12 radial
// geometry.)
template <bool WIDE, class PL, int NZ>
__global__ __launch_bounds__(NCC) void k_nc_col_adj(NcGeom g, const float2 *__restrict__ z,
                                                    float2 *__restrict__ tmp) {
  constexpr int CPL = NCE / NCC;  // cells per lane (XBa Gy <= NCE)
  constexpr int U = 4;            // products per lane issued together
  constexpr int CH = NCE / NZ;    // entries per chunk
  __shared__ float2 tile[NCE];
  __shared__ float2 twl[nc_table_len(WIDE)];
  extern __shared__ int cps[];  // XBa * Gy + 1 ints, sized at launch
  const int tid = threadIdx.x;
  const int ngrp = (g.C + NZ - 1) / NZ;  // coil groups per frame
  const int t = blockIdx.y / ngrp, cz = (blockIdx.y % ngrp) * NZ;
  const int nz = min(NZ, g.C - cz);  // coils of this block
  const int X = blockIdx.x, nX = gridDim.x;
  const float2 *tw = stage_twiddles(g.py, twl, tid, NCC);
  constexpr bool CT = PL::L != 0;
  const int XB = CT ? PL::XB : g.XBa, ncell = XB * (CT ? PL::L : g.Gy);
  const int *tc = g.tcell + ((long)t * nX + X) * (ncell + 1);
  for (int k = tid; k <= ncell; k += NCC) cps[k] = tc[k];
  const int2 *ent = g.tce + g.tstart[t * nX + X];
  const float2 *zp = z + ((long)t * g.C + cz) * g.M;  // coil cz + q at zp + q M
  __syncthreads();
  const int E = cps[ncell];
  float2 acc[NZ][CPL];
#pragma unroll
  for (int q = 0; q < NZ; ++q)
#pragma unroll
    for (int j = 0; j < CPL; ++j) acc[q][j] = czero();
  for (int c0 = 0; c0 < E; c0 += CH) {
    const int n = min(CH, E - c0);
    for (int e0 = 0; e0 < n; e0 += U * NCC) {
      int2 mw[U];
      float2 zv[NZ][U];
#pragma unroll
      for (int u = 0; u < U; ++u) mw[u] = ent[c0 + min(e0 + u * NCC + tid, n - 1)];
#pragma unroll
      for (int q = 0; q < NZ; ++q)
#pragma unroll
        for (int u = 0; u < U; ++u) zv[q][u] = zp[(long)(q < nz ? q : 0) * g.M + mw[u].x];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int e = e0 + u * NCC + tid;
        if (e < n)
#pragma unroll
          for (int q = 0; q < NZ; ++q) tile[q * CH + e] = cscale(zv[q][u], __int_as_float(mw[u].y));
      }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < CPL; ++j) {
      const int k = tid + j * NCC;
      if (k < ncell) {
        const int e1 = min(cps[k + 1], c0 + n);
        for (int e = max(cps[k], c0); e < e1; ++e)
#pragma unroll
          for (int q = 0; q < NZ; ++q) acc[q][j] = cadd(acc[q][j], tile[q * CH + e - c0]);
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int q = 0; q < NZ; ++q) {
    if (q >= nz) break;  // uniform over the block
#pragma unroll
    for (int j = 0; j < CPL; ++j) {
      const int k = tid + j * NCC;
      if (k < ncell) tile[k] = acc[q][j];
    }
    __syncthreads();
    NcColAdjStore st{tmp + ((long)t * g.C + cz + q) * g.H * g.Gx, g.H, CT ? PL::L : g.Gx,
                     CT ? PL::L : g.Gy, X * XB};
    NoLoad nl;
    if constexpr (CT)
      fft_io_ct<NCC, PL, PL::XB, true, false, true>(tile, PL::XB, 1, PL::XB, +1.f, tid,
                                                       tw, nl, st);
    else
      fft_io<NCC, NCE, true, false, true, WIDE>(tile, g.py, XB, 1, XB, +1.f, tid, tw, nl, st);
    __syncthreads();  // every lane has read the tile before the next coil writes it
  }
}

// ------------------------------------------------------------ K^H, row pass
struct NcRowAdjLoad {
  const float2 *tr;  // tmp rows y0.. of plane (t, c)
  int Gx;
  __device__ float2 operator()(int b, int i, int) const { return tr[(unsigned)(b * Gx + i)]; }
};
template <int SLOTS>
struct NcRowAdjStore {
  float2 acc[SLOTS];
  const float2 *bc;
  float2 *img;  // frame t
  const float *dxi, *dyi;
  int W, Gx, y0;
  bool last;
  __device__ void operator()(int b, int i, float2 v, int slot) {
    const int x = wrap_up(i + W / 2, Gx);
    if (x >= W) return;
    const int y = y0 + b;
    const unsigned o = (unsigned)(y * W + x);
    const float2 a = cadd(acc[slot], cmulc(bc[o], v));
    if (last)
      img[o] = cscale(a, dxi[x] * dyi[y]);
    else
      acc[slot] = a;
  }
};

template <bool WIDE, class PL>
__global__ __launch_bounds__(NCR) void k_nc_row_adj(NcGeom g, const float2 *__restrict__ tmp,
                                                    const float2 *__restrict__ b1,
                                                    float2 *__restrict__ img) {
  __shared__ float2 tile[NCE];
  __shared__ float2 twl[nc_table_len(WIDE)];
  const int tid = threadIdx.x;
  const int t = blockIdx.y;
  const int y0 = blockIdx.x * g.RBa;
  const int nr = min(g.RBa, g.H - y0);
  const long HW = (long)g.H * g.W;
  constexpr bool CT = PL::L != 0;
  const int Gx = CT ? PL::L : g.Gx;
  const float2 *tw = stage_twiddles(g.px, twl, tid, NCR);
  __syncthreads();
  NcRowAdjLoad ld{nullptr, Gx};
  NcRowAdjStore<row_slots_nc(WIDE)> st;
#pragma unroll
  for (int k = 0; k < row_slots_nc(WIDE); ++k) st.acc[k] = czero();
  st.img = img + t * HW;
  st.dxi = g.dxi;
  st.dyi = g.dyi;
  st.W = g.W;
  st.Gx = Gx;
  st.y0 = y0;
  for (int c = 0; c < g.C; ++c) {
    ld.tr = tmp + (((long)t * g.C + c) * g.H + y0) * Gx;
    st.bc = b1 + c * HW;
    st.last = (c == g.C - 1);
    if constexpr (CT) {
      static_assert(((NCE / PL::L) * PL::L / PL::RL + NCR - 1) / NCR * PL::RL <= row_slots_nc(WIDE),
                    "coil-sum slots of the last stage");
      fft_io_ct<NCR, PL, NCE / PL::L, false, true, true>(tile, nr, PL::L, 1, +1.f, tid, tw, ld, st);
    }
    else
      fft_io<NCR, NCE, false, true, true, WIDE>(tile, g.px, nr, Gx, 1, +1.f, tid, tw, ld, st);
    __syncthreads();
  }
}

}  // namespace

// ---------------------------------------------------------------------------
// host side

namespace {
// The compiled grid length that applies (0: none): both grid plans are
// make_fft_plan's plan of that length and the tiles have the compiled sizes.
template <class PL>
bool nc_is(const NcGeom &g) {
  auto is = [](const FftPlan &p) {
    const int r[3] = {PL::R0, PL::R1, PL::R2};
    const int n = PL::R2 > 1 ? 3 : 2;
    if (p.L != PL::L || p.nst != n) return false;
    for (int k = 0; k < n; ++k)
      if (p.radix[k] != r[k]) return false;
    return true;
  };
  return is(g.px) && is(g.py) && g.XBf == PL::XB && g.XBa == PL::XB && g.RBf <= NCE / PL::L &&
         g.RBa <= NCE / PL::L;
}
int nc_ct_len(const NcGeom &g) {
  if (g.wide) return 0;
  if (nc_is<NcPlan384>(g)) return 384;
  if (nc_is<NcPlan512>(g)) return 512;
  if (nc_is<NcPlan192>(g)) return 192;
  if (nc_is<NcPlan256>(g)) return 256;
  return 0;
}

// K's grid passes on a compiled plan: row blocks of two coils, NCE / L / 2
// grid rows each.
template <class PL>
void nc_forward_ct(const NcGeom &g, const float2 *u, const float2 *usub, const float2 *b1,
                   float2 *tmp, float2 *part, hipStream_t s) {
  constexpr int rb2 = NCE / PL::L / 2;
  const dim3 gc((g.W + g.XBf - 1) / g.XBf, g.T, g.CG);
  const dim3 gr2((g.Gy + rb2 - 1) / rb2, g.T * ((g.C + 1) / 2));
  hipLaunchKernelGGL((k_nc_col_fwd<false, PL>), gc, dim3(NCF), 0, s, g, u, usub, b1, tmp);
  AV_LAUNCH_CHECK();
  hipLaunchKernelGGL((k_nc_row_fwd<false, PL, 2>), gr2, dim3(NCR), 0, s, g, (const float2 *)tmp, part);
  AV_LAUNCH_CHECK();
}
// K^H's grid passes on a compiled plan: adjoint column blocks of three coils
// (two at grid length 512, whose tiles hold more entries: a chunk of 2048
// entries per coil holds most of a tile's; measured 867 -> 923 it/s against
// three at the 256x256x30x12 radial geometry with osf 2).
template <class PL>
void nc_adjoint_ct(const NcGeom &g, const float2 *z, const float2 *b1, float2 *tmp, float2 *img,
                   hipStream_t s) {
  const size_t cps_bytes = sizeof(int) * ((size_t)g.XBa * g.Gy + 1);
  const int nXa = (g.Gx + g.XBa - 1) / g.XBa;
  constexpr int NZ = PL::L == 512 ? 2 : 3;
  hipLaunchKernelGGL((k_nc_col_adj<false, PL, NZ>), dim3(nXa, g.T * ((g.C + NZ - 1) / NZ)), dim3(NCC),
                     cps_bytes, s, g, z, tmp);
  AV_LAUNCH_CHECK();
  hipLaunchKernelGGL((k_nc_row_adj<false, PL>), dim3((g.H + g.RBa - 1) / g.RBa, g.T), dim3(NCR), 0,
                     s, g, (const float2 *)tmp, b1, img);
  AV_LAUNCH_CHECK();
}
}  // namespace

void launch_nc_forward_grid(const NcGeom &g, const float2 *u, const float2 *usub,
                            const float2 *b1, float2 *tmp, float2 *part, hipStream_t s) {
  switch (nc_ct_len(g)) {
    case 384: nc_forward_ct<NcPlan384>(g, u, usub, b1, tmp, part, s); return;
    case 512: nc_forward_ct<NcPlan512>(g, u, usub, b1, tmp, part, s); return;
    case 192: nc_forward_ct<NcPlan192>(g, u, usub, b1, tmp, part, s); return;
    case 256: nc_forward_ct<NcPlan256>(g, u, usub, b1, tmp, part, s); return;
    default: break;
  }
  const dim3 gc((g.W + g.XBf - 1) / g.XBf, g.T, g.CG), gr((g.Gy + g.RBf - 1) / g.RBf, g.T * g.C);
  if (g.wide)
    hipLaunchKernelGGL((k_nc_col_fwd<true, NcPlanRt>), gc, dim3(NCF), 0, s, g, u, usub, b1, tmp);
  else
    hipLaunchKernelGGL((k_nc_col_fwd<false, NcPlanRt>), gc, dim3(NCF), 0, s, g, u, usub, b1, tmp);
  AV_LAUNCH_CHECK();
  if (g.wide)
    hipLaunchKernelGGL((k_nc_row_fwd<true, NcPlanRt, 1>), gr, dim3(NCR), 0, s, g, (const float2 *)tmp, part);
  else
    hipLaunchKernelGGL((k_nc_row_fwd<false, NcPlanRt, 1>), gr, dim3(NCR), 0, s, g, (const float2 *)tmp, part);
  AV_LAUNCH_CHECK();
}

int nc_sample_blocks(const NcGeom &g) {
  const long n = (long)g.T * g.C * g.M;
  return (int)std::max<long>(1, std::min<long>(1024, (n + NCR - 1) / NCR));
}

int launch_nc_samples(int mode, const NcGeom &g, const float2 *part, float2 *out, float2 *z,
                      const float2 *d, const DevScalars *sc, double *partial, hipStream_t s) {
  const int nb = nc_sample_blocks(g);
  switch (mode) {
    case kNcApplyK: hipLaunchKernelGGL(k_nc_samples<kNcApplyK>, dim3(nb), dim3(NCR), 0, s, g, part, out, z, d, sc, partial); break;
    case kNcDualUpdate: hipLaunchKernelGGL(k_nc_samples<kNcDualUpdate>, dim3(nb), dim3(NCR), 0, s, g, part, out, z, d, sc, partial); break;
    case kNcNorm2: hipLaunchKernelGGL(k_nc_samples<kNcNorm2>, dim3(nb), dim3(NCR), 0, s, g, part, out, z, d, sc, partial); break;
    case kNcResidual: hipLaunchKernelGGL(k_nc_samples<kNcResidual>, dim3(nb), dim3(NCR), 0, s, g, part, out, z, d, sc, partial); break;
    default: AV_CHECK(false, "bad sample-pass mode");
  }
  AV_LAUNCH_CHECK();
  return nb;
}

void launch_nc_adjoint(const NcGeom &g, const float2 *z, const float2 *b1, float2 *tmp,
                       float2 *img, hipStream_t s) {
  switch (nc_ct_len(g)) {
    case 384: nc_adjoint_ct<NcPlan384>(g, z, b1, tmp, img, s); return;
    case 512: nc_adjoint_ct<NcPlan512>(g, z, b1, tmp, img, s); return;
    case 192: nc_adjoint_ct<NcPlan192>(g, z, b1, tmp, img, s); return;
    case 256: nc_adjoint_ct<NcPlan256>(g, z, b1, tmp, img, s); return;
    default: break;
  }
  const size_t cps_bytes = sizeof(int) * ((size_t)g.XBa * g.Gy + 1);
  const int nXa = (g.Gx + g.XBa - 1) / g.XBa;
  const dim3 gr((g.H + g.RBa - 1) / g.RBa, g.T);
  if (g.wide)
    hipLaunchKernelGGL((k_nc_col_adj<true, NcPlanRt, 1>), dim3(nXa, g.T * g.C), dim3(NCC),
                       cps_bytes, s, g, z, tmp);
  else
    hipLaunchKernelGGL((k_nc_col_adj<false, NcPlanRt, 1>), dim3(nXa, g.T * g.C), dim3(NCC),
                       cps_bytes, s, g, z, tmp);
  AV_LAUNCH_CHECK();
  if (g.wide)
    hipLaunchKernelGGL((k_nc_row_adj<true, NcPlanRt>), gr, dim3(NCR), 0, s, g, (const float2 *)tmp, b1, img);
  else
    hipLaunchKernelGGL((k_nc_row_adj<false, NcPlanRt>), gr, dim3(NCR), 0, s, g, (const float2 *)tmp, b1, img);
  AV_LAUNCH_CHECK();
}

}  // namespace avionic
