// In-LDS mixed-radix Stockham FFT for one workgroup (gfx950, 64-wide waves).
//
// A tile of `B` transforms of length L lives in LDS; element (b, i) sits at
// b*bstride + i*istride.  Each stage of radix R reads R inputs per butterfly
// into registers, twiddles them, runs an in-register DFT_R, synchronises, and
// writes the outputs to their Stockham positions in place; the next stage
// starts after a second barrier.  Lanes are mapped so that consecutive lanes
// touch consecutive LDS words (unit stride on whichever of b / i has it).
#pragma once

#include "common.hpp"

namespace avionic {

template <int R>
struct RootTable;  // cos / sin (2 pi m / R), m < R
template <> struct RootTable<3> {
  static constexpr float c[3] = {1.000000000e+00f, -5.000000000e-01f, -5.000000000e-01f};
  static constexpr float s[3] = {0.000000000e+00f, 8.660254038e-01f, -8.660254038e-01f};
};
template <> struct RootTable<5> {
  static constexpr float c[5] = {1.000000000e+00f, 3.090169944e-01f, -8.090169944e-01f, -8.090169944e-01f, 3.090169944e-01f};
  static constexpr float s[5] = {0.000000000e+00f, 9.510565163e-01f, 5.877852523e-01f, -5.877852523e-01f, -9.510565163e-01f};
};
template <> struct RootTable<7> {
  static constexpr float c[7] = {1.000000000e+00f, 6.234898019e-01f, -2.225209340e-01f, -9.009688679e-01f, -9.009688679e-01f, -2.225209340e-01f, 6.234898019e-01f};
  static constexpr float s[7] = {0.000000000e+00f, 7.818314825e-01f, 9.749279122e-01f, 4.338837391e-01f, -4.338837391e-01f, -9.749279122e-01f, -7.818314825e-01f};
};
template <> struct RootTable<11> {
  static constexpr float c[11] = {1.000000000e+00f, 8.412535328e-01f, 4.154150130e-01f, -1.423148383e-01f, -6.548607339e-01f, -9.594929736e-01f, -9.594929736e-01f, -6.548607339e-01f, -1.423148383e-01f, 4.154150130e-01f, 8.412535328e-01f};
  static constexpr float s[11] = {0.000000000e+00f, 5.406408175e-01f, 9.096319954e-01f, 9.898214419e-01f, 7.557495744e-01f, 2.817325568e-01f, -2.817325568e-01f, -7.557495744e-01f, -9.898214419e-01f, -9.096319954e-01f, -5.406408175e-01f};
};
template <> struct RootTable<13> {
  static constexpr float c[13] = {1.000000000e+00f, 8.854560257e-01f, 5.680647467e-01f, 1.205366803e-01f, -3.546048870e-01f, -7.485107482e-01f, -9.709418174e-01f, -9.709418174e-01f, -7.485107482e-01f, -3.546048870e-01f, 1.205366803e-01f, 5.680647467e-01f, 8.854560257e-01f};
  static constexpr float s[13] = {0.000000000e+00f, 4.647231720e-01f, 8.229838659e-01f, 9.927088741e-01f, 9.350162427e-01f, 6.631226582e-01f, 2.393156643e-01f, -2.393156643e-01f, -6.631226582e-01f, -9.350162427e-01f, -9.927088741e-01f, -8.229838659e-01f, -4.647231720e-01f};
};
// exp(sign * i * theta) applied as (c, sign * s).
__device__ inline float2 crot(float2 a, float c, float s, float sign) {
  float ss = sign * s;
  return make_float2(a.x * c - a.y * ss, a.x * ss + a.y * c);
}

// 4-point DFT, X[k] = sum_n x[n] exp(sign 2 pi i n k / 4).
__device__ inline void dft4(float2 &x0, float2 &x1, float2 &x2, float2 &x3, float sign) {
  float2 a = cadd(x0, x2), b = csub(x0, x2), c = cadd(x1, x3), d = csub(x1, x3);
  float2 di = make_float2(-sign * d.y, sign * d.x);  // d * (sign i)
  x0 = cadd(a, c);
  x2 = csub(a, c);
  x1 = cadd(b, di);
  x3 = csub(b, di);
}

template <int R>
struct Dft {
  // Odd radix via conjugate pairs: X_r = v0 + sum_{q<=(R-1)/2}
  //   (v_q + v_{R-q}) cos(2 pi q r / R) + sign i (v_q - v_{R-q}) sin(2 pi q r / R).
  __device__ static inline void run(float2 *v, float sign) {
    constexpr int Hh = (R - 1) / 2;
    float2 sp[Hh], sm[Hh];
#pragma unroll
    for (int q = 1; q <= Hh; ++q) {
      sp[q - 1] = cadd(v[q], v[R - q]);
      sm[q - 1] = csub(v[q], v[R - q]);
    }
    float2 o0 = v[0];
#pragma unroll
    for (int q = 0; q < Hh; ++q) o0 = cadd(o0, sp[q]);
#pragma unroll
    for (int r = 1; r <= Hh; ++r) {
      float ar = v[0].x, ai = v[0].y, br = 0.f, bi = 0.f;
#pragma unroll
      for (int q = 1; q <= Hh; ++q) {
        const int m = (q * r) % R;
        const float c = RootTable<R>::c[m], sn = RootTable<R>::s[m];
        ar += sp[q - 1].x * c;
        ai += sp[q - 1].y * c;
        br += sm[q - 1].x * sn;
        bi += sm[q - 1].y * sn;
      }
      // sign i (br + i bi) = (-sign bi, sign br)
      v[r] = make_float2(ar - sign * bi, ai + sign * br);
      v[R - r] = make_float2(ar + sign * bi, ai - sign * br);
    }
    v[0] = o0;
  }
};

template <>
struct Dft<2> {
  __device__ static inline void run(float2 *v, float) {
    float2 a = v[0];
    v[0] = cadd(a, v[1]);
    v[1] = csub(a, v[1]);
  }
};

template <>
struct Dft<4> {
  __device__ static inline void run(float2 *v, float sign) { dft4(v[0], v[1], v[2], v[3], sign); }
};

template <>
struct Dft<8> {
  // radix-2 split of two 4-point DFTs.
  __device__ static inline void run(float2 *v, float sign) {
    float2 e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6];
    float2 o0 = v[1], o1 = v[3], o2 = v[5], o3 = v[7];
    dft4(e0, e1, e2, e3, sign);
    dft4(o0, o1, o2, o3, sign);
    const float h = 0.70710678118654752f;
    o1 = crot(o1, h, h, sign);
    o2 = make_float2(-sign * o2.y, sign * o2.x);
    o3 = crot(o3, -h, h, sign);
    v[0] = cadd(e0, o0); v[4] = csub(e0, o0);
    v[1] = cadd(e1, o1); v[5] = csub(e1, o1);
    v[2] = cadd(e2, o2); v[6] = csub(e2, o2);
    v[3] = cadd(e3, o3); v[7] = csub(e3, o3);
  }
};

template <>
struct Dft<16> {
  // 4 x 4: n = 4 n1 + n2, k = k1 + 4 k2.
  __device__ static inline void run(float2 *v, float sign) {
    float2 a[4][4];
#pragma unroll
    for (int n2 = 0; n2 < 4; ++n2) {
      a[n2][0] = v[n2]; a[n2][1] = v[4 + n2]; a[n2][2] = v[8 + n2]; a[n2][3] = v[12 + n2];
      dft4(a[n2][0], a[n2][1], a[n2][2], a[n2][3], sign);
    }
    // twiddles w16^(n2 k1)
    const float c1 = 0.92387953251128674f, s1 = 0.38268343236508978f;
    const float h = 0.70710678118654752f;
    a[1][1] = crot(a[1][1], c1, s1, sign);
    a[1][2] = crot(a[1][2], h, h, sign);
    a[1][3] = crot(a[1][3], s1, c1, sign);
    a[2][1] = crot(a[2][1], h, h, sign);
    a[2][2] = make_float2(-sign * a[2][2].y, sign * a[2][2].x);
    a[2][3] = crot(a[2][3], -h, h, sign);
    a[3][1] = crot(a[3][1], s1, c1, sign);
    a[3][2] = crot(a[3][2], -h, h, sign);
    a[3][3] = crot(a[3][3], -c1, -s1, sign);
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) {
      float2 b0 = a[0][k1], b1 = a[1][k1], b2 = a[2][k1], b3 = a[3][k1];
      dft4(b0, b1, b2, b3, sign);
      v[k1] = b0; v[k1 + 4] = b1; v[k1 + 8] = b2; v[k1 + 12] = b3;
    }
  }
};

// One Stockham stage.  Must be called by all NT threads of the block.
// BFAST: consecutive lanes take consecutive transforms (b) instead of
// consecutive butterflies (j); pick the one whose LDS stride is 1.
template <int R, int NT, int EMAX, bool BFAST>
__device__ inline void fft_stage(float2 *buf, int L, int B, int bstride, int istride, int Ns,
                                 float sign, const float2 *__restrict__ tw, int tid) {
  constexpr int MAXB = (EMAX / R + NT - 1) / NT;
  // Opaque copy of the lane id: keeps the compiler from hoisting each radix's
  // per-lane index arithmetic out of the stage loop (which would keep every
  // radix's indices live at once and multiply the register footprint).
  asm volatile("" : "+v"(tid));
  const int nbf = L / R;
  const int total = nbf * B;
  const int twstep = L / (Ns * R);
  const int rstride = nbf * istride;
  float2 v[MAXB][R];
#pragma unroll
  for (int c = 0; c < MAXB; ++c) {
    const int idx = tid + c * NT;
    if (idx < total) {
      int b, j;
      if (BFAST) { b = idx % B; j = idx / B; } else { j = idx % nbf; b = idx / nbf; }
      const int k = j % Ns;
      const int base = b * bstride + j * istride;
      v[c][0] = buf[base];
#pragma unroll
      for (int r = 1; r < R; ++r) {
        float2 w = tw[r * k * twstep];
        w.y = -sign * w.y;  // tw holds exp(-i theta); want exp(sign i theta)
        v[c][r] = cmul(buf[base + r * rstride], w);
      }
      Dft<R>::run(v[c], sign);
    }
  }
  __syncthreads();
#pragma unroll
  for (int c = 0; c < MAXB; ++c) {
    const int idx = tid + c * NT;
    if (idx < total) {
      int b, j;
      if (BFAST) { b = idx % B; j = idx / B; } else { j = idx % nbf; b = idx / nbf; }
      const int k = j % Ns;
      const int base = b * bstride + ((j / Ns) * Ns * R + k) * istride;
      const int ostride = Ns * istride;
#pragma unroll
      for (int r = 0; r < R; ++r) buf[base + r * ostride] = v[c][r];
    }
  }
  __syncthreads();
}

// Full transform of B batched length-L sequences in LDS (unnormalised).
// Radices 11 and 13 are compiled in only with WIDE (see plan_is_wide).
template <int NT, int EMAX, bool BFAST, bool WIDE = true>
__device__ inline void fft_lds(float2 *buf, const FftPlan &p, int B, int bstride, int istride,
                               float sign, int tid) {
  for (int s = 0; s < p.nst; ++s) {
    const int R = p.radix[s], Ns = p.ns[s];
    switch (R) {
      case 16: fft_stage<16, NT, EMAX, BFAST>(buf, p.L, B, bstride, istride, Ns, sign, p.tw, tid); break;
      case 8: fft_stage<8, NT, EMAX, BFAST>(buf, p.L, B, bstride, istride, Ns, sign, p.tw, tid); break;
      case 4: fft_stage<4, NT, EMAX, BFAST>(buf, p.L, B, bstride, istride, Ns, sign, p.tw, tid); break;
      case 2: fft_stage<2, NT, EMAX, BFAST>(buf, p.L, B, bstride, istride, Ns, sign, p.tw, tid); break;
      case 3: fft_stage<3, NT, EMAX, BFAST>(buf, p.L, B, bstride, istride, Ns, sign, p.tw, tid); break;
      case 5: fft_stage<5, NT, EMAX, BFAST>(buf, p.L, B, bstride, istride, Ns, sign, p.tw, tid); break;
      case 7: fft_stage<7, NT, EMAX, BFAST>(buf, p.L, B, bstride, istride, Ns, sign, p.tw, tid); break;
      case 11:
        if constexpr (WIDE) fft_stage<11, NT, EMAX, BFAST>(buf, p.L, B, bstride, istride, Ns, sign, p.tw, tid);
        break;
      case 13:
        if constexpr (WIDE) fft_stage<13, NT, EMAX, BFAST>(buf, p.L, B, bstride, istride, Ns, sign, p.tw, tid);
        break;
      default: break;
    }
  }
}

// ---------------------------------------------------------------------------
// Fused transforms: the first stage may take its inputs from a load functor
// (global memory) instead of LDS, and the last stage may hand its outputs to
// a store functor instead of writing LDS.  Only the exchanges between stages
// touch LDS.  `tw` may point to LDS (a per-block copy of the twiddles).
//
//   ld(b, i, slot) -> float2  input element i of transform b   (natural order)
//   st(b, i, v, slot)        output element i of transform b   (natural order)
// The lane -> (b, j) mapping of a stage is the same for every call with the
// same plan and B, so a store functor may keep per-lane state across calls:
// `slot` (a compile-time constant once the stage is unrolled) numbers the
// outputs a lane produces, at most kMaxSlots of them.

template <int R, int NT, int EMAX, bool BFAST, bool FROMG, bool TOG, class LD, class ST>
__device__ __forceinline__ void fft_stage_io(float2 *buf, int L, int B, int bstride, int istride, int Ns,
                                    float sign, const float2 *__restrict__ tw, int tid, LD &ld,
                                    ST &st) {
  constexpr int MAXB = (EMAX / R + NT - 1) / NT;
  asm volatile("" : "+v"(tid));
  const int nbf = L / R;
  const int total = nbf * B;
  const int twstep = L / (Ns * R);
  const int rstride = nbf * istride;
  float2 v[MAXB][R];
#pragma unroll
  for (int c = 0; c < MAXB; ++c) {
    const int idx = tid + c * NT;
    if (idx < total) {
      int b, j;
      if (BFAST) { b = idx % B; j = idx / B; } else { j = idx % nbf; b = idx / nbf; }
      const int k = j % Ns;
      if (FROMG) {
#pragma unroll
        for (int r = 0; r < R; ++r) v[c][r] = ld(b, j + r * nbf, c * R + r);
      } else {
        const int base = b * bstride + j * istride;
#pragma unroll
        for (int r = 0; r < R; ++r) v[c][r] = buf[base + r * rstride];
      }
      if (Ns > 1) {
#pragma unroll
        for (int r = 1; r < R; ++r) {
          float2 w = tw[r * k * twstep];
          w.y = -sign * w.y;
          v[c][r] = cmul(v[c][r], w);
        }
      }
      Dft<R>::run(v[c], sign);
    }
  }
  if (!FROMG && !TOG) __syncthreads();  // in place: all reads before any write
#pragma unroll
  for (int c = 0; c < MAXB; ++c) {
    const int idx = tid + c * NT;
    if (idx < total) {
      int b, j;
      if (BFAST) { b = idx % B; j = idx / B; } else { j = idx % nbf; b = idx / nbf; }
      const int k = j % Ns;
      const int o = (j / Ns) * Ns * R + k;
      if (TOG) {
#pragma unroll
        for (int r = 0; r < R; ++r) st(b, o + r * Ns, v[c][r], c * R + r);
      } else {
        const int base = b * bstride + o * istride;
        const int ostride = Ns * istride;
#pragma unroll
        for (int r = 0; r < R; ++r) buf[base + r * ostride] = v[c][r];
      }
    }
  }
  if (!TOG) __syncthreads();
}

template <int NT, int EMAX, bool BFAST, bool FROMG, bool TOG, bool WIDE, class LD, class ST>
__device__ inline void fft_stage_io_dispatch(int R, float2 *buf, int L, int B, int bstride,
                                             int istride, int Ns, float sign, const float2 *tw,
                                             int tid, LD &ld, ST &st) {
  switch (R) {
    case 16: fft_stage_io<16, NT, EMAX, BFAST, FROMG, TOG>(buf, L, B, bstride, istride, Ns, sign, tw, tid, ld, st); break;
    case 8: fft_stage_io<8, NT, EMAX, BFAST, FROMG, TOG>(buf, L, B, bstride, istride, Ns, sign, tw, tid, ld, st); break;
    case 4: fft_stage_io<4, NT, EMAX, BFAST, FROMG, TOG>(buf, L, B, bstride, istride, Ns, sign, tw, tid, ld, st); break;
    case 2: fft_stage_io<2, NT, EMAX, BFAST, FROMG, TOG>(buf, L, B, bstride, istride, Ns, sign, tw, tid, ld, st); break;
    case 3: fft_stage_io<3, NT, EMAX, BFAST, FROMG, TOG>(buf, L, B, bstride, istride, Ns, sign, tw, tid, ld, st); break;
    case 5: fft_stage_io<5, NT, EMAX, BFAST, FROMG, TOG>(buf, L, B, bstride, istride, Ns, sign, tw, tid, ld, st); break;
    case 7: fft_stage_io<7, NT, EMAX, BFAST, FROMG, TOG>(buf, L, B, bstride, istride, Ns, sign, tw, tid, ld, st); break;
    case 11:
      if constexpr (WIDE) fft_stage_io<11, NT, EMAX, BFAST, FROMG, TOG>(buf, L, B, bstride, istride, Ns, sign, tw, tid, ld, st);
      break;
    case 13:
      if constexpr (WIDE) fft_stage_io<13, NT, EMAX, BFAST, FROMG, TOG>(buf, L, B, bstride, istride, Ns, sign, tw, tid, ld, st);
      break;
    default: break;
  }
}

struct NoLoad {
  __device__ float2 operator()(int, int, int) const { return make_float2(0.f, 0.f); }
};
struct NoStore {
  __device__ void operator()(int, int, float2, int) {}
};

// Outputs per lane of one stage: max over radices of ceil(EMAX/R/NT) * R
// (radices 11 and 13 only in wide kernels).
constexpr int max_slots(int emax, int nt, bool wide) {
  int m = 0;
  const int rs[] = {16, 8, 4, 2, 3, 5, 7, 11, 13};
  for (int r : rs) {
    if (!wide && r > 8 && r != 16) continue;
    const int v = ((emax / r + nt - 1) / nt) * r;
    m = v > m ? v : m;
  }
  return m;
}

// Transform with optional global input / output.  With FROMG the LDS tile's
// previous contents may still be being read by other lanes: callers sync
// before re-using a tile.  With TOG the LDS tile is free once this returns
// (the last stage read it before the function's final store loop; lanes may
// still be reading it, so callers sync before writing the tile again).
template <int NT, int EMAX, bool BFAST, bool FROMG, bool TOG, bool WIDE, class LD, class ST>
__device__ inline void fft_io(float2 *buf, const FftPlan &p, int B, int bstride, int istride,
                              float sign, int tid, const float2 *tw, LD &ld, ST &st) {
  const int n = p.nst;
  if (n == 0) {  // length 1: identity
    for (int b = tid; b < B; b += NT) {
      const float2 v = FROMG ? ld(b, 0, 0) : buf[b * bstride];
      if (TOG) st(b, 0, v, 0); else buf[b * bstride] = v;
    }
    if (!TOG) __syncthreads();
    return;
  }
  if (n == 1) {
    fft_stage_io_dispatch<NT, EMAX, BFAST, FROMG, TOG, WIDE>(p.radix[0], buf, p.L, B, bstride, istride,
                                                       p.ns[0], sign, tw, tid, ld, st);
    return;
  }
  NoStore ns;
  NoLoad nl;
  fft_stage_io_dispatch<NT, EMAX, BFAST, FROMG, false, WIDE>(p.radix[0], buf, p.L, B, bstride, istride,
                                                       p.ns[0], sign, tw, tid, ld, ns);
  for (int s = 1; s < n - 1; ++s)
    fft_stage_io_dispatch<NT, EMAX, BFAST, false, false, WIDE>(p.radix[s], buf, p.L, B, bstride, istride,
                                                         p.ns[s], sign, tw, tid, nl, ns);
  fft_stage_io_dispatch<NT, EMAX, BFAST, false, TOG, WIDE>(p.radix[n - 1], buf, p.L, B, bstride, istride,
                                                     p.ns[n - 1], sign, tw, tid, nl, st);
}

// A plan known at compile time: length L = R0 R1 R2, stages in that order
// (R2 = 1: two stages).  fft_io_ct runs the stages fft_io runs on
// make_fft_plan's plan of the same radices (same butterflies, twiddle indices
// and Stockham positions), with the length, the stage strides and, for
// BFAST, the transform count folded into constants: the per-element index
// divisions and remainders of the runtime plan become multiplies and shifts.
// `tw` holds the length-L twiddle table (stage_twiddles of that plan).
template <int L_, int R0_, int R1_, int R2_ = 1>
struct FftCt {
  static constexpr int L = L_, R0 = R0_, R1 = R1_, R2 = R2_;
  static_assert(R0_ * R1_ * R2_ == L_, "radices must multiply to the length");
};
// BMAX: most transforms per call (sizes the per-lane register arrays).  For
// BFAST the transform count is the constant BMAX and `B` is ignored.
template <int NT, class P, int BMAX, bool BFAST, bool FROMG, bool TOG, class LD, class ST>
__device__ __forceinline__ void fft_io_ct(float2 *buf, int B, int bstride, int istride, float sign,
                                          int tid, const float2 *tw, LD &ld, ST &st) {
  constexpr int L = P::L, E = L * BMAX;
  const int nb = BFAST ? BMAX : B;
  NoStore ns;
  NoLoad nl;
  if constexpr (P::R2 == 1) {
    fft_stage_io<P::R0, NT, E, BFAST, FROMG, false>(buf, L, nb, bstride, istride, 1, sign, tw, tid, ld, ns);
    fft_stage_io<P::R1, NT, E, BFAST, false, TOG>(buf, L, nb, bstride, istride, P::R0, sign, tw, tid, nl, st);
  } else {
    fft_stage_io<P::R0, NT, E, BFAST, FROMG, false>(buf, L, nb, bstride, istride, 1, sign, tw, tid, ld, ns);
    fft_stage_io<P::R1, NT, E, BFAST, false, false>(buf, L, nb, bstride, istride, P::R0, sign, tw, tid, nl, ns);
    fft_stage_io<P::R2, NT, E, BFAST, false, TOG>(buf, L, nb, bstride, istride, P::R0 * P::R1, sign, tw, tid, nl, st);
  }
}

// Copy a plan's twiddles to LDS (L <= kMaxFftLen) and return the LDS view.
__device__ inline const float2 *stage_twiddles(const FftPlan &p, float2 *lds, int tid, int nt) {
  for (int i = tid; i < p.L; i += nt) lds[i] = p.tw[i];
  return lds;
}

}  // namespace avionic
