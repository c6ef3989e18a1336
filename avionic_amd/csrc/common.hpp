// Shared device/host definitions of libavionic (gfx950).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <type_traits>

namespace avionic {

// ----------------------------------------------------------------- errors
struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

#define AV_HIP(expr)                                                          \
  do {                                                                        \
    hipError_t e_ = (expr);                                                   \
    if (e_ != hipSuccess)                                                     \
      throw ::avionic::Error(-2, std::string(#expr) + ": " +                  \
                                     hipGetErrorString(e_));                  \
  } while (0)

#define AV_CHECK(cond, msg)                                                   \
  do {                                                                        \
    if (!(cond)) throw ::avionic::Error(-1, msg);                             \
  } while (0)

#define AV_LAUNCH_CHECK() AV_HIP(hipGetLastError())

// A named range on the host timeline for profilers (rocprofv3
// --marker-trace shows it beside the kernels); ends with its scope.
struct TraceRange {
  explicit TraceRange(const char *name);
  ~TraceRange();
  TraceRange(const TraceRange &) = delete;
  TraceRange &operator=(const TraceRange &) = delete;
};

// ---------------------------------------------------------------- complex
__host__ __device__ inline float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__host__ __device__ inline float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__host__ __device__ inline float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
__host__ __device__ inline float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// conj(a) * b
__host__ __device__ inline float2 cmulc(float2 a, float2 b) {
  return make_float2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}
__host__ __device__ inline float cabs2(float2 a) { return a.x * a.x + a.y * a.y; }
__host__ __device__ inline float2 czero() { return make_float2(0.f, 0.f); }

// Two complex values of consecutive elements (x, x + 1) handled by one lane;
// the arithmetic is elementwise, in the same order as on one value.
struct C2 {
  float2 a, b;
};
__host__ __device__ inline C2 cadd(C2 p, C2 q) { return {cadd(p.a, q.a), cadd(p.b, q.b)}; }
__host__ __device__ inline C2 csub(C2 p, C2 q) { return {csub(p.a, q.a), csub(p.b, q.b)}; }
__host__ __device__ inline C2 cscale(C2 p, float s) { return {cscale(p.a, s), cscale(p.b, s)}; }

// ------------------------------------------------------ dual storage types
// bfloat16 complex pair: round-to-nearest-even from float32.
struct bf16x2 {
  uint16_t re, im;
};

__host__ __device__ inline uint32_t f2u(float f) { return __builtin_bit_cast(uint32_t, f); }
__host__ __device__ inline float u2f(uint32_t u) { return __builtin_bit_cast(float, u); }

__host__ __device__ inline uint16_t f32_to_bf16(float f) {
  uint32_t u = f2u(f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);  // NaN
  uint32_t r = u + 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(r >> 16);
}
__host__ __device__ inline float bf16_to_f32(uint16_t h) { return u2f(((uint32_t)h) << 16); }

// 16-bit fixed-point complex pair scaled by a bound r: v = q * r / 32767.
// After the norm_3 / norm_6 projections every dual component satisfies
// |re|, |im| <= r (the projection radius), so the dual variables fit this
// format with an absolute resolution of r / 32767.
struct q16x2 {
  int16_t re, im;
};
constexpr float kQ16Max = 32767.f;

__host__ __device__ inline int16_t f32_to_q16(float v, float inv_scale) {
  float q = rintf(v * inv_scale);
  q = fminf(fmaxf(q, -kQ16Max), kQ16Max);
  return (int16_t)q;
}

// Global-memory access at a 32-bit element offset from a base pointer.  The
// kernels read their pointers from a task table in memory, so the compiler
// cannot tell their address space and would emit flat loads with 64-bit
// per-lane addresses; through these helpers every access is a global load or
// store with the (uniform) base in scalar registers and a 32-bit lane offset.
// i * sizeof(T) must stay below 4 GiB (callers bound their extents).
// (The host pass only parses these bodies; it cannot form address-space
// qualified class objects, so it sees a plain access.)
template <class T>
__device__ __forceinline__ T gld(const T *base, unsigned i) {
#if defined(__HIP_DEVICE_COMPILE__)
  typedef __attribute__((address_space(1))) const char gchar;
  typedef __attribute__((address_space(1))) const T gT;
  return *(gT *)((gchar *)base + i * (unsigned)sizeof(T));
#else
  return base[i];
#endif
}
template <class T>
__device__ __forceinline__ void gst(T *base, unsigned i, const T &v) {
#if defined(__HIP_DEVICE_COMPILE__)
  typedef __attribute__((address_space(1))) char gchar;
  typedef __attribute__((address_space(1))) T gT;
  *(gT *)((gchar *)base + i * (unsigned)sizeof(T)) = v;
#else
  base[i] = v;
#endif
}
// Integer (or integer vector) type of a 4-, 8- or 16-byte value, the operand
// type of the non-temporal builtins.
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
template <class T>
using StreamWord = typename std::conditional<
    sizeof(T) == 16, u32x4,
    typename std::conditional<sizeof(T) == 8, unsigned long long, unsigned>::type>::type;

// Streaming (non-temporal) load of a 4-, 8- or 16-byte value that no other
// lane reads.
template <class T>
__device__ __forceinline__ T gld_stream(const T *base, unsigned i) {
  static_assert(sizeof(T) == 4 || sizeof(T) == 8 || sizeof(T) == 16, "4-, 8- or 16-byte values");
#if defined(__HIP_DEVICE_COMPILE__)
  typedef StreamWord<T> U;
  typedef __attribute__((address_space(1))) const char gchar;
  typedef __attribute__((address_space(1))) const U gU;
  return __builtin_bit_cast(T, __builtin_nontemporal_load((gU *)((gchar *)base + i * (unsigned)sizeof(T))));
#else
  return base[i];
#endif
}

// Streaming (non-temporal) store of a 4-, 8- or 16-byte value, for outputs
// that are not read again before they would have left the caches.
template <class T>
__device__ __forceinline__ void gst_stream(T *base, unsigned i, const T &v) {
  static_assert(sizeof(T) == 4 || sizeof(T) == 8 || sizeof(T) == 16, "4-, 8- or 16-byte values");
#if defined(__HIP_DEVICE_COMPILE__)
  typedef StreamWord<T> U;
  typedef __attribute__((address_space(1))) char gchar;
  typedef __attribute__((address_space(1))) U gU;
  __builtin_nontemporal_store(__builtin_bit_cast(U, v), (gU *)((gchar *)base + i * (unsigned)sizeof(T)));
#else
  base[i] = v;
#endif
}

// Both complex values of a pair of elements (i, i + 1), i even, in one
// 16-byte access (the base must be 16-byte aligned).
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned u32x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ C2 c2_of(f32x4 v) {
  return {make_float2(v.x, v.y), make_float2(v.z, v.w)};
}
__device__ __forceinline__ f32x4 f4_of(C2 v) { return f32x4{v.a.x, v.a.y, v.b.x, v.b.y}; }
__device__ __forceinline__ C2 gld2(const float2 *base, unsigned i) {
  return c2_of(gld(reinterpret_cast<const f32x4 *>(base), i >> 1));
}
__device__ __forceinline__ C2 gld2_stream(const float2 *base, unsigned i) {
  return c2_of(gld_stream(reinterpret_cast<const f32x4 *>(base), i >> 1));
}
__device__ __forceinline__ void gst2_stream(float2 *base, unsigned i, C2 v) {
  gst_stream(reinterpret_cast<f32x4 *>(base), i >> 1, f4_of(v));
}

// Dual storage kinds (avionic_params::dualStorage).
enum DualStorage { kDualF32 = 0, kDualBF16 = 1, kDualQ16 = 2 };
inline size_t dual_bytes(int storage) { return storage == kDualF32 ? 8 : 4; }

// Loads and stores of one complex dual value; r is the projection radius of
// its component group (used by the fixed-point format only).
template <typename DT>
struct DualIO;
template <>
struct DualIO<float2> {
  __device__ static inline float2 load(const float2 *p, unsigned i, float) { return gld(p, i); }
  __device__ static inline float2 load_stream(const float2 *p, unsigned i, float) {
    return gld_stream(p, i);
  }
  __device__ static inline void store(float2 *p, unsigned i, float2 v, float) { gst_stream(p, i, v); }
  // pairs (i, i + 1), i even
  __device__ static inline C2 load2(const float2 *p, unsigned i, float) { return gld2(p, i); }
  __device__ static inline C2 load2_stream(const float2 *p, unsigned i, float) {
    return gld2_stream(p, i);
  }
  __device__ static inline void store2(float2 *p, unsigned i, C2 v, float) { gst2_stream(p, i, v); }
};
template <>
struct DualIO<bf16x2> {
  __device__ static inline float2 load(const bf16x2 *p, unsigned i, float) {
    return decode(gld(reinterpret_cast<const uint32_t *>(p), i));
  }
  __device__ static inline float2 load_stream(const bf16x2 *p, unsigned i, float) {
    return decode(gld_stream(reinterpret_cast<const uint32_t *>(p), i));
  }
  __device__ static inline float2 decode(uint32_t w) {
    return make_float2(u2f(w << 16), u2f(w & 0xffff0000u));
  }
  __device__ static inline uint32_t encode(float2 v) {
    return (uint32_t)f32_to_bf16(v.x) | ((uint32_t)f32_to_bf16(v.y) << 16);
  }
  __device__ static inline void store(bf16x2 *p, unsigned i, float2 v, float) {
    gst_stream(reinterpret_cast<uint32_t *>(p), i, encode(v));
  }
  // pairs (i, i + 1), i even: one 8-byte access
  __device__ static inline C2 load2(const bf16x2 *p, unsigned i, float) {
    const u32x2 w = gld(reinterpret_cast<const u32x2 *>(p), i >> 1);
    return {decode(w.x), decode(w.y)};
  }
  __device__ static inline C2 load2_stream(const bf16x2 *p, unsigned i, float) {
    const u32x2 w = gld_stream(reinterpret_cast<const u32x2 *>(p), i >> 1);
    return {decode(w.x), decode(w.y)};
  }
  __device__ static inline void store2(bf16x2 *p, unsigned i, C2 v, float) {
    gst_stream(reinterpret_cast<u32x2 *>(p), i >> 1, u32x2{encode(v.a), encode(v.b)});
  }
};
template <>
struct DualIO<q16x2> {
  __device__ static inline float2 load(const q16x2 *p, unsigned i, float r) {
    return decode(gld(reinterpret_cast<const uint32_t *>(p), i), r);
  }
  __device__ static inline float2 load_stream(const q16x2 *p, unsigned i, float r) {
    return decode(gld_stream(reinterpret_cast<const uint32_t *>(p), i), r);
  }
  __device__ static inline float2 decode(uint32_t w, float r) {
    const float s = r * (1.f / kQ16Max);
    return make_float2((float)(int16_t)(w & 0xffffu) * s, (float)(int16_t)(w >> 16) * s);
  }
  // The encode scale 32767 / r through the hardware reciprocal (1 ulp): the
  // stored value may differ from the correctly rounded quotient's by one
  // step of the int16 grid, far below the format's own resolution.
  __device__ static inline uint32_t encode(float2 v, float inv) {
    return (uint32_t)(uint16_t)f32_to_q16(v.x, inv) |
           ((uint32_t)(uint16_t)f32_to_q16(v.y, inv) << 16);
  }
  __device__ static inline void store(q16x2 *p, unsigned i, float2 v, float r) {
    const float inv = kQ16Max * __builtin_amdgcn_rcpf(r);
    gst_stream(reinterpret_cast<uint32_t *>(p), i, encode(v, inv));
  }
  // pairs (i, i + 1), i even: one 8-byte access
  __device__ static inline C2 load2(const q16x2 *p, unsigned i, float r) {
    const u32x2 w = gld(reinterpret_cast<const u32x2 *>(p), i >> 1);
    return {decode(w.x, r), decode(w.y, r)};
  }
  __device__ static inline C2 load2_stream(const q16x2 *p, unsigned i, float r) {
    const u32x2 w = gld_stream(reinterpret_cast<const u32x2 *>(p), i >> 1);
    return {decode(w.x, r), decode(w.y, r)};
  }
  __device__ static inline void store2(q16x2 *p, unsigned i, C2 v, float r) {
    const float inv = kQ16Max * __builtin_amdgcn_rcpf(r);
    gst_stream(reinterpret_cast<u32x2 *>(p), i >> 1, u32x2{encode(v.a, inv), encode(v.b, inv)});
  }
};

// ------------------------------------------------------------ FFT tiling
// Elements (complex) held in LDS by one FFT tile, and threads per tile.
constexpr int kFftThreads = 256;
constexpr int kFftTileElems = 4096;  // 32 KiB of float2 (column passes)
constexpr int kRowTileElems = 2048;  // 16 KiB of float2 (row passes)
constexpr int kRowMaxRows = 256;     // rows of one row-pass tile at most
constexpr int kMaxStages = 12;
constexpr int kMaxFftLen = 1024;  // longest 1D transform (twiddles staged in LDS)
// Longest transform of a "narrow" plan.  Narrow plans (length <= 512, radices
// 16, 8, 4, 2, 3, 5, 7) run in kernels whose LDS twiddle tables and register
// arrays are sized for them; longer plans and plans with a radix-11 or -13
// stage ("wide") run in separately compiled kernel variants, so that those
// stages' larger butterflies and tables cost nothing where they are not used.
constexpr int kMaxFftLenNarrow = 512;
// LDS tile of the hot loop's fused column passes (complex elements).
#ifndef AVIONIC_COL_EMAX
#define AVIONIC_COL_EMAX 4096
#endif
constexpr int kColTileElems = AVIONIC_COL_EMAX;

// A 1D FFT plan for one length: Stockham stages (radix, Ns) and twiddles.
struct FftPlan {
  int L;
  int nst;
  int radix[kMaxStages];
  int ns[kMaxStages];
  const float2 *tw;  // tw[m] = exp(-2 pi i m / L), m < L (device)
};
__host__ __device__ inline bool plan_is_wide(const FftPlan &p) {
  if (p.L > kMaxFftLenNarrow) return true;
  for (int s = 0; s < p.nst; ++s)
    if (p.radix[s] == 11 || p.radix[s] == 13) return true;
  return false;
}

// ----------------------------------------------------- reductions / state
constexpr int kRedThreads = 256;

// Per-reconstruction scalars kept on the device so that the step-size rule
// runs without a host round trip (and inside captured graphs).
struct DevScalars {
  double sigma, tau, lambda, ratio;
  double nKx2, nx2;  // last step-rule reductions
  int iter;          // iteration counter (incremented once per iteration)
  int diverged;      // set when a reduction produced NaN/Inf
  int maxit;         // iterations to run (task-list execution idles past it)
  int schedule;      // 0: adapt when iter < 10 || iter % 50 == 0; 1: always
};

}  // namespace avionic
