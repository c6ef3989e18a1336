// Volumetric Cartesian operator pair for TGV2_3D
// (CartesianOperator3D, cartesian_operator3d.cpp:64-345):
//   K  (image -> k-space): z_c = M . F^{+}(b1_c . u) / sqrt(N)
//   K^H (k-space -> image): u = sum_c conj(b1_c) . F^{-}(M . z_c) / sqrt(N)
// with 3D transforms over [depth][height][width] (F^{+}: e^{+i}, F^{-}:
// e^{-i}) and k-space stored coil-major [c][depth][height][width], as the
// reference stores it.  The reference's `centered` flag changes nothing in
// this operator (both branches run the same transform), so there is none here.
// A 3D transform is the batched 2D transform of every slice followed by a
// transform along depth; the coil expansion, mask and coil sum are small
// elementwise kernels.
#include <cmath>
#include <vector>

#include "cartesian.hpp"
#include "reg_device.hpp"
#include "solver.hpp"

namespace avionic {

namespace {

constexpr int kT3 = 256;

__global__ __launch_bounds__(kT3) void k3_expand(const float2 *__restrict__ u,
                                                 const float2 *__restrict__ usub,
                                                 const float2 *__restrict__ b1,
                                                 float2 *__restrict__ out, long N, long n) {
  for (long e = (long)blockIdx.x * kT3 + threadIdx.x; e < n; e += (long)gridDim.x * kT3) {
    const long i = e % N;
    float2 v = u[i];
    if (usub) v = csub(v, usub[i]);
    out[e] = cmul(b1[e], v);
  }
}

// out = in . mask (mask over one coil's N elements; null = ones)
__global__ __launch_bounds__(kT3) void k3_mask(const float2 *__restrict__ in,
                                               const float *__restrict__ mask,
                                               float2 *__restrict__ out, long N, long n) {
  for (long e = (long)blockIdx.x * kT3 + threadIdx.x; e < n; e += (long)gridDim.x * kT3)
    out[e] = mask ? cscale(in[e], mask[e % N]) : in[e];
}

__global__ __launch_bounds__(kT3) void k3_coil_sum(const float2 *__restrict__ t,
                                                   const float2 *__restrict__ b1,
                                                   float2 *__restrict__ img, long N, int C) {
  for (long i = (long)blockIdx.x * kT3 + threadIdx.x; i < N; i += (long)gridDim.x * kT3) {
    float2 acc = czero();
    for (int c = 0; c < C; ++c) acc = cadd(acc, cmulc(b1[c * N + i], t[c * N + i]));
    img[i] = acc;
  }
}

// z = (z + sigma K u - sigma d) / (1 + sigma / lambda)  (tgv2_3d.cpp loop)
__global__ __launch_bounds__(kT3) void k3_dual_z(float2 *__restrict__ z,
                                                 const float2 *__restrict__ ku,
                                                 const float2 *__restrict__ d,
                                                 const DevScalars *__restrict__ sc, long n) {
  const float sigma = (float)sc->sigma;
  const float inv = (float)(1.0 / (1.0 + sc->sigma / sc->lambda));
  for (long e = (long)blockIdx.x * kT3 + threadIdx.x; e < n; e += (long)gridDim.x * kT3) {
    float2 v = cadd(z[e], cscale(ku[e], sigma));
    z[e] = cscale(csub(v, cscale(d[e], sigma)), inv);
  }
}

// partial[b] = sum over the block's elements of |a - b|^2 (b may be null)
__global__ __launch_bounds__(kT3) void k3_abs2(const float2 *__restrict__ a,
                                               const float2 *__restrict__ b, long n,
                                               double *__restrict__ partial) {
  __shared__ double red[kT3];
  double acc = 0.0;
  for (long e = (long)blockIdx.x * kT3 + threadIdx.x; e < n; e += (long)gridDim.x * kT3) {
    const float2 v = b ? csub(a[e], b[e]) : a[e];
    acc += (double)cabs2(v);
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = kT3 / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

int blocks_for(long n) { return (int)std::max<long>(1, std::min<long>(1024, (n + kT3 - 1) / kT3)); }

class CartesianOperator3D : public Operator {
 public:
  CartesianOperator3D(int W_, int H_, int D_, int C_, const float *mask) {
    W = W_; H = H_; T = D_; C = C_;
    AV_CHECK(W > 0 && H > 0 && T > 0 && C > 0, "invalid operator dimensions");
    check_grid(Grid3{W, H, T, (long)W * H * T});
    AV_CHECK((long)C * N() < (1L << 31), "coils x volume exceeds 2^31 elements");
    const long n = N();
    if (mask) {
      std::vector<float> hm((size_t)n);
      AV_HIP(hipMemcpy(hm.data(), mask, sizeof(float) * n, hipMemcpyDeviceToHost));
      mask_norm2_ = 0.0;
      for (float v : hm) mask_norm2_ += (double)v * v;
      mask_.ensure(sizeof(float) * n);
      AV_HIP(hipMemcpy(mask_.p, hm.data(), sizeof(float) * n, hipMemcpyHostToDevice));
    }
    make_fft_plan(W, pw_, &twx_.p);
    make_fft_plan(H, ph_, &twy_.p);
    make_fft_plan(T, pd_, &twz_.p);
    AV_HIP(hipStreamSynchronize(nullptr));  // the copies above have landed
    tmp_.ensure(sizeof(float2) * (size_t)C * n);
  }
  long dense_elems() const override { return (long)C * N(); }
  long compact_elems() const override { return dense_elems(); }

  // K^H (ForwardOperation, cartesian_operator3d.cpp:64-133): the input is
  // masked into scratch, not in place.
  void forward(const float2 *kd, float2 *img, const float2 *b1, hipStream_t s) override {
    const long n = dense_elems();
    hipLaunchKernelGGL(k3_mask, dim3(blocks_for(n)), dim3(kT3), 0, s, kd, mask_.as<float>(),
                       tmp_.as<float2>(), N(), n);
    AV_LAUNCH_CHECK();
    fft3(tmp_.as<float2>(), -1, s);
    hipLaunchKernelGGL(k3_coil_sum, dim3(blocks_for(N())), dim3(kT3), 0, s, tmp_.as<float2>(), b1,
                       img, N(), C);
    AV_LAUNCH_CHECK();
  }
  // K (BackwardOperation, cartesian_operator3d.cpp:216-280)
  void backward(const float2 *img, float2 *kd, const float2 *b1, hipStream_t s) override {
    apply_K(img, nullptr, b1, kd, s);
  }
  // CartesianOperator3D::AdaptLambda (cartesian_operator3d.cpp:47-62)
  double adapt_lambda(double k, double d) const override {
    const double n = (double)N();
    const double subfac = mask_.p ? n / mask_norm2_ : 1.0;
    return subfac * k + d;
  }
  void gather(const float2 *kd, float2 *out, hipStream_t s) override {
    const long n = dense_elems();
    hipLaunchKernelGGL(k3_mask, dim3(blocks_for(n)), dim3(kT3), 0, s, kd, mask_.as<float>(), out,
                       N(), n);
    AV_LAUNCH_CHECK();
  }
  void dual_update(const float2 *u, const float2 *b1, float2 *z, const float2 *d, float2 *khz,
                   const DevScalars *sc, hipStream_t s) override {
    const long n = dense_elems();
    ku_.ensure(sizeof(float2) * (size_t)n);
    apply_K(u, nullptr, b1, ku_.as<float2>(), s);
    hipLaunchKernelGGL(k3_dual_z, dim3(blocks_for(n)), dim3(kT3), 0, s, z, ku_.as<float2>(), d, sc,
                       n);
    AV_LAUNCH_CHECK();
    forward(z, khz, b1, s);
  }
  void adjoint_compact(const float2 *z, const float2 *b1, float2 *img, hipStream_t s) override {
    forward(z, img, b1, s);
  }
  int norm2(const float2 *u, const float2 *usub, const float2 *b1, double *partial,
            hipStream_t s) override {
    const long n = dense_elems();
    ku_.ensure(sizeof(float2) * (size_t)n);
    apply_K(u, usub, b1, ku_.as<float2>(), s);
    const int nb = blocks_for(n);
    hipLaunchKernelGGL(k3_abs2, dim3(nb), dim3(kT3), 0, s, ku_.as<float2>(), nullptr, n, partial);
    AV_LAUNCH_CHECK();
    return nb;
  }
  int residual2(const float2 *u, const float2 *b1, const float2 *d, double *partial,
                hipStream_t s) override {
    const long n = dense_elems();
    ku_.ensure(sizeof(float2) * (size_t)n);
    apply_K(u, nullptr, b1, ku_.as<float2>(), s);
    const int nb = blocks_for(n);
    hipLaunchKernelGGL(k3_abs2, dim3(nb), dim3(kT3), 0, s, ku_.as<float2>(), d, n, partial);
    AV_LAUNCH_CHECK();
    return nb;
  }
  int max_partials() const override { return blocks_for(dense_elems()); }
  bool volumetric() const override { return true; }
  bool khk_fp64(const double2 *u, double2 *out, const float2 *b1, Fp64Scratch &ws,
                hipStream_t s) const override {
    cart3_khk_fp64(W, H, T, C, mask_.as<float>(), u, out, b1, ws, s);
    return true;
  }

 private:
  DevBuf mask_, twx_, twy_, twz_, tmp_, ku_;
  FftPlan pw_{}, ph_{}, pd_{};
  double mask_norm2_ = 0.0;

  // unitary 3D transform of the C volumes in place
  void fft3(float2 *v, int sign, hipStream_t s) {
    launch_fft2(v, pw_, ph_, H, W, C * T, sign, 0, 1.f, s);
    launch_fft_axis0(v, pd_, T, H * W, C, sign, (float)(1.0 / std::sqrt((double)N())), s);
  }
  void apply_K(const float2 *u, const float2 *usub, const float2 *b1, float2 *out, hipStream_t s) {
    const long n = dense_elems();
    hipLaunchKernelGGL(k3_expand, dim3(blocks_for(n)), dim3(kT3), 0, s, u, usub, b1, out, N(), n);
    AV_LAUNCH_CHECK();
    fft3(out, +1, s);
    hipLaunchKernelGGL(k3_mask, dim3(blocks_for(n)), dim3(kT3), 0, s, out, mask_.as<float>(), out,
                       N(), n);
    AV_LAUNCH_CHECK();
  }
};

}  // namespace

Operator *make_cartesian3d(int W, int H, int D, int C, const float *mask) {
  return new CartesianOperator3D(W, H, D, C, mask);
}

}  // namespace avionic
