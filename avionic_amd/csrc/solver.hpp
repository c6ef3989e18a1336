// Host side of the PD reconstruction: operators and solvers over the HIP
// kernels.  Mirrors PDRecon / TV / TGV2 / ICTGV2 (CUDA/include/pd_recon.h,
// tv.h, tgv2.h, ictgv2.h) and BaseOperator / CartesianOperator /
// NoncartesianOperator (CUDA/include/base_operator.h, cartesian_operator.h,
// noncartesian_operator.h).
#pragma once

#include <memory>
#include <string>
#include <vector>

#include "../../include/avionic.h"
#include "cartesian.hpp"
#include "common.hpp"
#include "opnorm.hpp"
#include "reg.hpp"

namespace avionic {

// Owning device allocation.
struct DevBuf {
  void *p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  ~DevBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  void ensure(size_t b) {
    if (b <= bytes && p) return;
    release();
    AV_HIP(hipMalloc(&p, b > 0 ? b : 16));
    bytes = b;
  }
  template <typename T>
  T *as() const { return (T *)p; }
};

// Abstract data operator K (image -> k-space) with the compact k-space the
// solver iterates on.
class Operator {
 public:
  int W = 0, H = 0, C = 0, T = 0;
  virtual ~Operator() = default;
  long N() const { return (long)W * H * T; }
  // Elements of the caller-visible (dense) k-space.
  virtual long dense_elems() const = 0;
  // Elements of the compact k-space the solver stores.
  virtual long compact_elems() const = 0;
  // K^H: dense k-space -> image (ForwardOperation).
  virtual void forward(const float2 *kdense, float2 *img, const float2 *b1, hipStream_t s) = 0;
  // K: image -> dense k-space (BackwardOperation).
  virtual void backward(const float2 *img, float2 *kdense, const float2 *b1, hipStream_t s) = 0;
  virtual double adapt_lambda(double k, double d) const = 0;
  // Dense (already masked) data -> compact.
  virtual void gather(const float2 *kdense, float2 *compact, hipStream_t s) = 0;
  // z = (z + sigma K u - sigma d) / (1 + sigma/lambda); khz = K^H z.
  virtual void dual_update(const float2 *u, const float2 *b1, float2 *z, const float2 *d,
                           float2 *khz, const DevScalars *sc, hipStream_t s) = 0;
  // khz = K^H z for compact z.
  virtual void adjoint_compact(const float2 *z, const float2 *b1, float2 *img, hipStream_t s) = 0;
  // Partial sums of |K (u - usub)|^2 (usub may be null); returns the count.
  virtual int norm2(const float2 *u, const float2 *usub, const float2 *b1, double *partial,
                    hipStream_t s) = 0;
  // Partial sums of |K u - d|^2 on compact d; returns the count.
  virtual int residual2(const float2 *u, const float2 *b1, const float2 *d, double *partial,
                        hipStream_t s) = 0;
  // Upper bound of the partial-sum counts above.
  virtual int max_partials() const = 0;
  // True for the volumetric operator (third axis = depth).
  virtual bool volumetric() const { return false; }
  // out += K^H K u in double precision (operator-norm power iteration);
  // false when the operator has no such path.
  virtual bool khk_fp64(const double2 *u, double2 *out, const float2 *b1, Fp64Scratch &ws,
                        hipStream_t s) const {
    (void)u; (void)out; (void)b1; (void)ws; (void)s;
    return false;
  }
};

class CartesianOperator : public Operator {
 public:
  CartesianOperator(int W, int H, int C, int T, const float *mask, int centered);
  long dense_elems() const override { return (long)C * N(); }
  long compact_elems() const override { return plan_.kspace_elems(); }
  void forward(const float2 *kdense, float2 *img, const float2 *b1, hipStream_t s) override;
  void backward(const float2 *img, float2 *kdense, const float2 *b1, hipStream_t s) override;
  double adapt_lambda(double k, double d) const override;
  void gather(const float2 *kdense, float2 *compact, hipStream_t s) override;
  void dual_update(const float2 *u, const float2 *b1, float2 *z, const float2 *d, float2 *khz,
                   const DevScalars *sc, hipStream_t s) override;
  void adjoint_compact(const float2 *z, const float2 *b1, float2 *img, hipStream_t s) override;
  int norm2(const float2 *u, const float2 *usub, const float2 *b1, double *partial,
            hipStream_t s) override;
  int residual2(const float2 *u, const float2 *b1, const float2 *d, double *partial,
                hipStream_t s) override;
  int max_partials() const override { return std::max(1, plan_.row_blocks()); }
  bool khk_fp64(const double2 *u, double2 *out, const float2 *b1, Fp64Scratch &ws,
                hipStream_t s) const override {
    cart_khk_fp64(plan_.g, u, out, b1, ws, s);
    return true;
  }
  const CartPlan &plan() const { return plan_; }

 private:
  CartPlan plan_;
  DevBuf kcol_, zhat_;
  void scratch();
};

// PD gap of a running reconstruction on its compact state: |G*(z, y) + R(x)|
// normalised by the voxel count (tv.cpp:202-235, tgv2.cpp:130-176,
// ictgv2.cpp:178-249 and the pdGap / N of ictgv2.cpp:466-467), with the
// regulariser norm R(x) and the data fidelity lambda/2 ||K x - d||
// (ictgv2.cpp:163-175) that the debug exports record beside it.  The duals
// are read in their storage format on the device.
struct GapValues {
  double gap, norm, fidelity;
};
struct GapState {
  int method = 0, ds = 0;
  Grid3 g{};
  RegParams rp{};
  double alpha0 = 0, alpha1 = 0, alpha = 0, lambda = 0;
  Operator *op = nullptr;
  const float2 *x = nullptr;  // primal, P components
  const void *y = nullptr;    // dual, D components in storage ds
  const float2 *z = nullptr, *d = nullptr, *khz = nullptr, *b1 = nullptr;  // compact z, d; K^H z
  double *part_a = nullptr, *part_b = nullptr;  // gap_partials() doubles each
};
size_t gap_partials(const Grid3 &g, const Operator &op);
GapValues compact_gap(const GapState &st, hipStream_t s);
// Regulariser norm R(x) (utils::TVNorm / TGV2Norm / ICTGV2Norm).
double regularizer_norm_of(const GapState &st, hipStream_t s);
// lambda/2 ||K x - d||_2 on the compact data (ictgv2.cpp:163-175).
double data_fidelity_of(const GapState &st, hipStream_t s);

// Primal-dual solver (PDRecon + the three regularisers).
// True for operators built by avionic_noncartesian_operator_create.
bool is_noncartesian(const Operator &op);

class Recon {
 public:
  Recon(int method, int W, int H, int C, int T, const avionic_params *params, Operator *op);
  int method;            // kernel method (kTV for temporal TV)
  bool temporal = false;  // TVTEMP: spatial weights zero, dt as given
  bool volumetric = false;  // TGV2_3D: (dx, dy, dz) as given, third axis = depth
  int W, H, C, T;
  long N;
  Operator *op;
  avionic_params params{};
  bool verbose = false, debug = false;
  int debugstep = 1;
  // exports / logs of the last run
  std::vector<double> sig_log, tau_log, pdgap_log, norm_log, fid_log;
  std::vector<float2> x3_host;

  void run(const float2 *data, float2 *x, const float2 *b1, hipStream_t s);
  // ||K_full|| by the FP64 power iteration (adaptSchedule 2), with the
  // derivative weights this solver uses.
  double operator_norm(const float2 *b1, int iterations, hipStream_t s);
  void adapt_step(const float2 *delta, const float2 *b1, hipStream_t s);
  double gstar_dense(const float2 *x1, const float2 *y, const float2 *z, const float2 *data,
                     const float2 *b1, hipStream_t s);
  double pdgap_dense(const float2 *x, const float2 *y, const float2 *z, const float2 *data,
                     const float2 *b1, hipStream_t s);
  double regularizer_norm(const float2 *x, hipStream_t s);

  static void defaults(int method, avionic_params *p);

 private:
  void prepare_weights();
  void d6(double out[6]) const;
  void ensure_partials();
  double reduce_pair(const double *partial, int n, int which, hipStream_t s);
  DevBuf part_a_, part_b_, scal_;
};

}  // namespace avionic

// The C-ABI operator handle (include/avionic.h).
struct avionic_operator {
  avionic::Operator *op;
};
