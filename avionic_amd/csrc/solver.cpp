// PD solver host logic: the iteration sequence, step-size schedule and
// functional / gap evaluation over the fused HIP kernels.
#include "solver.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>

namespace avionic {

// ============================================================ Cartesian op
CartesianOperator::CartesianOperator(int W_, int H_, int C_, int T_, const float *mask,
                                     int centered) {
  W = W_; H = H_; C = C_; T = T_;
  plan_.init(W, H, C, T, mask, centered, nullptr);
}

void CartesianOperator::scratch() {
  const size_t b = sizeof(float2) * (size_t)std::max(1L, compact_elems());
  kcol_.ensure(b);
  zhat_.ensure(b);
}

// cartesian_operator.cpp:47-62: lambda = k * W H T / ||mask||^2 + d.
double CartesianOperator::adapt_lambda(double k, double d) const {
  double subfac = (double)W * H * T;
  if (plan_.has_mask)
    subfac /= plan_.mask_norm2;
  else
    subfac /= (double)W * H * T;
  return subfac * k + d;
}

void CartesianOperator::gather(const float2 *kdense, float2 *compact, hipStream_t s) {
  launch_compact_gather(plan_.g, kdense, compact, s);
}

void CartesianOperator::backward(const float2 *img, float2 *kdense, const float2 *b1,
                                 hipStream_t s) {
  scratch();
  AV_HIP(hipMemsetAsync(kdense, 0, sizeof(float2) * (size_t)dense_elems(), s));
  launch_col_forward(plan_.g, img, nullptr, b1, kcol_.as<float2>(), s);
  launch_row_pass(kRowApplyK, plan_.g, kcol_.as<float2>(), nullptr, nullptr, nullptr, kdense,
                  nullptr, nullptr, nullptr, s);
}

void CartesianOperator::forward(const float2 *kdense, float2 *img, const float2 *b1,
                                hipStream_t s) {
  scratch();
  launch_row_pass(kRowAdjDense, plan_.g, nullptr, nullptr, nullptr, zhat_.as<float2>(), nullptr,
                  kdense, nullptr, nullptr, s);
  launch_col_adjoint(plan_.g, zhat_.as<float2>(), b1, img, s);
}

void CartesianOperator::dual_update(const float2 *u, const float2 *b1, float2 *z, const float2 *d,
                                    float2 *khz, const DevScalars *sc, hipStream_t s) {
  scratch();
  launch_col_forward(plan_.g, u, nullptr, b1, kcol_.as<float2>(), s);
  launch_row_pass(kRowDualUpdate, plan_.g, kcol_.as<float2>(), z, d, zhat_.as<float2>(), nullptr,
                  nullptr, sc, nullptr, s);
  launch_col_adjoint(plan_.g, zhat_.as<float2>(), b1, khz, s);
}

void CartesianOperator::adjoint_compact(const float2 *z, const float2 *b1, float2 *img,
                                        hipStream_t s) {
  scratch();
  launch_row_pass(kRowAdjCompact, plan_.g, nullptr, nullptr, nullptr, zhat_.as<float2>(), nullptr,
                  z, nullptr, nullptr, s);
  launch_col_adjoint(plan_.g, zhat_.as<float2>(), b1, img, s);
}

int CartesianOperator::norm2(const float2 *u, const float2 *usub, const float2 *b1,
                             double *partial, hipStream_t s) {
  scratch();
  launch_col_forward(plan_.g, u, usub, b1, kcol_.as<float2>(), s);
  launch_row_pass(kRowNorm2, plan_.g, kcol_.as<float2>(), nullptr, nullptr, nullptr, nullptr,
                  nullptr, nullptr, partial, s);
  return plan_.row_blocks();
}

int CartesianOperator::residual2(const float2 *u, const float2 *b1, const float2 *d,
                                 double *partial, hipStream_t s) {
  scratch();
  launch_col_forward(plan_.g, u, nullptr, b1, kcol_.as<float2>(), s);
  launch_row_pass(kRowResidual, plan_.g, kcol_.as<float2>(), nullptr, d, nullptr, nullptr,
                  nullptr, nullptr, partial, s);
  return plan_.row_blocks();
}

// ================================================================== gap

size_t gap_partials(const Grid3 &g, const Operator &op) {
  const int n = std::max({step_blocks(g), norm_blocks(g), op.max_partials(),
                          reduce_blocks(op.compact_elems())});
  return 2 * (size_t)n;
}

double regularizer_norm_of(const GapState &st, hipStream_t s) {
  launch_reg_norm(st.method, st.g, st.rp, st.alpha0, st.alpha1, st.alpha, st.x, st.part_a, s);
  return reduce_host(st.part_a, norm_blocks(st.g), 2, 0, s);
}

double data_fidelity_of(const GapState &st, hipStream_t s) {
  const int nk = st.op->residual2(st.x, st.b1, st.d, st.part_b, s);
  return st.lambda / 2.0 * std::sqrt(reduce_host(st.part_b, nk, 1, 0, s));
}

GapValues compact_gap(const GapState &st, hipStream_t s) {
  // G* = lambda/2 ||K x - d||^2 + Re<d, z> + ||z||^2 / (2 lambda) + dual-residual l1 terms
  const int nk = st.op->residual2(st.x, st.b1, st.d, st.part_b, s);
  const double r2 = reduce_host(st.part_b, nk, 1, 0, s);
  const long K = st.op->compact_elems();
  launch_kspace_dots(st.d, st.z, K, st.part_a, s);
  const int nb = reduce_blocks(K);
  const double re_dz = reduce_host(st.part_a, nb, 2, 0, s);
  const double zz = reduce_host(st.part_a, nb, 2, 1, s);
  launch_gstar_terms(st.method, st.ds, st.g, st.rp, st.khz, st.y, st.part_a, s);
  const double l1 = reduce_host(st.part_a, norm_blocks(st.g), 2, 0, s);
  const double lam = st.lambda;
  const double gstar = 0.5 * lam * r2 + re_dz + zz / (2.0 * lam) + l1;
  const double r = regularizer_norm_of(st, s);
  GapValues v;
  v.gap = std::fabs(gstar + r) / (double)st.g.N;
  v.norm = r;
  v.fidelity = lam / 2.0 * std::sqrt(r2);
  return v;
}

// ================================================================== Recon

static double ellipke(double m) {
  if (m < 0.0 || m > 1.0) return NAN;
  if (m == 1.0) return 1.0;
  double a = 1.0, b = std::sqrt(1.0 - m), s = m, pw = 1.0;
  for (int i = 0; i < 64; ++i) {
    const double c = 0.5 * (a - b), an = 0.5 * (a + b);
    b = std::sqrt(a * b);
    a = an;
    pw *= 2.0;
    const double w = pw * c * c;
    s += w;
    if (w < 1e-17) break;
  }
  return (M_PI / (2.0 * a)) * (1.0 - 0.5 * s);
}

double ellipke_host(double m) { return ellipke(m); }

// PDRecon::ComputeTimeSpaceWeights, pd_recon.cpp:72-98.
bool time_space_weights(double tsw, double &ds, double &dt) {
  double r = 1.0 / tsw, w1, w2;
  if (r > 0.0 && r <= 1.0) {
    w2 = M_PI / (2.0 * ellipke(1.0 - r * r));
    w1 = r * w2;
  } else if (r > 1.0) {
    r = 1.0 / r;
    w1 = M_PI / (2.0 * ellipke(1.0 - r * r));
    w2 = r * w1;
  } else {
    return false;
  }
  ds = 1.0 / w1;
  dt = 1.0 / w2;
  return true;
}

// InitParams: tv.cpp:39-57, tgv2.cpp:36-56, ictgv2.cpp:37-64, ictv.cpp:37-62
// TVTEMP (tv_temp.cpp:39-54) sets the TV values it uses (dt = 1).
void Recon::defaults(int method, avionic_params *p) {
  *p = avionic_params{};
  p->maxIt = 500;
  p->stopPDGap = 0;
  p->sigma = 1.0 / 3.0;
  p->tau = 1.0 / 3.0;
  p->sigmaTauRatio = 1.0;
  p->timeSpaceWeight = 5.0;
  p->dx = p->dy = p->dt = p->dz = 1.0;
  p->ds = 1.0;
  p->alpha0 = std::sqrt(2.0);
  p->alpha1 = 1.0;
  p->alpha = 0.5;
  p->timeSpaceWeight2 = 1.0;
  p->dx2 = p->dy2 = p->dt2 = 1.0;
  p->adaptLambda = 1;
  p->adaptLambdaK = 0.4 * 0.2991;
  p->adaptLambdaD = 10.0 * 0.2991;
  p->lambda = 0.0;
  if (method == kICTGV2) {
    p->timeSpaceWeight = 6.5;
    p->timeSpaceWeight2 = 2.5;
    p->alpha = 0.6423;
    p->adaptLambdaK = 0.4 / (5.1191 * 0.35);
    p->adaptLambdaD = 10.0 / (5.1191 * 0.35);
  }
  if (method == kICTV) {
    p->timeSpaceWeight = 4.0;
    p->timeSpaceWeight2 = 0.5;
    p->alpha = 0.5;
    p->alpha1 = 1.0;
    p->adaptLambdaK = 0.4 / (5.1191 * 0.35);
    p->adaptLambdaD = 10.0 / (5.1191 * 0.35);
  }
}

Recon::Recon(int method_, int W_, int H_, int C_, int T_, const avionic_params *p, Operator *op_)
    : method(kernel_method(method_)), temporal(method_ == kApiTVTEMP),
      volumetric(method_ == kApiTGV2_3D), W(W_), H(H_), C(C_), T(T_), N((long)W_ * H_ * T_),
      op(op_) {
  AV_CHECK(method_ >= 0 && method_ <= kApiTGV2_3D, "unknown method");
  AV_CHECK(op != nullptr, "operator is null");
  AV_CHECK(op->W == W && op->H == H && op->C == C && op->T == T,
           "operator dimensions do not match the solver");
  AV_CHECK(op->volumetric() == volumetric,
           "TGV2_3D runs on the volumetric operator, and that operator only with TGV2_3D");
  if (p) {
    params = *p;
    if (params.adaptLambda) params.lambda = op->adapt_lambda(params.adaptLambdaK, params.adaptLambdaD);
  } else {
    defaults(method, &params);
    if (volumetric) params.alpha0 = std::sqrt(3.0);  // tgv2_3d.cpp InitParams
    params.lambda = op->adapt_lambda(params.adaptLambdaK, params.adaptLambdaD);  // InitLambda(true)
  }
}

void Recon::d6(double out[6]) const {
  out[0] = params.dx; out[1] = params.dy; out[2] = params.dt;
  if (temporal) out[0] = out[1] = INFINITY;  // 1/dx = 1/dy = 0
  if (volumetric) out[2] = params.dz;         // third axis: depth
  out[3] = params.dx2; out[4] = params.dy2; out[5] = params.dt2;
}

void Recon::ensure_partials() {
  const Grid3 g{W, H, T, N};
  const size_t n = std::max(gap_partials(g, *op), (size_t)2 * reduce_blocks((long)C * N));
  part_a_.ensure(sizeof(double) * n);
  part_b_.ensure(sizeof(double) * n);
  scal_.ensure(sizeof(DevScalars));
}

double Recon::reduce_pair(const double *partial, int n, int which, hipStream_t s) {
  return reduce_host(partial, n, 2, which, s);
}

static void upload_scalars(DevScalars *d, const avionic_params &p, hipStream_t s) {
  DevScalars h{};
  h.sigma = p.sigma;
  h.tau = p.tau;
  h.lambda = p.lambda;
  h.ratio = p.sigmaTauRatio;
  AV_HIP(hipMemcpyAsync(d, &h, sizeof(h), hipMemcpyHostToDevice, s));
}

// AdaptStepSize (tv.cpp:79-97, tgv2.cpp:89-128, ictgv2.cpp:97-161): delta holds
// P primal components; x = 0 so that ext - x = delta.
void Recon::adapt_step(const float2 *delta, const float2 *b1, hipStream_t s) {
  ensure_partials();
  const int P = primal_components(method);
  DevBuf zero;
  zero.ensure(sizeof(float2) * (size_t)P * N);
  AV_HIP(hipMemsetAsync(zero.p, 0, zero.bytes, s));
  double dd[6];
  d6(dd);
  const RegParams rp = make_reg_params(method, dd, params.alpha0, params.alpha1, params.alpha);
  const Grid3 g{W, H, T, N};
  DevScalars *sc = scal_.as<DevScalars>();
  upload_scalars(sc, params, s);
  launch_step_reg(method, g, rp, delta, zero.as<float2>(), part_a_.as<double>(), s);
  const int nk = op->norm2(delta, nullptr, b1, part_b_.as<double>(), s);
  launch_step_finalize(part_a_.as<double>(), step_blocks(g), part_b_.as<double>(), nk, sc, nullptr,
                       nullptr, 0, s);
  DevScalars h{};
  AV_HIP(hipMemcpyAsync(&h, sc, sizeof(h), hipMemcpyDeviceToHost, s));
  AV_HIP(hipStreamSynchronize(s));
  params.sigma = h.sigma;
  params.tau = h.tau;
}

double Recon::regularizer_norm(const float2 *x, hipStream_t s) {
  ensure_partials();
  double dd[6];
  d6(dd);
  const RegParams rp = make_reg_params(method, dd, params.alpha0, params.alpha1, params.alpha);
  const Grid3 g{W, H, T, N};
  launch_reg_norm(method, g, rp, params.alpha0, params.alpha1, params.alpha, x,
                  part_a_.as<double>(), s);
  return reduce_pair(part_a_.as<double>(), norm_blocks(g), 0, s);
}

// G* on caller-provided dense arrays (tv.cpp:202-225, tgv2.cpp:130-165,
// ictgv2.cpp:178-232).
double Recon::gstar_dense(const float2 *x1, const float2 *y, const float2 *z, const float2 *data,
                          const float2 *b1, hipStream_t s) {
  ensure_partials();
  const long CN = op->dense_elems();
  DevBuf kx, img;
  kx.ensure(sizeof(float2) * (size_t)CN);
  img.ensure(sizeof(float2) * (size_t)N);
  // g1 = lambda/2 ||K x - d||^2 over the dense k-space
  op->backward(x1, kx.as<float2>(), b1, s);
  launch_kspace_dots(data, z, CN, part_a_.as<double>(), s);
  const int nb = reduce_blocks(CN);
  const double re_dz = reduce_pair(part_a_.as<double>(), nb, 0, s);
  const double zz = reduce_pair(part_a_.as<double>(), nb, 1, s);
  // |Kx - d|^2 = sum over (Kx - d) via the dots kernel on the difference
  std::vector<float2> hk(CN), hd(CN);
  AV_HIP(hipMemcpyAsync(hk.data(), kx.p, sizeof(float2) * CN, hipMemcpyDeviceToHost, s));
  AV_HIP(hipMemcpyAsync(hd.data(), data, sizeof(float2) * CN, hipMemcpyDeviceToHost, s));
  AV_HIP(hipStreamSynchronize(s));
  double r2 = 0.0;
  for (long i = 0; i < CN; ++i) {
    const double a = (double)hk[i].x - hd[i].x, b = (double)hk[i].y - hd[i].y;
    r2 += a * a + b * b;
  }
  const double lam = params.lambda;
  double g = 0.5 * lam * r2 + re_dz + zz / (2.0 * lam);
  // dual-residual l1 terms with K^H z
  op->forward(z, img.as<float2>(), b1, s);
  double dd[6];
  d6(dd);
  const RegParams rp = make_reg_params(method, dd, params.alpha0, params.alpha1, params.alpha);
  const Grid3 gr{W, H, T, N};
  launch_gstar_terms(method, kDualF32, gr, rp, img.as<float2>(), y, part_b_.as<double>(), s);
  g += reduce_pair(part_b_.as<double>(), norm_blocks(gr), 0, s);
  return g;
}

double Recon::pdgap_dense(const float2 *x, const float2 *y, const float2 *z, const float2 *data,
                          const float2 *b1, hipStream_t s) {
  const double gs = gstar_dense(x, y, z, data, b1, s);
  const double r = regularizer_norm(x, s);
  return std::fabs(gs + r);
}

// Derivative weights of a run: from the time/space ratios (tv.cpp:103-104,
// ictgv2.cpp:300-306), dt as given for TVTEMP, (dx, dy, dz) for TGV2_3D.
void Recon::prepare_weights() {
  double ds, dt;
  if (temporal) {
    // TVTEMP uses dt as configured, no time/space weights (tv_temp.cpp:98)
    AV_CHECK(params.dt > 0.0, "dt must be positive");
  } else if (volumetric) {
    // TGV2_3D uses dx, dy, dz as configured (tgv2_3d.cpp loop)
    AV_CHECK(params.dx > 0.0 && params.dy > 0.0 && params.dz > 0.0, "dx, dy, dz must be positive");
  } else {
    AV_CHECK(time_space_weights(params.timeSpaceWeight, ds, dt),
             "ComputeTimeSpaceWeights: Invalid ratio.");
    params.dx = params.dy = ds;
    params.dt = dt;
  }
  if (method == kICTGV2 || method == kICTV) {
    AV_CHECK(time_space_weights(params.timeSpaceWeight2, ds, dt),
             "ComputeTimeSpaceWeights: Invalid ratio.");
    params.dx2 = params.dy2 = ds;
    params.dt2 = dt;
  }
}

double Recon::operator_norm(const float2 *b1, int iterations, hipStream_t s) {
  prepare_weights();
  double dd[6];
  d6(dd);
  return operator_norm_fp64(method, W, H, T, dd, *op, b1, iterations, s);
}

// IterativeReconstruction: tv.cpp:99-200, tgv2.cpp:180-338, ictgv2.cpp:294-492.
void Recon::run(const float2 *data, float2 *xio, const float2 *b1, hipStream_t s) {
  TraceRange tr("avionic_recon_run");
  prepare_weights();
  AV_CHECK(params.adaptSchedule >= 0 && params.adaptSchedule <= 2, "adaptSchedule must be 0, 1 or 2");
  if (params.adaptSchedule == 2) {
    // fixed steps sigma = tau = 0.99 / ||K_full|| (FP64 power iteration)
    const int its = params.opnormIterations ? (int)params.opnormIterations : kOpnormDefaultIterations;
    const double nk = operator_norm(b1, its, s);
    AV_CHECK(nk > 0.0, "operator norm is zero");
    params.sigma = params.tau = 0.99 / nk;
  }
  AV_CHECK(params.lambda > 0.0, "lambda must be positive");
  ensure_partials();
  const int P = primal_components(method), D = dual_components(method);
  const int dstore = params.dualStorage;
  const long K = op->compact_elems();
  DevBuf x, ext, y, z, d, khz, slog, tlog;
  x.ensure(sizeof(float2) * (size_t)P * N);
  ext.ensure(sizeof(float2) * (size_t)P * N);
  y.ensure(dual_bytes(dstore) * (size_t)D * N);
  z.ensure(sizeof(float2) * (size_t)std::max(1L, K));
  d.ensure(sizeof(float2) * (size_t)std::max(1L, K));
  khz.ensure(sizeof(float2) * (size_t)N);
  const unsigned maxIt = params.maxIt;
  slog.ensure(sizeof(double) * std::max(1u, maxIt));
  tlog.ensure(sizeof(double) * std::max(1u, maxIt));

  AV_HIP(hipMemsetAsync(x.p, 0, x.bytes, s));
  AV_HIP(hipMemsetAsync(y.p, 0, y.bytes, s));
  AV_HIP(hipMemsetAsync(z.p, 0, z.bytes, s));
  AV_HIP(hipMemcpyAsync(x.p, xio, sizeof(float2) * N, hipMemcpyDeviceToDevice, s));
  AV_HIP(hipMemcpyAsync(ext.p, x.p, x.bytes, hipMemcpyDeviceToDevice, s));
  op->gather(data, d.as<float2>(), s);
  DevScalars *sc = scal_.as<DevScalars>();
  upload_scalars(sc, params, s);

  double dd[6];
  d6(dd);
  const RegParams rp = make_reg_params(method, dd, params.alpha0, params.alpha1, params.alpha);
  const Grid3 g{W, H, T, N};
  float2 *xs = x.as<float2>(), *es = ext.as<float2>();
  std::vector<char> adapted(maxIt, 0);
  pdgap_log.clear();
  norm_log.clear();
  fid_log.clear();

  GapState gs;
  gs.method = method;
  gs.ds = dstore;
  gs.g = g;
  gs.rp = rp;
  gs.alpha0 = params.alpha0;
  gs.alpha1 = params.alpha1;
  gs.alpha = params.alpha;
  gs.lambda = params.lambda;
  gs.op = op;
  gs.x = xs;
  gs.y = y.p;
  gs.z = z.as<float2>();
  gs.d = d.as<float2>();
  gs.khz = khz.as<float2>();
  gs.b1 = b1;
  gs.part_a = part_a_.as<double>();
  gs.part_b = part_b_.as<double>();
  if (method == kICTGV2) {
    fid_log.push_back(data_fidelity_of(gs, s));
    norm_log.push_back(regularizer_norm_of(gs, s));
  }

  unsigned k = 0;
  for (; k < maxIt; ++k) {
    // dual ascent + projections (regulariser)
    launch_dual(method, dstore, g, rp, es, y.p, sc, s);
    // data term: z update, prox and K^H z
    op->dual_update(es, b1, z.as<float2>(), d.as<float2>(), khz.as<float2>(), sc, s);
    // primal descent + extragradient
    launch_primal(method, dstore, g, rp, xs, es, y.p, khz.as<float2>(), sc, s);
    const bool adapt = params.adaptSchedule == 1 ||
                       (params.adaptSchedule == 0 && (k < 10 || k % 50 == 0));
    if (adapt) {
      adapted[k] = 1;
      launch_step_reg(method, g, rp, es, xs, part_a_.as<double>(), s);
      const int nk = op->norm2(es, xs, b1, part_b_.as<double>(), s);
      launch_step_finalize(part_a_.as<double>(), step_blocks(g), part_b_.as<double>(), nk, sc,
                           slog.as<double>(), tlog.as<double>(), (int)k, s);
    }
    const bool gap = (verbose && (k < 10 || k % 50 == 0)) || (debug && k % std::max(1, debugstep) == 0) ||
                     (params.stopPDGap > 0 && k % 20 == 0);
    if (gap) {
      gs.x = xs;
      const GapValues gv = compact_gap(gs, s);
      pdgap_log.push_back(gv.gap);
      if (verbose) std::printf("Normalized Primal-Dual Gap after %u iterations: %.4e\n", k, gv.gap);
      if (method == kICTGV2 || method == kICTV) {  // ictv.cpp:347-355 too
        norm_log.push_back(gv.norm);
        fid_log.push_back(gv.fidelity);
      }
      if (params.stopPDGap > 0 && gv.gap < params.stopPDGap) {
        ++k;
        break;
      }
    }
  }
  const unsigned ran = k;
  // results
  AV_HIP(hipMemcpyAsync(xio, xs, sizeof(float2) * N, hipMemcpyDeviceToDevice, s));
  DevScalars h{};
  AV_HIP(hipMemcpyAsync(&h, sc, sizeof(h), hipMemcpyDeviceToHost, s));
  std::vector<double> hs(std::max(1u, maxIt)), ht(std::max(1u, maxIt));
  AV_HIP(hipMemcpyAsync(hs.data(), slog.p, sizeof(double) * hs.size(), hipMemcpyDeviceToHost, s));
  AV_HIP(hipMemcpyAsync(ht.data(), tlog.p, sizeof(double) * ht.size(), hipMemcpyDeviceToHost, s));
  if (method == kICTGV2 || method == kICTV) {  // w: component 4 (ICTGV2) or 1 (ICTV)
    x3_host.resize(N);
    AV_HIP(hipMemcpyAsync(x3_host.data(), xs + (method == kICTV ? 1 : 4) * N, sizeof(float2) * N,
                          hipMemcpyDeviceToHost, s));
  }
  AV_HIP(hipStreamSynchronize(s));
  sig_log.assign(ran, 0.0);
  tau_log.assign(ran, 0.0);
  double cs = params.sigma, ct = params.tau;
  for (unsigned i = 0; i < ran; ++i) {
    if (adapted[i]) {
      cs = hs[i];
      ct = ht[i];
    }
    sig_log[i] = cs;
    tau_log[i] = ct;
  }
  if (h.diverged) throw Error(AVIONIC_E_DIVERGED, "NaN/Inf in the step-size reduction");
  params.sigma = h.sigma;
  params.tau = h.tau;
}

}  // namespace avionic
