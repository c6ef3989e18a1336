// NoncartesianOperator (CUDA/src/noncartesian_operator.cpp:34-185) with the
// operator definition of matlab/backward_opt_noncart.m:2-21 and
// forward_opt_noncart.m:2-18: K applies the coil maps and a type-2 NUFFT with
// sqrt(density) weighting, K^H is its exact adjoint.  The NUFFT is
// Kaiser-Bessel gridding (kernel width J <= 3, grid oversampling osf,
// beta = pi sqrt((J/osf)^2 (osf - 1/2)^2 - 0.8) of Beatty et al., IEEE TMI
// 2005), deapodised by the kernel's continuous Fourier transform; the GPU
// passes are described in noncart.hpp.
//
// The trajectory is static per operator, so the tables the passes use (the
// separable tap weights of every point, the row -> point list of the forward
// row pass and the cell -> point list of the adjoint's spreading) are built
// once on the host here and kept on the device.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <vector>

#include "cartesian.hpp"
#include "noncart.hpp"
#include "solver.hpp"

namespace avionic {

namespace {

double bessel_i0(double x) {
  double s = 1.0, t = 1.0;
  const double q = 0.25 * x * x;
  for (int k = 1; k < 200; ++k) {
    t *= q / ((double)k * k);
    s += t;
    if (t < 1e-17 * s) break;
  }
  return s;
}

// Launch shapes of the gridding passes that are picked per geometry: the
// coil groups of the forward column pass (CG, blocks per image tile) and the
// image rows per adjoint row block (RBa).  No output depends on them (every
// element is computed the same way whatever the shape), so they are tuned by
// timing: the first operator of a geometry runs each candidate a few times
// on scratch data and keeps the fastest for the process.
// AVIONIC_NC_TILES="cg,rows" forces a shape; AVIONIC_NC_TUNE=0 keeps the
// defaults (4 coil groups, 4 rows; the fastest at the 256x256x30x12 radial
// geometry when measured by hand on MI355X).
constexpr int kNcAdjRows = 4;
constexpr int kNcFwdCoilGroups = 4;

struct NcTiles {
  int cg, rows;
};
std::mutex g_tune_mu;
std::map<std::array<long, 7>, NcTiles> g_tune_cache;

// Mean duration in ms of `reps` runs of f on stream s (after one warm-up).
template <class F>
double time_ms(F &&f, hipStream_t s, int reps) {
  hipEvent_t a, b;
  AV_HIP(hipEventCreate(&a));
  AV_HIP(hipEventCreate(&b));
  f();
  AV_HIP(hipEventRecord(a, s));
  for (int r = 0; r < reps; ++r) f();
  AV_HIP(hipEventRecord(b, s));
  AV_HIP(hipEventSynchronize(b));
  float ms = 0.f;
  AV_HIP(hipEventElapsedTime(&ms, a, b));
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  return ms / reps;
}

NcTiles tune_tiles(NcGeom g, float2 *tmp, float2 *part) {
  const char *force = std::getenv("AVIONIC_NC_TILES");
  if (force) {
    NcTiles t{kNcFwdCoilGroups, kNcAdjRows};
    if (std::sscanf(force, "%d,%d", &t.cg, &t.rows) == 2) {
      t.cg = std::max(1, std::min(t.cg, g.C));
      t.rows = std::max(1, std::min(t.rows, std::max(1, 4096 / g.Gx)));
      return t;
    }
  }
  const NcTiles def{std::max(1, std::min(g.C, kNcFwdCoilGroups)),
                    std::max(1, std::min(4096 / g.Gx, kNcAdjRows))};
  const char *tune = std::getenv("AVIONIC_NC_TUNE");
  if (tune && std::strcmp(tune, "0") == 0) return def;
  const std::array<long, 7> key{g.W, g.H, g.C, g.T, g.M, g.Gx, g.Gy};
  std::lock_guard<std::mutex> lock(g_tune_mu);
  auto it = g_tune_cache.find(key);
  if (it != g_tune_cache.end()) return it->second;
  // scratch inputs (zeros: the passes' work does not depend on the values)
  DevBuf img, b1, z;
  img.ensure(sizeof(float2) * (size_t)g.T * g.H * g.W);
  b1.ensure(sizeof(float2) * (size_t)g.C * g.H * g.W);
  z.ensure(sizeof(float2) * (size_t)g.T * g.C * g.M);
  hipStream_t s;
  AV_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  AV_HIP(hipMemsetAsync(img.p, 0, img.bytes, s));
  AV_HIP(hipMemsetAsync(b1.p, 0, b1.bytes, s));
  AV_HIP(hipMemsetAsync(z.p, 0, z.bytes, s));
  NcTiles best = def;
  double bt = 1e30;
  for (int cg : {1, 2, 4, 8}) {
    if (cg > g.C) break;
    g.CG = cg;
    const double t = time_ms([&] {
      launch_nc_forward_grid(g, img.as<float2>(), nullptr, b1.as<float2>(), tmp, part, s);
    }, s, 3);
    if (t < bt) { bt = t; best.cg = cg; }
  }
  g.CG = best.cg;
  bt = 1e30;
  for (int rows : {1, 2, 4, 8}) {
    if (rows * g.Gx > 4096) break;
    g.RBa = rows;
    const double t = time_ms([&] {
      launch_nc_adjoint(g, z.as<float2>(), b1.as<float2>(), tmp, img.as<float2>(), s);
    }, s, 3);
    if (t < bt) { bt = t; best.rows = rows; }
  }
  AV_HIP(hipStreamSynchronize(s));
  (void)hipStreamDestroy(s);
  g_tune_cache[key] = best;
  return best;
}

int pick_cols(int L, int n) {
  int xb = 1;
  while (xb * 2 * L <= 4096 && xb * 2 <= 64 && xb < n) xb *= 2;
  return xb;
}

class NoncartesianOperator : public Operator {
 public:
  NoncartesianOperator(int W_, int H_, int C_, int T_, int nFE, int spf, const float *traj,
                       const float *dens, int J, double osf)
      : spf_(spf), nfe_(nFE) {
    W = W_; H = H_; C = C_; T = T_;
    AV_CHECK(W > 0 && H > 0 && C > 0 && T > 0 && nFE > 0 && spf > 0,
             "invalid operator dimensions");
    AV_CHECK(J >= 1 && J <= 3, "kernel width must be 1, 2 or 3");
    AV_CHECK(osf >= 1.0 && osf <= 4.0, "grid oversampling must be in [1, 4]");
    const long M = (long)nFE * spf;
    AV_CHECK(M < (1L << 26), "too many trajectory points per frame");
    const int Gx = (int)std::lround(osf * W), Gy = (int)std::lround(osf * H);
    AV_CHECK(Gx >= 4 && Gy >= 4, "oversampled grid too small (at least 4 x 4)");
    AV_CHECK(Gx <= kMaxFftLen && Gy <= kMaxFftLen,
             "oversampled grid exceeds the longest FFT (1024)");
    const long G2 = (long)Gx * Gy;
    AV_CHECK((long)T * M * kNcTaps * kNcTaps < (1L << 31), "trajectory tables exceed 2^31 entries");

    const double b2 = std::pow((double)J / osf * (osf - 0.5), 2) - 0.8;
    AV_CHECK(b2 > 0.0, "kernel width too small for this grid oversampling (beta undefined)");
    const double beta = M_PI * std::sqrt(b2);
    auto kb = [&](double u) {
      const double r = 2.0 * u / J;
      return std::fabs(r) > 1.0 ? 0.0 : bessel_i0(beta * std::sqrt(1.0 - r * r));
    };
    auto kb_ft = [&](double nu) {
      const double a = M_PI * J * nu, d = beta * beta - a * a;
      if (d > 1e-12) return J * std::sinh(std::sqrt(d)) / std::sqrt(d);
      if (d < -1e-12) return J * std::sin(std::sqrt(-d)) / std::sqrt(-d);
      return (double)J;
    };
    auto wrap = [](long q, long G) {
      long r = q % G;
      return r < 0 ? r + G : r;
    };

    std::vector<float> dxi(W), dyi(H);
    for (int i = 0; i < W; ++i) dxi[i] = (float)(1.0 / kb_ft((double)(i - W / 2) / Gx));
    for (int i = 0; i < H; ++i) dyi[i] = (float)(1.0 / kb_ft((double)(i - H / 2) / Gy));

    std::vector<float> tr((size_t)T * 2 * M), dn((size_t)T * M, 1.f);
    AV_HIP(hipMemcpy(tr.data(), traj, sizeof(float) * tr.size(), hipMemcpyDeviceToHost));
    if (dens) AV_HIP(hipMemcpy(dn.data(), dens, sizeof(float) * dn.size(), hipMemcpyDeviceToHost));
    traj_h_ = tr;
    dens_h_ = dn;
    const double sc = 1.0 / std::sqrt((double)W * H);

    std::vector<float> sw((size_t)T * M);
    std::vector<int2> q0((size_t)T * M);
    std::vector<float4> wx4((size_t)T * M), wy4((size_t)T * M);
    std::vector<int> rptr((size_t)T * (Gy + 1));
    std::vector<int4> rhdr;
    std::vector<float4> rwx;
    std::vector<int> cptr((size_t)T * (G2 + 1));
    std::vector<int2> ce;
    // double-precision copies for the FP64 power iteration (opnorm.hpp)
    std::vector<double> w64((size_t)T * M * 2 * kNcTaps), sw64((size_t)T * M), ce_w64;
    std::vector<double> dxi64(W), dyi64(H);
    for (int i = 0; i < W; ++i) dxi64[i] = 1.0 / kb_ft((double)(i - W / 2) / Gx);
    for (int i = 0; i < H; ++i) dyi64[i] = 1.0 / kb_ft((double)(i - H / 2) / Gy);
    rhdr.reserve((size_t)T * M * kNcTaps);
    rwx.reserve((size_t)T * M * kNcTaps);
    for (int t = 0; t < T; ++t) {
      std::vector<double> wxd((size_t)M * kNcTaps), wyd((size_t)M * kNcTaps);
      std::vector<long> qx((size_t)M), qy((size_t)M);
      for (long m = 0; m < M; ++m) {
        const size_t p = (size_t)t * M + m;
        const double d = dn[p];
        AV_CHECK(std::isfinite(d) && d >= 0.0, "density weights must be finite and >= 0");
        sw[p] = (float)(std::sqrt(d) * sc);
        sw64[p] = std::sqrt(d) * sc;
        const double sx = (double)tr[((size_t)t * 2 + 0) * M + m] * Gx;
        const double sy = (double)tr[((size_t)t * 2 + 1) * M + m] * Gy;
        AV_CHECK(std::isfinite(sx) && std::isfinite(sy), "trajectory must be finite");
        qx[m] = (long)std::ceil(sx - J / 2.0);
        qy[m] = (long)std::ceil(sy - J / 2.0);
        float w4x[4], w4y[4];
        for (int j = 0; j < kNcTaps; ++j) {
          wxd[m * kNcTaps + j] = kb(sx - (double)(qx[m] + j));
          wyd[m * kNcTaps + j] = kb(sy - (double)(qy[m] + j));
          w4x[j] = (float)wxd[m * kNcTaps + j];
          w4y[j] = (float)wyd[m * kNcTaps + j];
          w64[p * 2 * kNcTaps + j] = wxd[m * kNcTaps + j];
          w64[p * 2 * kNcTaps + kNcTaps + j] = wyd[m * kNcTaps + j];
        }
        q0[p] = make_int2((int)wrap(qx[m], Gx), (int)wrap(qy[m], Gy));
        wx4[p] = make_float4(w4x[0], w4x[1], w4x[2], w4x[3]);
        wy4[p] = make_float4(w4y[0], w4y[1], w4y[2], w4y[3]);
      }
      // row -> (point, tap): every point lists all kNcTaps rows (a zero
      // weight row still writes its partial, so none is left stale)
      std::vector<std::vector<int2>> rows(Gy);
      for (long m = 0; m < M; ++m)
        for (int dy = 0; dy < kNcTaps; ++dy)
          rows[wrap(qy[m] + dy, Gy)].push_back(make_int2((int)m, dy));
      int *rp = &rptr[(size_t)t * (Gy + 1)];
      for (int r = 0; r < Gy; ++r) {
        rp[r] = (int)rhdr.size();
        for (const int2 &md : rows[r]) {
          const size_t pi = (size_t)t * M + md.x;
          const float wyd = (&wy4[pi].x)[md.y];
          rhdr.push_back(make_int4(md.x * kNcTaps + md.y, r, q0[pi].x, __builtin_bit_cast(int, wyd)));
          rwx.push_back(wx4[pi]);
        }
      }
      rp[Gy] = (int)rhdr.size();
      // cell -> (point, weight) for the spreading, points in increasing order
      std::vector<int> count(G2, 0);
      for (long m = 0; m < M; ++m)
        for (int dy = 0; dy < kNcTaps; ++dy)
          for (int dx = 0; dx < kNcTaps; ++dx)
            if (wyd[m * kNcTaps + dy] * wxd[m * kNcTaps + dx] != 0.0)
              count[wrap(qy[m] + dy, Gy) * Gx + wrap(qx[m] + dx, Gx)]++;
      int *cp = &cptr[(size_t)t * (G2 + 1)];
      cp[0] = (int)ce.size();
      for (long g = 0; g < G2; ++g) cp[g + 1] = cp[g] + count[g];
      ce.resize(cp[G2]);
      ce_w64.resize(cp[G2]);
      std::vector<int> fill(G2, 0);
      for (long m = 0; m < M; ++m)
        for (int dy = 0; dy < kNcTaps; ++dy)
          for (int dx = 0; dx < kNcTaps; ++dx) {
            const double w = wyd[m * kNcTaps + dy] * wxd[m * kNcTaps + dx];
            if (w == 0.0) continue;
            const long cell = wrap(qy[m] + dy, Gy) * Gx + wrap(qx[m] + dx, Gx);
            const int pos = cp[cell] + fill[cell]++;
            const float wf = (float)(w * std::sqrt((double)dn[(size_t)t * M + m]) * sc);
            ce[pos] = make_int2((int)m, __builtin_bit_cast(int, wf));
            ce_w64[pos] = w * std::sqrt((double)dn[(size_t)t * M + m]) * sc;
          }
    }

    auto up = [](DevBuf &b, const void *src, size_t bytes) {
      b.ensure(std::max<size_t>(bytes, 16));
      if (bytes) AV_HIP(hipMemcpy(b.p, src, bytes, hipMemcpyHostToDevice));
    };
    up(dxi_, dxi.data(), sizeof(float) * dxi.size());
    up(dyi_, dyi.data(), sizeof(float) * dyi.size());
    up(sw_, sw.data(), sizeof(float) * sw.size());
    up(rptr_, rptr.data(), sizeof(int) * rptr.size());
    up(rhdr_, rhdr.data(), sizeof(int4) * rhdr.size());
    up(rwx_, rwx.data(), sizeof(float4) * rwx.size());
    up(cptr_, cptr.data(), sizeof(int) * cptr.size());
    up(ce_, ce.data(), sizeof(int2) * ce.size());
    up(q64_, q0.data(), sizeof(int2) * q0.size());
    up(w64_, w64.data(), sizeof(double) * w64.size());
    up(sw64_, sw64.data(), sizeof(double) * sw64.size());
    up(cew64_, ce_w64.data(), sizeof(double) * ce_w64.size());
    up(dxi64_, dxi64.data(), sizeof(double) * dxi64.size());
    up(dyi64_, dyi64.data(), sizeof(double) * dyi64.size());
    make_fft_plan(Gx, g_.px, &twx_.p);
    make_fft_plan(Gy, g_.py, &twy_.p);
    // The tables come from pageable host memory: such a copy may return before
    // its data has landed.  The row pass writes partials at indices it reads
    // from them, and the launch-shape timing below (like a caller's
    // non-blocking stream) does not wait for the copies by itself.
    AV_HIP(hipStreamSynchronize(nullptr));

    g_.W = W; g_.H = H; g_.C = C; g_.T = T; g_.M = (int)M;
    g_.Gx = Gx; g_.Gy = Gy;
    g_.wide = (plan_is_wide(g_.px) || plan_is_wide(g_.py)) ? 1 : 0;
    g_.XBf = pick_cols(Gy, W);
    g_.XBa = pick_cols(Gy, Gx);
    g_.RBf = std::max(1, 4096 / Gx);
    g_.RBa = std::max(1, std::min(4096 / Gx, kNcAdjRows));
    g_.CG = std::max(1, std::min(C, kNcFwdCoilGroups));
    g_.scale = (float)sc;
    g_.dxi = dxi_.as<float>();
    g_.dyi = dyi_.as<float>();
    g_.sw = sw_.as<float>();
    g_.rptr = rptr_.as<int>();
    g_.rhdr = rhdr_.as<int4>();
    g_.rwx = rwx_.as<float4>();
    g_.cptr = cptr_.as<int>();
    g_.ce = ce_.as<int2>();
    {
      // the spreading entries regrouped by adjoint column tile (NcGeom::tce)
      const int XB = g_.XBa, nX = (Gx + XB - 1) / XB, ncell = XB * Gy;
      std::vector<int> tstart((size_t)T * nX), tcell((size_t)T * nX * (ncell + 1));
      std::vector<int2> tce;
      tce.reserve(ce.size());
      for (int t = 0; t < T; ++t) {
        const int *cp = &cptr[(size_t)t * (G2 + 1)];
        for (int X = 0; X < nX; ++X) {
          tstart[(size_t)t * nX + X] = (int)tce.size();
          int *tc = &tcell[((size_t)t * nX + X) * (ncell + 1)];
          int n = 0;
          for (int r = 0; r < Gy; ++r)
            for (int b = 0; b < XB; ++b) {
              tc[r * XB + b] = n;
              const int gx = X * XB + b;
              if (gx >= Gx) continue;
              const long cell = (long)r * Gx + gx;
              for (int e = cp[cell]; e < cp[cell + 1]; ++e) tce.push_back(ce[e]);
              n += cp[cell + 1] - cp[cell];
            }
          tc[ncell] = n;
        }
      }
      up(tstart_, tstart.data(), sizeof(int) * tstart.size());
      up(tcell_, tcell.data(), sizeof(int) * tcell.size());
      up(tce_, tce.data(), sizeof(int2) * tce.size());
      AV_HIP(hipStreamSynchronize(nullptr));
      g_.tstart = tstart_.as<int>();
      g_.tcell = tcell_.as<int>();
      g_.tce = tce_.as<int2>();
    }
    scratch();
    const NcTiles tiles = tune_tiles(g_, tmp_.as<float2>(), part_.as<float2>());
    g_.CG = tiles.cg;
    g_.RBa = tiles.rows;
  }

  long dense_elems() const override { return (long)T * C * g_.M; }
  long compact_elems() const override { return dense_elems(); }

  // noncartesian_operator.cpp:88-95: lambda = k W / (spokesPerFrame pi / 2) + d
  double adapt_lambda(double k, double d) const override {
    return (double)W / (spf_ * M_PI / 2.0) * k + d;
  }

  void backward(const float2 *img, float2 *kd, const float2 *b1, hipStream_t s) override {
    grid_pass(img, nullptr, b1, s);
    launch_nc_samples(kNcApplyK, g_, part_.as<float2>(), kd, nullptr, nullptr, nullptr, nullptr, s);
  }
  void forward(const float2 *kd, float2 *img, const float2 *b1, hipStream_t s) override {
    adjoint_compact(kd, b1, img, s);
  }
  void gather(const float2 *kd, float2 *compact, hipStream_t s) override {
    AV_HIP(hipMemcpyAsync(compact, kd, sizeof(float2) * dense_elems(), hipMemcpyDeviceToDevice, s));
  }
  void dual_update(const float2 *u, const float2 *b1, float2 *z, const float2 *d, float2 *khz,
                   const DevScalars *sc, hipStream_t s) override {
    grid_pass(u, nullptr, b1, s);
    launch_nc_samples(kNcDualUpdate, g_, part_.as<float2>(), nullptr, z, d, sc, nullptr, s);
    adjoint_compact(z, b1, khz, s);
  }
  void adjoint_compact(const float2 *z, const float2 *b1, float2 *img, hipStream_t s) override {
    scratch();
    launch_nc_adjoint(g_, z, b1, tmp_.as<float2>(), img, s);
  }
  int norm2(const float2 *u, const float2 *usub, const float2 *b1, double *partial,
            hipStream_t s) override {
    grid_pass(u, usub, b1, s);
    return launch_nc_samples(kNcNorm2, g_, part_.as<float2>(), nullptr, nullptr, nullptr, nullptr,
                             partial, s);
  }
  int residual2(const float2 *u, const float2 *b1, const float2 *d, double *partial,
                hipStream_t s) override {
    grid_pass(u, nullptr, b1, s);
    return launch_nc_samples(kNcResidual, g_, part_.as<float2>(), nullptr, nullptr, d, nullptr,
                             partial, s);
  }
  int max_partials() const override { return nc_sample_blocks(g_); }
  bool khk_fp64(const double2 *u, double2 *out, const float2 *b1, Fp64Scratch &ws,
                hipStream_t s) const override {
    Nc64 n{W, H, C, T, g_.M, g_.Gx, g_.Gy, q64_.as<int2>(), w64_.as<double>(),
           sw64_.as<double>(), dxi64_.as<double>(), dyi64_.as<double>(), cptr_.as<int>(),
           ce_.as<int2>(), cew64_.as<double>()};
    nc_khk_fp64(n, u, out, b1, ws, s);
    return true;
  }

  // The trajectory and density as given (host copies) and the geometry, for
  // the all-frames operator of the time-averaged coil images.
  int spokes_per_frame() const { return spf_; }
  int readout_points() const { return nfe_; }
  const std::vector<float> &traj_host() const { return traj_h_; }
  const std::vector<float> &dens_host() const { return dens_h_; }

 private:
  int spf_, nfe_;
  std::vector<float> traj_h_, dens_h_;
  NcGeom g_{};
  DevBuf dxi_, dyi_, sw_, rptr_, rhdr_, rwx_, cptr_, ce_, tstart_, tcell_, tce_, twx_, twy_, tmp_,
      part_;
  DevBuf q64_, w64_, sw64_, cew64_, dxi64_, dyi64_;  // FP64 tables

  void scratch() {
    const size_t planes = (size_t)T * C;
    tmp_.ensure(sizeof(float2) * planes * std::max((size_t)g_.Gy * W, (size_t)H * g_.Gx));
    part_.ensure(sizeof(float2) * planes * g_.M * kNcTaps);
  }
  void grid_pass(const float2 *u, const float2 *usub, const float2 *b1, hipStream_t s) {
    scratch();
    launch_nc_forward_grid(g_, u, usub, b1, tmp_.as<float2>(), part_.as<float2>(), s);
  }
};

}  // namespace

Operator *make_noncartesian(int W, int H, int C, int T, int nFE, int spf, const float *traj,
                            const float *dens, int kw, double osf) {
  return new NoncartesianOperator(W, H, C, T, nFE, spf, traj, dens, kw, osf);
}

// One coil, one frame holding every frame's points (frame-major), kernel
// width 3 and grid oversampling 2: the operator of NoncartesianCoilConstruction::
// TimeAveragedReconstruction (noncartesian_coil_construction.cpp:26-86).
// nullptr when `op` is not a non-Cartesian operator.
bool is_noncartesian(const Operator &op) {
  return dynamic_cast<const NoncartesianOperator *>(&op) != nullptr;
}

Operator *make_noncartesian_all_frames(const Operator &op) {
  const auto *nc = dynamic_cast<const NoncartesianOperator *>(&op);
  if (!nc) return nullptr;
  const int T = nc->T, spf = nc->spokes_per_frame(), nFE = nc->readout_points();
  const size_t M = (size_t)spf * nFE, P = (size_t)T * M;
  const std::vector<float> &tr = nc->traj_host();  // [T][2][M]
  std::vector<float> all(2 * P);                  // [2][T*M]
  for (size_t t = 0; t < (size_t)T; ++t)
    for (size_t m = 0; m < M; ++m) {
      all[t * M + m] = tr[2 * t * M + m];
      all[P + t * M + m] = tr[2 * t * M + M + m];
    }
  DevBuf dtr, dd;
  dtr.ensure(sizeof(float) * all.size());
  dd.ensure(sizeof(float) * P);
  AV_HIP(hipMemcpy(dtr.p, all.data(), sizeof(float) * all.size(), hipMemcpyHostToDevice));
  AV_HIP(hipMemcpy(dd.p, nc->dens_host().data(), sizeof(float) * P, hipMemcpyHostToDevice));
  return new NoncartesianOperator(nc->W, nc->H, 1, 1, nFE, spf * T, dtr.as<float>(), dd.as<float>(),
                                  3, 2.0);
}

}  // namespace avionic
