// extern "C" boundary of libavionic (include/avionic.h).  No C++ exception
// crosses it: every entry point maps errors to a status code and records the
// message for avionic_last_error().
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include <rocprofiler-sdk-roctx/roctx.h>

#include "../../include/avionic.h"
#include "coil.hpp"
#include "solver.hpp"

namespace avionic {
TraceRange::TraceRange(const char *name) { roctxRangePushA(name); }
TraceRange::~TraceRange() { roctxRangePop(); }

double ellipke_host(double m);
bool time_space_weights(double tsw, double &ds, double &dt);
class NoncartesianOperator;
Operator *make_cartesian3d(int W, int H, int D, int C, const float *mask);
Operator *make_noncartesian(int W, int H, int C, int T, int nFE, int spf, const float *traj,
                            const float *dens, int kw, double osf);
void cartesian_backward_noalloc(const CartesianOperator &op, const float2 *img, float2 *kdense,
                                const float2 *b1, hipStream_t s);
}  // namespace avionic

using namespace avionic;

struct avionic_recon {
  Recon *r;
};

static thread_local std::string g_err;

void avionic_set_error(const char *msg) { g_err = msg ? msg : ""; }

template <class F>
static int guard(F &&f) {
  try {
    f();
    g_err.clear();
    return AVIONIC_OK;
  } catch (const Error &e) {
    g_err = e.what();
    return e.code;
  } catch (const std::bad_alloc &) {
    g_err = "out of memory";
    return AVIONIC_E_NOMEM;
  } catch (const std::exception &e) {
    g_err = e.what();
    return AVIONIC_E_INVALID;
  } catch (...) {
    g_err = "unknown error";
    return AVIONIC_E_INVALID;
  }
}

#define NEED(p) AV_CHECK((p) != nullptr, #p " is null")

static hipStream_t S(void *s) { return (hipStream_t)s; }
static const float2 *CF(const float *p) { return reinterpret_cast<const float2 *>(p); }
static float2 *F2(float *p) { return reinterpret_cast<float2 *>(p); }

extern "C" {

const char *avionic_last_error(void) { return g_err.c_str(); }
const char *avionic_version(void) { return "avionic-mi355x 0.1 (gfx950)"; }

int avionic_cartesian3d_operator_create(unsigned width, unsigned height, unsigned depth,
                                        unsigned coils, const float *mask, avionic_operator **op) {
  return guard([&] {
    NEED(op);
    *op = nullptr;
    auto *o = new avionic_operator{nullptr};
    try {
      o->op = make_cartesian3d((int)width, (int)height, (int)depth, (int)coils, mask);
    } catch (...) {
      delete o;
      throw;
    }
    *op = o;
  });
}

int avionic_cartesian_operator_create(unsigned width, unsigned height, unsigned coils,
                                      unsigned frames, const float *mask, int centered,
                                      avionic_operator **op) {
  return guard([&] {
    NEED(op);
    *op = nullptr;
    auto *o = new avionic_operator{nullptr};
    try {
      o->op = new CartesianOperator((int)width, (int)height, (int)coils, (int)frames, mask, centered);
    } catch (...) {
      delete o;
      throw;
    }
    *op = o;
  });
}

int avionic_noncartesian_operator_create(unsigned width, unsigned height, unsigned coils,
                                         unsigned frames, unsigned nFE, unsigned spokesPerFrame,
                                         const float *traj, const float *dens, int kernel_width,
                                         double osf, avionic_operator **op) {
  return guard([&] {
    NEED(op);
    NEED(traj);
    *op = nullptr;
    auto *o = new avionic_operator{nullptr};
    try {
      o->op = make_noncartesian((int)width, (int)height, (int)coils, (int)frames, (int)nFE,
                                (int)spokesPerFrame, traj, dens, kernel_width, osf);
    } catch (...) {
      delete o;
      throw;
    }
    *op = o;
  });
}

int avionic_operator_destroy(avionic_operator *op) {
  return guard([&] {
    if (!op) return;
    delete op->op;
    delete op;
  });
}

int avionic_operator_forward(avionic_operator *op, const float *kdata, float *image,
                             const float *b1, void *stream) {
  return guard([&] {
    NEED(op); NEED(kdata); NEED(image); NEED(b1);
    op->op->forward(CF(kdata), F2(image), CF(b1), S(stream));
  });
}

int avionic_operator_backward(avionic_operator *op, const float *image, float *kdata,
                              const float *b1, void *stream) {
  return guard([&] {
    NEED(op); NEED(kdata); NEED(image); NEED(b1);
    op->op->backward(CF(image), F2(kdata), CF(b1), S(stream));
  });
}

int avionic_operator_backward_noalloc(avionic_operator *op, const float *image, float *kdata,
                                      const float *b1, void *stream) {
  return guard([&] {
    NEED(op); NEED(kdata); NEED(image); NEED(b1);
    const auto *cop = dynamic_cast<const CartesianOperator *>(op->op);
    if (!cop)
      throw Error(AVIONIC_E_UNSUPPORTED,
                  "avionic_operator_backward_noalloc: Cartesian 2D+t operators only");
    cartesian_backward_noalloc(*cop, CF(image), F2(kdata), CF(b1), S(stream));
  });
}

int avionic_operator_adapt_lambda(avionic_operator *op, double k, double d, double *lambda) {
  return guard([&] {
    NEED(op); NEED(lambda);
    *lambda = op->op->adapt_lambda(k, d);
  });
}

size_t avionic_operator_kspace_size(const avionic_operator *op) {
  if (!op || !op->op) return 0;
  return (size_t)(op->op->dense_elems() / ((long)op->op->C * op->op->T));
}

int avionic_default_params(avionic_method method, avionic_params *params) {
  return guard([&] {
    NEED(params);
    AV_CHECK(method >= 0 && method <= AVIONIC_TGV2_3D, "unknown method");
    Recon::defaults(kernel_method((int)method), params);
    if (method == AVIONIC_TGV2_3D) params->alpha0 = std::sqrt(3.0);  // tgv2_3d.cpp InitParams
  });
}

int avionic_recon_create(avionic_method method, unsigned width, unsigned height, unsigned coils,
                         unsigned frames, const avionic_params *params, avionic_operator *op,
                         avionic_recon **recon) {
  return guard([&] {
    NEED(recon); NEED(op);
    *recon = nullptr;
    auto *r = new Recon((int)method, (int)width, (int)height, (int)coils, (int)frames, params, op->op);
    *recon = new avionic_recon{r};
  });
}

int avionic_recon_destroy(avionic_recon *recon) {
  return guard([&] {
    if (!recon) return;
    delete recon->r;
    delete recon;
  });
}

int avionic_recon_get_params(const avionic_recon *recon, avionic_params *params) {
  return guard([&] {
    NEED(recon); NEED(params);
    *params = recon->r->params;
  });
}

int avionic_recon_set_params(avionic_recon *recon, const avionic_params *params) {
  return guard([&] {
    NEED(recon); NEED(params);
    AV_CHECK(params->dualStorage >= 0 && params->dualStorage <= 2,
             "dualStorage must be 0, 1 or 2");
    recon->r->params = *params;
  });
}

int avionic_recon_set_verbose(avionic_recon *recon, int verbose) {
  return guard([&] { NEED(recon); recon->r->verbose = verbose != 0; });
}

int avionic_recon_set_debug(avionic_recon *recon, int debug, int debugstep) {
  return guard([&] {
    NEED(recon);
    AV_CHECK(debugstep >= 1, "debugstep must be >= 1");
    recon->r->debug = debug != 0;
    recon->r->debugstep = debugstep;
  });
}

int avionic_recon_adapt_lambda(avionic_recon *recon, double k, double d, double *lambda) {
  return guard([&] {
    NEED(recon); NEED(lambda);
    *lambda = recon->r->op->adapt_lambda(k, d);
  });
}

int avionic_recon_run(avionic_recon *recon, const float *data, float *x, const float *b1,
                      void *stream) {
  return guard([&] {
    NEED(recon); NEED(data); NEED(x); NEED(b1);
    recon->r->run(CF(data), F2(x), CF(b1), S(stream));
  });
}

int avionic_recon_operator_norm(avionic_recon *recon, const float *b1, unsigned iterations,
                                double *norm, void *stream) {
  return guard([&] {
    NEED(recon); NEED(b1); NEED(norm);
    *norm = recon->r->operator_norm(CF(b1), (int)iterations, S(stream));
  });
}

int avionic_recon_adapt_step_size(avionic_recon *recon, const float *delta, const float *b1,
                                  void *stream) {
  return guard([&] {
    NEED(recon); NEED(delta); NEED(b1);
    recon->r->adapt_step(CF(delta), CF(b1), S(stream));
  });
}

int avionic_recon_compute_gstar(avionic_recon *recon, const float *x1, const float *y,
                                const float *z, const float *data, const float *b1, double *gstar,
                                void *stream) {
  return guard([&] {
    NEED(recon); NEED(x1); NEED(y); NEED(z); NEED(data); NEED(b1); NEED(gstar);
    *gstar = recon->r->gstar_dense(CF(x1), CF(y), CF(z), CF(data), CF(b1), S(stream));
  });
}

int avionic_recon_compute_pdgap(avionic_recon *recon, const float *x, const float *y,
                                const float *z, const float *data, const float *b1, double *pdgap,
                                void *stream) {
  return guard([&] {
    NEED(recon); NEED(x); NEED(y); NEED(z); NEED(data); NEED(b1); NEED(pdgap);
    *pdgap = recon->r->pdgap_dense(CF(x), CF(y), CF(z), CF(data), CF(b1), S(stream));
  });
}

int avionic_recon_get_step_log(const avionic_recon *recon, double *sigma, double *tau, size_t *n) {
  return guard([&] {
    NEED(recon); NEED(n);
    const size_t m = std::min(*n, recon->r->sig_log.size());
    for (size_t i = 0; i < m; ++i) {
      if (sigma) sigma[i] = recon->r->sig_log[i];
      if (tau) tau[i] = recon->r->tau_log[i];
    }
    *n = m;
  });
}

int avionic_recon_export(avionic_recon *recon, avionic_export_cb cb, void *user) {
  return guard([&] {
    NEED(recon); NEED(cb);
    Recon *r = recon->r;
    auto emit = [&](const char *name, const std::vector<double> &v) {
      std::vector<float> tmp(2 * v.size());
      for (size_t i = 0; i < v.size(); ++i) {
        tmp[2 * i] = (float)v[i];
        tmp[2 * i + 1] = 0.f;
      }
      cb(name, tmp.data(), v.size(), user);
    };
    if ((r->method == kICTGV2 || r->method == kICTV) && !r->x3_host.empty())
      cb("x3_component", reinterpret_cast<const float *>(r->x3_host.data()), r->x3_host.size(), user);
    if (r->debug) {
      emit("PDGap", r->pdgap_log);
      if (r->method == kICTGV2 || r->method == kICTV) {  // ictv.cpp:370-388
        emit(r->method == kICTV ? "ictvnorm" : "ICTGVnorm", r->norm_log);
        emit("DATAfidelity", r->fid_log);
      }
    }
  });
}

// ----------------------------------------------------------------- utils
int avionic_gradient(const float *u, float *g, unsigned width, unsigned height, unsigned frames,
                     float dx, float dy, float dt, void *stream) {
  return guard([&] {
    NEED(u); NEED(g);
    const Grid3 gr{(int)width, (int)height, (int)frames, (long)width * height * frames};
    launch_gradient(gr, 1.f / dx, 1.f / dy, 1.f / dt, CF(u), F2(g), S(stream));
  });
}

int avionic_symmetric_gradient(const float *v, float *e, unsigned width, unsigned height,
                               unsigned frames, float dx, float dy, float dt, void *stream) {
  return guard([&] {
    NEED(v); NEED(e);
    const Grid3 gr{(int)width, (int)height, (int)frames, (long)width * height * frames};
    launch_symgrad(gr, 1.f / dx, 1.f / dy, 1.f / dt, CF(v), F2(e), S(stream));
  });
}

int avionic_divergence(const float *p, float *d, unsigned width, unsigned height, unsigned frames,
                       float dx, float dy, float dt, void *stream) {
  return guard([&] {
    NEED(p); NEED(d);
    const Grid3 gr{(int)width, (int)height, (int)frames, (long)width * height * frames};
    launch_divergence(gr, 1.f / dx, 1.f / dy, 1.f / dt, CF(p), F2(d), S(stream));
  });
}

int avionic_symmetric_divergence(const float *q, float *d, unsigned width, unsigned height,
                                 unsigned frames, float dx, float dy, float dt, void *stream) {
  return guard([&] {
    NEED(q); NEED(d);
    const Grid3 gr{(int)width, (int)height, (int)frames, (long)width * height * frames};
    launch_symdiv(gr, 1.f / dx, 1.f / dy, 1.f / dt, CF(q), F2(d), S(stream));
  });
}

int avionic_proximal_map3(float *y, size_t n, float scale, void *stream) {
  return guard([&] { NEED(y); launch_prox(3, F2(y), (long)n, scale, S(stream)); });
}

int avionic_proximal_map6(float *y, size_t n, float scale, void *stream) {
  return guard([&] { NEED(y); launch_prox(6, F2(y), (long)n, scale, S(stream)); });
}

static double norm_via(int method, const float2 *x, unsigned W, unsigned H, unsigned T,
                       double a0, double a1, double alpha, const double d6[6], hipStream_t s) {
  const long N = (long)W * H * T;
  const Grid3 g{(int)W, (int)H, (int)T, N};
  const RegParams rp = make_reg_params(method, d6, a0, a1, alpha);
  DevBuf part;
  part.ensure(sizeof(double) * 2 * (size_t)norm_blocks(g));
  launch_reg_norm(method, g, rp, a0, a1, alpha, x, part.as<double>(), s);
  return reduce_host(part.as<double>(), norm_blocks(g), 2, 0, s);
}

int avionic_tv_norm(const float *u, unsigned width, unsigned height, unsigned frames, float dx,
                    float dy, float dt, double *norm, void *stream) {
  return guard([&] {
    NEED(u); NEED(norm);
    const double d6[6] = {dx, dy, dt, 1, 1, 1};
    *norm = norm_via(kTV, CF(u), width, height, frames, 0, 0, 0, d6, S(stream));
  });
}

int avionic_tgv2_norm(const float *u, const float *v, unsigned width, unsigned height,
                      unsigned frames, double alpha0, double alpha1, float dx, float dy, float dt,
                      double *norm, void *stream) {
  return guard([&] {
    NEED(u); NEED(v); NEED(norm);
    const long N = (long)width * height * frames;
    DevBuf x;
    x.ensure(sizeof(float2) * 4 * N);
    AV_HIP(hipMemcpyAsync(x.p, u, sizeof(float2) * N, hipMemcpyDeviceToDevice, S(stream)));
    AV_HIP(hipMemcpyAsync(x.as<float2>() + N, v, sizeof(float2) * 3 * N, hipMemcpyDeviceToDevice, S(stream)));
    const double d6[6] = {dx, dy, dt, 1, 1, 1};
    *norm = norm_via(kTGV2, x.as<float2>(), width, height, frames, alpha0, alpha1, 0, d6, S(stream));
  });
}

int avionic_ictgv2_norm(const float *u, const float *v, const float *w, const float *v2,
                        unsigned width, unsigned height, unsigned frames, double alpha0,
                        double alpha1, double alpha, float dx, float dy, float dt, float dx2,
                        float dy2, float dt2, double *norm, void *stream) {
  return guard([&] {
    NEED(u); NEED(v); NEED(w); NEED(v2); NEED(norm);
    const long N = (long)width * height * frames;
    DevBuf x;
    x.ensure(sizeof(float2) * 8 * N);
    hipStream_t s = S(stream);
    AV_HIP(hipMemcpyAsync(x.p, u, sizeof(float2) * N, hipMemcpyDeviceToDevice, s));
    AV_HIP(hipMemcpyAsync(x.as<float2>() + N, v, sizeof(float2) * 3 * N, hipMemcpyDeviceToDevice, s));
    AV_HIP(hipMemcpyAsync(x.as<float2>() + 4 * N, w, sizeof(float2) * N, hipMemcpyDeviceToDevice, s));
    AV_HIP(hipMemcpyAsync(x.as<float2>() + 5 * N, v2, sizeof(float2) * 3 * N, hipMemcpyDeviceToDevice, s));
    const double d6[6] = {dx, dy, dt, dx2, dy2, dt2};
    *norm = norm_via(kICTGV2, x.as<float2>(), width, height, frames, alpha0, alpha1, alpha, d6, s);
  });
}

double avionic_ellipke(double m) { return ellipke_host(m); }

int avionic_time_space_weights(double tsw, double *ds, double *dt) {
  return guard([&] {
    NEED(ds); NEED(dt);
    AV_CHECK(time_space_weights(tsw, *ds, *dt), "ComputeTimeSpaceWeights: Invalid ratio.");
  });
}

int avionic_fft2(float *data, unsigned height, unsigned width, unsigned batch, int sign,
                 int centered, int unitary, void *stream) {
  return guard([&] {
    NEED(data);
    AV_CHECK(sign == 1 || sign == -1, "sign must be +1 or -1");
    AV_CHECK(height >= 1 && width >= 1 && batch >= 1, "empty transform");
    FftPlan pw, ph;
    DevBuf tw1, tw2;
    make_fft_plan((int)width, pw, &tw1.p);
    make_fft_plan((int)height, ph, &tw2.p);
    const float sc = unitary ? (float)(1.0 / std::sqrt((double)width * height)) : 1.f;
    launch_fft2(F2(data), pw, ph, (int)height, (int)width, (int)batch, sign, centered, sc, S(stream));
    AV_HIP(hipStreamSynchronize(S(stream)));
  });
}

// ------------------------------------------------------- coil construction
static const Operator &coil_op(avionic_operator *op) {
  NEED(op);
  if (!dynamic_cast<const CartesianOperator *>(op->op) && !is_noncartesian(*op->op))
    throw Error(AVIONIC_E_UNSUPPORTED,
                "coil construction needs a Cartesian 2D+t or a non-Cartesian operator");
  return *op->op;
}
static avionic_coil_params coil_params(const avionic_coil_params *p) {
  avionic_coil_params d;
  coil_default_params(&d);
  return p ? *p : d;
}

int avionic_default_coil_params(avionic_coil_params *params) {
  return guard([&] {
    NEED(params);
    coil_default_params(params);
  });
}

int avionic_coil_time_averaged(avionic_operator *op, const float *kdata, float *u, float *crec,
                               int apply_phase, void *stream) {
  return guard([&] {
    NEED(kdata); NEED(u); NEED(crec);
    coil_time_averaged(coil_op(op), CF(kdata), F2(u), F2(crec), apply_phase != 0, S(stream));
  });
}

int avionic_coil_b1_from_uh1(unsigned width, unsigned height, unsigned coils, const float *u,
                             const float *crec, double mu, float *b1, void *stream) {
  return guard([&] {
    NEED(u); NEED(crec); NEED(b1);
    AV_CHECK(width && height && coils, "empty image");
    const CoilCtx k{(int)width, (int)height, (int)coils, (long)width * height, S(stream)};
    coil_b1_from_uh1(k, CF(u), CF(crec), mu, F2(b1));
  });
}

int avionic_coil_find_max_sum(unsigned width, unsigned height, unsigned coils, const float *crec,
                              int *index, void *stream) {
  return guard([&] {
    NEED(crec); NEED(index);
    AV_CHECK(width && height && coils, "empty image");
    const CoilCtx k{(int)width, (int)height, (int)coils, (long)width * height, S(stream)};
    *index = coil_find_max_sum(k, CF(crec));
  });
}

int avionic_coil_utgv2(unsigned width, unsigned height, unsigned coils, const float *b1,
                       const float *crec, const int *inds, int n_inds,
                       const avionic_coil_params *params, float *x, void *stream) {
  return guard([&] {
    NEED(b1); NEED(crec); NEED(inds); NEED(x);
    AV_CHECK(width && height && coils, "empty image");
    AV_CHECK(n_inds >= 1 && n_inds <= (int)coils, "n_inds must be in 1..coils");
    std::vector<int> v(inds, inds + n_inds);
    for (int i : v) AV_CHECK(i >= 0 && i < (int)coils, "coil index out of range");
    const CoilCtx k{(int)width, (int)height, (int)coils, (long)width * height, S(stream)};
    coil_utgv2(k, CF(b1), CF(crec), v, coil_params(params), F2(x));
  });
}

int avionic_coil_b1_recon(unsigned width, unsigned height, unsigned coils, const float *u0,
                          const float *crec, unsigned maxIt, double mu,
                          const avionic_coil_params *params, float *b1, void *stream) {
  return guard([&] {
    NEED(u0); NEED(crec); NEED(b1);
    AV_CHECK(width && height && coils, "empty image");
    AV_CHECK(mu > 0.0, "mu must be positive");
    const CoilCtx k{(int)width, (int)height, (int)coils, (long)width * height, S(stream)};
    coil_b1_recon(k, CF(u0), CF(crec), (int)coils, maxIt, mu, coil_params(params), F2(b1));
  });
}

int avionic_coil_construction(avionic_operator *op, const float *kdata,
                              const avionic_coil_params *params, float *u, float *b1, int *order,
                              void *stream) {
  return guard([&] {
    NEED(kdata); NEED(u); NEED(b1);
    const Operator &c = coil_op(op);
    const CoilCtx k{c.W, c.H, c.C, (long)c.W * c.H, S(stream)};
    DevBuf u0, crec;
    u0.ensure(sizeof(float2) * k.N);
    crec.ensure(sizeof(float2) * k.N * k.C);
    coil_time_averaged(c, CF(kdata), u0.as<float2>(), crec.as<float2>(), true, k.s);
    std::vector<int> ord;
    coil_construction_from(k, crec.as<float2>(), u0.as<float2>(), coil_params(params), F2(u),
                           F2(b1), &ord);
    if (order)
      for (int i = 0; i < k.C; ++i) order[i] = ord[i];
  });
}

int avionic_coil_construction_from_images(unsigned width, unsigned height, unsigned coils,
                                          const float *crec, const float *u0,
                                          const avionic_coil_params *params, float *u,
                                          float *b1, int *order, void *stream) {
  return guard([&] {
    NEED(crec); NEED(u0); NEED(u); NEED(b1);
    AV_CHECK(width && height && coils, "empty image");
    const CoilCtx k{(int)width, (int)height, (int)coils, (long)width * height, S(stream)};
    std::vector<int> ord;
    coil_construction_from(k, CF(crec), CF(u0), coil_params(params), F2(u), F2(b1), &ord);
    if (order)
      for (int i = 0; i < k.C; ++i) order[i] = ord[i];
  });
}

}  // extern "C"
