/*
 * FP64 CPU oracle for the AVIONIC primal-dual reconstruction loop.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP product
 * path is compared against, and the timed CPU baseline.  Only tests/, the
 * smoke check and the benchmark's cpu_baseline leg may load it.  The product
 * library (libavionic) never links, loads or calls anything in this directory.
 *
 * It restates, in double precision, the reference MATLAB scripts
 * (matlab/tv_dmri.m, tgv2_dmri.m, ictgv2_dmri.m, forward_opt.m, backward_opt.m
 * and their helpers) with the CUDA solver conventions (CUDA/src/tv.cpp,
 * tgv2.cpp, ictgv2.cpp, pd_recon.cpp, utils.cpp, cartesian_operator.cpp),
 * which the reference gtests pin.  Each function cites the lines it follows.
 *
 * Layout (the reference's CUDA layout, cartesian_operator.cpp:77-81,136-140):
 *   image   [t][y][x]      complex, x fastest, width W, height H, frames T
 *   k-space [t][c][y][x]   complex, C coils
 *   mask    [t][y][x]      real (binary in practice)
 *   b1      [c][y][x]      complex coil sensitivities
 * Complex arrays are interleaved (re, im) doubles.
 */
#ifndef AVIONIC_ORACLE_H_
#define AVIONIC_ORACLE_H_

#ifdef __cplusplus
extern "C" {
#endif

/* Method identifiers, same numbering as the product C-ABI. */
enum { ORC_TV = 0, ORC_TGV2 = 1, ORC_ICTGV2 = 2, ORC_TVTEMP = 3, ORC_ICTV = 4, ORC_TGV2_3D = 5 };

typedef struct orc_params {
  int method;            /* ORC_TV / ORC_TGV2 / ORC_ICTGV2 / ORC_TVTEMP / ORC_ICTV */
  int maxIt;             /* PDParams::maxIt (pd_recon.h:37) */
  double sigma, tau;     /* initial step sizes (tv.cpp:44-45) */
  double sigmaTauRatio;  /* pd_recon.h:64 */
  double tsw, tsw2;      /* timeSpaceWeight(2): (ds,dt) derived inside, like
                            ictgv2.cpp:300-306 */
  double lambda;         /* data-fidelity weight */
  double alpha0, alpha1; /* TGV2 weights (tgv2.h:11-14) */
  double alpha;          /* ICTGV2 balance (ictgv2.h:12) */
  int schedule;          /* 0: adapt when k<10 || k%50==0 (tv.cpp:172);
                            1: every iteration (tv_dmri.m:198, debug=1) */
  int gapEvery;          /* >0: record the normalised PD gap every gapEvery
                            iterations (tv.cpp:180-193, debugstep) */
  int centered;          /* 1: centred FFTs (cartesian_operator.cpp:90-99) */
  double dt;             /* TVTEMP's temporal step, used as given (tv_temp.cpp
                            :49, tvtemp.dt in options_parser.cpp:206) */
  double dx, dy, dz;     /* TGV2_3D's steps, used as given (tgv2_3d.cpp
                            InitParams) */
} orc_params;

/* ---- FFT (unnormalised, sign -1 = e^{-i...}, +1 = e^{+i...}) ---- */
void orc_fft1(double *a, int n, int sign);
void orc_fft2(double *a, int H, int W, int sign, int centered);

/* ---- 3D transform and volumetric operator pair (cartesian_operator3d.cpp):
 * arrays [D][H][W]; k-space and b1 [C][D][H][W]; K = M F3^{+}(b1 x)/sqrt(N),
 * K^H = sum_c conj(b1_c) F3^{-}(M z_c)/sqrt(N); unnormalised orc_fft3 ---- */
void orc_fft3(double *a, int D, int H, int W, int sign);
void orc_cart3_K(const double *x, const double *b1, const double *mask, int W, int H, int D,
                 int C, double *z);
void orc_cart3_KH(const double *z, const double *b1, const double *mask, int W, int H, int D,
                  int C, double *x);
int orc_recon3(const orc_params *p, int W, int H, int D, int C, const double *data,
               const double *b1, const double *mask, double *x_inout, double *x_out_all,
               double *sig_log, double *tau_log, int nthreads);

/* ---- Cartesian operator pair ---- */
void orc_cart_K(const double *x, const double *b1, const double *mask, int W,
                int H, int C, int T, int centered, double *z);
void orc_cart_KH(const double *z, const double *b1, const double *mask, int W,
                 int H, int C, int T, int centered, double *x);

/* ---- stencils (single-frame-stack, N = W*H*T) ---- */
void orc_grad(const double *u, int W, int H, int T, double dx, double dy,
              double dt, double *g);
void orc_symgrad(const double *v, int W, int H, int T, double dx, double dy,
                 double dt, double *e);
void orc_div(const double *p, int W, int H, int T, double dx, double dy,
             double dt, double *out);
void orc_symdiv(const double *q, int W, int H, int T, double dx, double dy,
                double dt, double *out);
void orc_prox3(double *y, long N, double radius);
void orc_prox6(double *y, long N, double radius);

/* ---- scalars ---- */
double orc_ellipke(double m);
int orc_time_space_weights(double tsw, double *ds, double *dt);
double orc_adapt_lambda_cart(const double *mask, int W, int H, int T, double k,
                             double d);
void orc_adapt_step(double nKx, double nx, double ratio, double *sigma,
                    double *tau);

/* ---- functionals ---- */
/* TVTEMP: temporal forward difference / 1-norm of it (tv_temp.cpp:130-147, 217-219) */
void orc_tgrad(const double *x, int W, int H, int T, double dt, double *out);
void orc_tgrad_T(const double *y, int W, int H, int T, double dt, double *out);
double orc_tvtemp_norm(const double *u, int W, int H, int T, double dt);
/* ICTV: a1 (alpha/m) TV1(u - w) + a1 ((1-alpha)/m) TV2(w), m = min(alpha,
 * 1-alpha) (utils.cpp:1221-1243) */
double orc_ictv_norm(const double *u, const double *w, int W, int H, int T, double alpha1,
                     double alpha, const double *d6);
double orc_tv_norm(const double *u, int W, int H, int T, double dx, double dy,
                   double dt);
double orc_tgv2_norm(const double *u, const double *v, int W, int H, int T,
                     double alpha0, double alpha1, double dx, double dy,
                     double dt);
double orc_ictgv2_norm(const double *u, const double *v, const double *w,
                       const double *v2, int W, int H, int T, double alpha0,
                       double alpha1, double alpha, double dx, double dy,
                       double dt, double dx2, double dy2, double dt2);
/* G* and PD gap with explicit (dx,dy,dt[,2]) like the reference methods. */
double orc_gstar(int method, const double *x1, const double *y, const double *z,
                 const double *data, const double *b1, const double *mask,
                 int W, int H, int C, int T, int centered, double lambda,
                 const double *d6);
double orc_pdgap(int method, const double *x, const double *y, const double *z,
                 const double *data, const double *b1, const double *mask,
                 int W, int H, int C, int T, int centered, double lambda,
                 double alpha0, double alpha1, double alpha, const double *d6);
/* Step-size adaptation of one solver (tv.cpp:79-97, tgv2.cpp:89-128,
   ictgv2.cpp:97-161) on primal difference `dx` (P components). */
void orc_solver_adapt_step(int method, const double *delta, const double *b1,
                           const double *mask, int W, int H, int C, int T,
                           int centered, const double *d6, double ratio,
                           double *sigma, double *tau, double *nKx, double *nx);

/* ---- the PD loop ----
 * x_inout: the image u (N complex).  x_out_all (nullable): all P primal
 * components at exit.  sig_log/tau_log (nullable): sigma/tau after each
 * iteration (maxIt entries).  gap_log (nullable, >= maxIt/gapEvery+1 entries):
 * normalised gaps; *n_gap receives the count.  Returns iterations run. */
int orc_recon(const orc_params *p, int W, int H, int C, int T,
              const double *data, const double *b1, const double *mask,
              double *x_inout, double *x_out_all, double *sig_log,
              double *tau_log, double *gap_log, int *n_gap, int nthreads);

/* Primal / dual component counts per method. */
int orc_primal_components(int method);
int orc_dual_components(int method);

/* ---- non-Cartesian (radial) operator, Kaiser-Bessel gridding ---- */
double orc_kb_beta(int width, double osf);
void orc_nufft_K(const double *x, const double *b1, const double *traj,
                 const double *dcf, int W, int H, int C, int T, int M,
                 int width, double osf, double *kdata);
void orc_nufft_KH(const double *kdata, const double *b1, const double *traj,
                  const double *dcf, int W, int H, int C, int T, int M,
                  int width, double osf, double *x);
void orc_nudft_K(const double *x, const double *b1, const double *traj,
                 const double *dcf, int W, int H, int C, int T, int M,
                 double *kdata);
/* The PD loop with the non-Cartesian operator (k-space [T][C][M]). */
int orc_recon_nc(const orc_params *p, int W, int H, int C, int T, int M, int J,
                 double osf, const double *traj, const double *dcf,
                 const double *data, const double *b1, double *x_inout,
                 double *x_out_all, double *sig_log, double *tau_log,
                 int nthreads);

/* Coil-sensitivity estimation (oracle_coil.c; coil_construction.cpp:22-557,
 * cartesian_coil_construction.cpp:23-87).  Complex arrays interleaved;
 * images [C][H][W]. */
void orc_laplace_periodic(const double *u, int W, int H, double *out);
void orc_coil_tar(const double *kdata, const double *mask, int W, int H, int C, int T,
                  int centered, int apply_phase, double *u, double *crec);
int orc_coil_b1_from_uh1(const double *u, const double *crec, int W, int H, int C, double mu,
                         double *b1);
int orc_coil_find_max_sum(const double *crec, int W, int H, int C);
void orc_coil_utgv2(const double *b1, const double *crec, const int *inds, int n_inds, int W,
                    int H, const double *p, double *x);
void orc_coil_b1_recon(const double *u0, const double *crec, int coils, int W, int H, int iters,
                       double mu, const double *p, double *x);
void orc_coil_construction_from(const double *crec, const double *u0, int W, int H, int C,
                                const double *prm, double *u, double *b1, int *order);
void orc_coil_construction(const double *kdata, const double *mask, int W, int H, int C, int T,
                           int centered, const double *prm, double *u, double *b1, int *order);

/* ||K_full|| of the stacked regulariser + data operator by power iteration
 * (matlab/getopnorm.m:1-16 on K_full^* K_full): Cartesian 2D+t, volumetric
 * (TGV2_3D) and non-Cartesian data operators.  d6 = {dx, dy, dt, dx2, dy2,
 * dt2}. */
double orc_opnorm(int method, const double *b1, const double *mask, int W, int H, int C, int T,
                  int centered, const double *d6, int iters);
double orc_opnorm3(const double *b1, const double *mask, int W, int H, int D, int C,
                   const double *d6, int iters);
double orc_opnorm_nc(int method, const double *b1, int W, int H, int C, int T, int M, int J,
                     double osf, const double *traj, const double *dcf, const double *d6,
                     int iters);

#ifdef __cplusplus
}
#endif
#endif
