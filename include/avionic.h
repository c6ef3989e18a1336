/*
 * avionic.h — C-ABI of the MI355X-native AVIONIC primal-dual reconstruction
 * library (libavionic.so, HIP kernels for gfx950).
 *
 * Drop-in boundary for the reference's L3 solver / L2 operator / L1 utility
 * APIs.  Every entry point names the reference interface it replaces.
 * Paths are relative to the AVIONIC source tree.
 *
 * Conventions
 *  - All array pointers are DEVICE pointers (the reference's CVector/RVector
 *    are GPU vectors, CUDA/include/types.h:17-23).  Complex data are
 *    interleaved float pairs (std::complex<float> / hipFloatComplex layout).
 *  - Layout (CUDA/src/cartesian_operator.cpp:77-81,136-140; main.cpp:509-519):
 *      image   [frames][height][width]           width fastest
 *      k-space [frames][coils][height][width]
 *      mask    [frames][height][width]           float
 *      coils   [coils][height][width]
 *    A multi-component primal/dual vector is its components one after the
 *    other, each N = width*height*frames long (the reference's
 *    std::vector<CVector> of equal-length vectors, concatenated).
 *  - `stream` is a hipStream_t passed as void* (NULL = default stream).
 *    Calls enqueue work on `stream`; functions that return a host scalar
 *    synchronise that stream.
 *  - Errors: every function returns AVIONIC_OK (0) or a negative status; no
 *    C++ exception crosses this boundary (the reference throws
 *    std::invalid_argument / runtime_error, pd_recon.cpp:94, utils.cpp:21).
 *    avionic_last_error() returns a thread-local message for the last failure.
 */
#ifndef AVIONIC_H_
#define AVIONIC_H_

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define AVIONIC_OK 0
#define AVIONIC_E_INVALID (-1)   /* bad argument (invalid_argument in the reference) */
#define AVIONIC_E_HIP (-2)       /* HIP runtime error */
#define AVIONIC_E_NOMEM (-3)     /* device allocation failed */
#define AVIONIC_E_DIVERGED (-4)  /* NaN/Inf in a step-size or gap reduction */
#define AVIONIC_E_CAPACITY (-5)  /* serving: request does not fit the HBM budget */
#define AVIONIC_E_NOTFOUND (-6)  /* serving: unknown request id */
#define AVIONIC_E_UNSUPPORTED (-7)

typedef enum avionic_method {
  AVIONIC_TV = 0,     /* TV     (CUDA/include/tv.h)     */
  AVIONIC_TGV2 = 1,   /* TGV2   (CUDA/include/tgv2.h)   */
  AVIONIC_ICTGV2 = 2, /* ICTGV2 (CUDA/include/ictgv2.h) */
  AVIONIC_TVTEMP = 3, /* TVTEMP, temporal TV (CUDA/include/tv_temp.h): dt is
                         used as given, dx, dy and timeSpaceWeight are unused */
  AVIONIC_ICTV = 4,   /* ICTV (CUDA/include/ictv.h): u, w; alpha1, alpha,
                         timeSpaceWeight and timeSpaceWeight2 apply */
  AVIONIC_TGV2_3D = 5 /* TGV2_3D (CUDA/include/tgv2_3d.h): TGV2 on a volume with the
                         operator of avionic_cartesian3d_operator_create, "frames" =
                         depth; dx, dy, dz apply; eager solver only (not the engine) */
} avionic_method;

/* Parameters: the union of PDParams (pd_recon.h:34-74), TGV2Params
 * (tgv2.h:9-16) and ICTGV2Params (ictgv2.h:9-21), same field names, held in
 * double.  dx/dy/dt (and dx2/dy2/dt2) are overwritten from timeSpaceWeight(2)
 * when a reconstruction starts, as the reference does (tv.cpp:103-104,
 * ictgv2.cpp:300-306).  The last two fields have no reference counterpart. */
typedef struct avionic_params {
  unsigned maxIt;
  double dx, dy, dz;
  double stopPDGap;
  double ds, dt;
  double sigma, tau, sigmaTauRatio;
  double timeSpaceWeight;
  double lambda;
  int adaptLambda;             /* AdaptLambdaParams (pd_recon.h:18-31) */
  double adaptLambdaK, adaptLambdaD;
  double alpha0, alpha1;       /* TGV2 */
  double alpha;                /* ICTGV2 */
  double timeSpaceWeight2;
  double dx2, dy2, dt2;
  int adaptSchedule;           /* 0: adapt when k<10 || k%50==0 (tv.cpp:172);
                                  1: every iteration (tv_dmri.m:198, debug=1);
                                  2: fixed steps sigma = tau = 0.99 / ||K_full||
                                     from an FP64 power iteration on the GPU
                                     (matlab/getopnorm.m:1-16 on the stacked
                                     regulariser + data operator), no
                                     adaptation */
  int dualStorage;             /* 0: float32 duals; 1: bfloat16 duals;
                                  2: 16-bit fixed point scaled by each dual's
                                  projection radius.  1 and 2 store 4 bytes
                                  per complex value; arithmetic is float32
                                  (no counterpart in the reference, which
                                  keeps float32 duals) */
  unsigned opnormIterations;   /* power-iteration steps of adaptSchedule 2
                                  (0 selects 50) */
} avionic_params;

typedef struct avionic_operator avionic_operator;
typedef struct avionic_recon avionic_recon;

/* Last error message of this thread ("" if none). */
const char *avionic_last_error(void);
/* Library version string. */
const char *avionic_version(void);

/* ------------------------------------------------------------------------ */
/* L2 operators (CUDA/include/base_operator.h:24-112)                        */
/* ------------------------------------------------------------------------ */

/* CartesianOperator(width, height, coils, frames, mask, centered)
 * (cartesian_operator.h:18-25, cartesian_operator.cpp:3-45).  mask may be
 * NULL (the reference's ZeroMask: fully sampled).  The mask is copied; the
 * caller may free it afterwards (the reference keeps a reference,
 * cartesian_operator.h:114). */
int avionic_cartesian_operator_create(unsigned width, unsigned height,
                                      unsigned coils, unsigned frames,
                                      const float *mask, int centered,
                                      avionic_operator **op);

/* CartesianOperator3D(width, height, depth, coils, mask)
 * (cartesian_operator3d.cpp:1-345): unitary 3D transforms over
 * [depth][height][width]; k-space and coil maps are coil-major
 * [coils][depth][height][width]; mask: device pointer [depth][height][width]
 * or NULL (fully sampled), copied at creation.  K^H (avionic_operator_forward)
 * masks a copy of its input, not the input. */
int avionic_cartesian3d_operator_create(unsigned width, unsigned height, unsigned depth,
                                        unsigned coils, const float *mask,
                                        avionic_operator **op);

/* NoncartesianOperator(width, height, coils, frames, nSpokes, nFE,
 * spokesPerFrame, kTraj, dens, sens) (noncartesian_operator.h,
 * noncartesian_operator.cpp:34-86).  traj: [frames][2][spokesPerFrame*nFE]
 * (x then y per frame, in [-0.5, 0.5)); dens: [frames][spokesPerFrame*nFE]
 * (may be NULL = 1); Kaiser-Bessel gridding with kernel width `kernel_width`
 * on a grid oversampled by `osf`. */
int avionic_noncartesian_operator_create(unsigned width, unsigned height,
                                         unsigned coils, unsigned frames,
                                         unsigned nFE, unsigned spokesPerFrame,
                                         const float *traj, const float *dens,
                                         int kernel_width, double osf,
                                         avionic_operator **op);

int avionic_operator_destroy(avionic_operator *op);

/* ForwardOperation (k-space -> image, K^H), cartesian_operator.cpp:65-111.
 * Unlike the reference the input is not modified. */
int avionic_operator_forward(avionic_operator *op, const float *kdata,
                             float *image, const float *b1, void *stream);

/* BackwardOperation (image -> k-space, K), cartesian_operator.cpp:122-160. */
int avionic_operator_backward(avionic_operator *op, const float *image,
                              float *kdata, const float *b1, void *stream);

/* K without any device allocation (no counterpart in the reference): for
 * callers that cannot allocate, such as code under HIP graph capture, tight
 * serving loops and harnesses that forbid hipMalloc.  Same result as
 * avionic_operator_backward for a Cartesian 2D+t operator
 * (avionic_cartesian_operator_create; others return AVIONIC_E_UNSUPPORTED).
 * The column pass's scratch is a 64 MiB device array of the library, one per
 * GPU, held until the process ends: the operator's compact k-space (coils x
 * sampled rows x width complex values) must fit it, else AVIONIC_E_INVALID.
 * The output is cleared by a kernel (no memset node).  Calls that overlap on
 * one GPU (other streams or host threads) race on the shared scratch; calls
 * ordered on one stream are safe. */
int avionic_operator_backward_noalloc(avionic_operator *op, const float *image,
                                      float *kdata, const float *b1, void *stream);

/* AdaptLambda(k, d): cartesian_operator.cpp:47-62,
 * noncartesian_operator.cpp:88-95. */
int avionic_operator_adapt_lambda(avionic_operator *op, double k, double d,
                                  double *lambda);

/* Number of k-space samples per (frame, coil) the operator expects. */
size_t avionic_operator_kspace_size(const avionic_operator *op);

/* ------------------------------------------------------------------------ */
/* L3 solvers (CUDA/include/pd_recon.h:88-185, tv.h, tgv2.h, ictgv2.h)       */
/* ------------------------------------------------------------------------ */

/* InitParams() defaults of each solver (tv.cpp:39-57, tgv2.cpp:36-56,
 * ictgv2.cpp:37-64); lambda is left at 0 and adaptLambda = 1. */
int avionic_default_params(avionic_method method, avionic_params *params);

/* TV/TGV2/ICTGV2(width, height, coils, frames, [params,] mrOp).
 * params == NULL selects the defaults and adapts lambda through the operator
 * (InitParams + InitLambda(true)); otherwise lambda is adapted only when
 * params->adaptLambda is set (InitLambda, tv.cpp:25-37).  The operator must
 * outlive the solver. */
int avionic_recon_create(avionic_method method, unsigned width, unsigned height,
                         unsigned coils, unsigned frames,
                         const avionic_params *params, avionic_operator *op,
                         avionic_recon **recon);
int avionic_recon_destroy(avionic_recon *recon);

/* GetParams() (pd_recon.h:115) and writes through its reference. */
int avionic_recon_get_params(const avionic_recon *recon, avionic_params *params);
int avionic_recon_set_params(avionic_recon *recon, const avionic_params *params);

/* SetVerbose / SetDebug(debug, debugstep) (pd_recon.h:148-151). */
int avionic_recon_set_verbose(avionic_recon *recon, int verbose);
int avionic_recon_set_debug(avionic_recon *recon, int debug, int debugstep);

/* AdaptLambda(k, d) (pd_recon.h:98). */
int avionic_recon_adapt_lambda(avionic_recon *recon, double k, double d,
                               double *lambda);

/* IterativeReconstruction(data, x, b1) (pd_recon.h:134-135).  data: k-space
 * of the operator's shape, already masked (main.cpp:509-519); x: in = initial
 * image, out = reconstruction.  Synchronises `stream` at the end. */
int avionic_recon_run(avionic_recon *recon, const float *data, float *x,
                      const float *b1, void *stream);

/* AdaptStepSize(extDiff..., b1) (tv.cpp:79-97, tgv2.cpp:89-128,
 * ictgv2.cpp:97-161) on a primal difference of P components (1 TV, 4 TGV2,
 * 8 ICTGV2), updating sigma/tau in the solver parameters.  Uses dx..dt2 as
 * currently set in the parameters. */
int avionic_recon_adapt_step_size(avionic_recon *recon, const float *delta,
                                  const float *b1, void *stream);

/* ComputeGStar / ComputePDGap (tv.cpp:202-235, tgv2.cpp:130-176,
 * ictgv2.cpp:178-249).  x: primal (P components for the gap, the image only
 * for G*), y: dual (3 / 9 / 18 components), z: k-space dual.  Returned as
 * double (the reference returns float). */
int avionic_recon_compute_gstar(avionic_recon *recon, const float *x1,
                                const float *y, const float *z,
                                const float *data, const float *b1,
                                double *gstar, void *stream);
int avionic_recon_compute_pdgap(avionic_recon *recon, const float *x,
                                const float *y, const float *z,
                                const float *data, const float *b1,
                                double *pdgap, void *stream);

/* ||K_full||, the norm of the stacked regulariser + data operator of this
 * solver (its method, time/space weights and operator), by `iterations` steps
 * of an FP64 power iteration on the GPU (matlab/getopnorm.m:1-16): the
 * Rayleigh quotient sqrt(<K_full^* K_full x, x>) for unit x from a
 * deterministic start.  Cartesian 2D+t and volumetric operators. */
int avionic_recon_operator_norm(avionic_recon *recon, const float *b1, unsigned iterations,
                                double *norm, void *stream);

/* Step-size history of the last run: sigma/tau after each iteration.
 * n: capacity in, entries written out. */
int avionic_recon_get_step_log(const avionic_recon *recon, double *sigma,
                               double *tau, size_t *n);

/* ExportAdditionalResults(outputDir, callback) (pd_recon.h:142,
 * ictgv2.cpp:494-516): calls cb(name, host data, complex count, user) for
 * "x3_component" (ICTGV2), and with debug on "PDGap", "ICTGVnorm",
 * "DATAfidelity".  Host buffers are valid during the call only. */
typedef void (*avionic_export_cb)(const char *name, const float *data,
                                  size_t count, void *user);
int avionic_recon_export(avionic_recon *recon, avionic_export_cb cb,
                         void *user);

/* ------------------------------------------------------------------------ */
/* L1 utilities (CUDA/include/utils.h; CUDA/src/utils.cpp)                   */
/* ------------------------------------------------------------------------ */

/* utils::Gradient (utils.cpp:99-118): u (N) -> g (3N). */
int avionic_gradient(const float *u, float *g, unsigned width, unsigned height,
                     unsigned frames, float dx, float dy, float dt,
                     void *stream);
/* utils::SymmetricGradient (utils.cpp:397-456): v (3N) -> e (6N). */
int avionic_symmetric_gradient(const float *v, float *e, unsigned width,
                               unsigned height, unsigned frames, float dx,
                               float dy, float dt, void *stream);
/* utils::Divergence (utils.cpp:733-757): p (3N) -> d (N). */
int avionic_divergence(const float *p, float *d, unsigned width,
                       unsigned height, unsigned frames, float dx, float dy,
                       float dt, void *stream);
/* utils::SymmetricDivergence (utils.cpp:961-1012): q (6N) -> d (3N). */
int avionic_symmetric_divergence(const float *q, float *d, unsigned width,
                                 unsigned height, unsigned frames, float dx,
                                 float dy, float dt, void *stream);
/* utils::ProximalMap3/6 (utils.cpp:1341-1351): y /= max(1, |y| * scale). */
int avionic_proximal_map3(float *y, size_t n, float scale, void *stream);
int avionic_proximal_map6(float *y, size_t n, float scale, void *stream);
/* utils::TVNorm / TGV2Norm / ICTGV2Norm (utils.cpp:1119-1287). */
int avionic_tv_norm(const float *u, unsigned width, unsigned height,
                    unsigned frames, float dx, float dy, float dt,
                    double *norm, void *stream);
int avionic_tgv2_norm(const float *u, const float *v, unsigned width,
                      unsigned height, unsigned frames, double alpha0,
                      double alpha1, float dx, float dy, float dt,
                      double *norm, void *stream);
int avionic_ictgv2_norm(const float *u, const float *v, const float *w,
                        const float *v2, unsigned width, unsigned height,
                        unsigned frames, double alpha0, double alpha1,
                        double alpha, float dx, float dy, float dt, float dx2,
                        float dy2, float dt2, double *norm, void *stream);
/* utils::Ellipke (utils.cpp:18-47); returns NaN outside [0, 1]. */
double avionic_ellipke(double m);
/* PDRecon::ComputeTimeSpaceWeights (pd_recon.cpp:72-98). */
int avionic_time_space_weights(double timeSpaceWeight, double *ds, double *dt);

/* In-place 2D FFT of `batch` height x width images (AGILE FFT, as used by
 * cartesian_operator.cpp:90-99,142-151).  sign -1: e^{-i}, +1: e^{+i};
 * unitary scales by 1/sqrt(width*height). */
int avionic_fft2(float *data, unsigned height, unsigned width, unsigned batch,
                 int sign, int centered, int unitary, void *stream);

/* ------------------------------------------------------------------------ */
/* Coil-sensitivity estimation (CUDA/include/coil_construction.h:58-141,      */
/* CUDA/src/coil_construction.cpp, cartesian_coil_construction.cpp)          */
/* ------------------------------------------------------------------------ */
/* Images are device arrays [coils][height][width] (one [height][width] image
 * for u0).  Arithmetic is float as in the reference. */

/* CoilConstructionParams (coil_construction.h:18-41). */
typedef struct avionic_coil_params {
  double uH1mu;        /* H1 weight of B1FromUH1 */
  double uReg;         /* TGV weight nu of UTGV2Recon */
  unsigned uNrIt;      /* UTGV2Recon iterations */
  double uSigma, uTau, uAlpha0, uAlpha1;
  double b1Reg;        /* mu of the per-coil B1Recon */
  unsigned b1NrIt;
  double b1Sigma, b1Tau;
  double b1FinalReg;   /* mu of the final B1Recon over all coils */
  unsigned b1FinalNrIt;
} avionic_coil_params;

/* CoilConstruction::InitParams (coil_construction.cpp:22-46). */
int avionic_default_coil_params(avionic_coil_params *params);

/* TimeAveragedReconstruction.  Cartesian 2D+t operator
 * (CartesianCoilConstruction, cartesian_coil_construction.cpp:23-87): crec_c =
 * the image transform of sum_t kdata[t][c] / max(1, sum_t M_t) with the
 * operator's mask and centring; kdata [frames][coils][h][w].  Non-Cartesian
 * operator (NoncartesianCoilConstruction, noncartesian_coil_construction.cpp
 * :21-123): crec_c = the adjoint gridding of every frame's points of coil c
 * together, with the operator's trajectory and density, kernel width 3 and
 * grid oversampling 2 whatever the operator's own; kdata [frames][coils][M],
 * sqrt(density)-weighted as the operator takes it.  Both: u = sqrt(sum_c
 * |crec_c|^2), with the phase of sum_c crec_c when apply_phase. */
int avionic_coil_time_averaged(avionic_operator *op, const float *kdata, float *u, float *crec,
                               int apply_phase, void *stream);
/* CoilConstruction::B1FromUH1 (coil_construction.cpp:363-446): per coil the
 * solution of (mu |u|^2 + L) x = mu conj(u) |crec_c| (L the periodic 2D
 * Laplacian; conjugate gradients, squared-residual tolerance 1e-12, at most
 * 800 iterations), returned as |x_c| / sqrt(sum_c |x_c|^2). */
int avionic_coil_b1_from_uh1(unsigned width, unsigned height, unsigned coils, const float *u,
                             const float *crec, double mu, float *b1, void *stream);
/* CoilConstruction::FindMaximumSum (coil_construction.cpp:52-73). */
int avionic_coil_find_max_sum(unsigned width, unsigned height, unsigned coils, const float *crec,
                              int *index, void *stream);
/* CoilConstruction::UTGV2Recon (coil_construction.cpp:75-216): params->uNrIt
 * PD iterations of min_u nu TGV2(u) + 1/2 sum_{c in inds} |b1_c u - crec_c|^2
 * from u = x (in/out).  inds: host array of coil indices into b1 / crec. */
int avionic_coil_utgv2(unsigned width, unsigned height, unsigned coils, const float *b1,
                       const float *crec, const int *inds, int n_inds,
                       const avionic_coil_params *params, float *x, void *stream);
/* CoilConstruction::B1Recon (coil_construction.cpp:217-362): maxIt PD
 * iterations per coil of min_b mu/2 |grad b|^2 + 1/2 |b u_n - crec_c|^2,
 * u_n = u0 / max |u0|, from b = 0, then normalised to unit sum of squares
 * over the coils.  b1 (out): [coils][h][w]. */
int avionic_coil_b1_recon(unsigned width, unsigned height, unsigned coils, const float *u0,
                          const float *crec, unsigned maxIt, double mu,
                          const avionic_coil_params *params, float *b1, void *stream);
/* CoilConstruction::PerformCoilConstruction (coil_construction.cpp:447-557)
 * on a Cartesian 2D+t or non-Cartesian operator, from its time-averaged coil
 * images (avionic_coil_time_averaged, with phase): u (out) [coils][h][w] holds the image
 * estimate after each coil step (the reference's CLI takes the last one as
 * u0, main.cpp:639), b1 (out) the coil maps; order (host, optional, `coils`
 * ints) the greedy coil order.  params NULL selects the defaults. */
int avionic_coil_construction(avionic_operator *op, const float *kdata,
                              const avionic_coil_params *params, float *u, float *b1, int *order,
                              void *stream);
/* PerformCoilConstruction from given time-averaged images: crec
 * [coils][h][w] and u0 [h][w] as TimeAveragedReconstruction returns them
 * (the non-Cartesian path of the reference grids every frame's points into
 * one image per coil, noncartesian_coil_construction.cpp:21-123). */
int avionic_coil_construction_from_images(unsigned width, unsigned height, unsigned coils,
                                          const float *crec, const float *u0,
                                          const avionic_coil_params *params, float *u,
                                          float *b1, int *order, void *stream);

/* ------------------------------------------------------------------------ */
/* Raw-data normalisation (host memory; no device needed)                    */
/* ------------------------------------------------------------------------ */
/* RawDataPreparation::FindNormalizationFactor (raw_data_preparation.cpp
 * :661-686, utils::Median utils.cpp:1442-1462): the median of the magnitudes
 * >= 0.9 * max, taken from the first maximum onwards (values before it are not
 * counted, as in the reference).  The CLI's -o scales the data by
 * 255 / median (Cartesian, :795-818) or 255 * frames / median (radial,
 * :736-793).  Fails on n == 0 or a non-positive median. */
int avionic_normalization_factor(const float *magnitude, size_t n, float *median);

/* ------------------------------------------------------------------------ */
/* .bin vector files (host memory; no device needed)                         */
/* ------------------------------------------------------------------------ */
/* The reference's vector file: uint8 iscomplex, uint32 n, n float64 real
 * parts, then (iscomplex) n float64 imaginary parts (matlab/writebin_vector.m
 * :35-44, readbin_vector.m:30-40, python/utils.py:23-48; the CLI writes its
 * results with it, main.cpp:831-839).
 * avionic_bin_read: with re == NULL only *n and *iscomplex are returned;
 * otherwise re (and im, if the file is complex and im != NULL) receive
 * min(n, capacity) values.  A real file read with im != NULL zero-fills im. */
int avionic_bin_read(const char *path, double *re, double *im, size_t capacity,
                     size_t *n, int *iscomplex);
/* im == NULL writes a real vector. */
int avionic_bin_write(const char *path, const double *re, const double *im,
                      size_t n);

/* ------------------------------------------------------------------------ */
/* Batched execution and serving (no reference counterpart: the reference   */
/* runs one reconstruction per process, main.cpp:789)                        */
/* ------------------------------------------------------------------------ */
/* Engine calls are thread-safe: each holds the engine's lock for its whole
 * duration, so one thread may submit while another steps or polls. */

typedef struct avionic_engine avionic_engine;

/* Admission.  A submitted request is held (its inputs copied, its state
 * allocated) while the bytes of all held requests stay within
 * hbm_budget_bytes; otherwise submit returns AVIONIC_E_CAPACITY.  Held
 * requests join the batched launch set in submission order (FIFO) as long as
 * the running requests number at most max_batch and their bytes stay within
 * resident_bytes_cap, which bounds the work of one step and so the latency of
 * each running request; the others wait in the queue and are admitted by
 * avionic_engine_step as running requests finish.  A request always runs
 * when nothing else does. */
typedef struct avionic_engine_config {
  size_t hbm_budget_bytes;  /* bytes of held requests; 0: 90% of the device
                               memory free at creation */
  int use_graphs;           /* capture each iteration in a hipGraph */
  int max_batch;            /* max running requests; <= 0: no limit */
  int iterations_per_tick;  /* iterations run per engine step */
  size_t resident_bytes_cap; /* bytes of running requests; 0: no limit */
  int max_queue;            /* waiting requests beyond which submit returns
                               AVIONIC_E_CAPACITY; < 0: no limit */
  int voxels_per_lane;      /* dual / primal updates: 1 voxel per lane, or
                               0: a pair of voxels of a row per lane (16- and
                               8-byte accesses) for requests of even width */
  int fixed_launch;         /* 1: each launch set runs the default launch
                               variant; 0: a launch set that has run 8 plain
                               iterations unchanged times its variants (pairs
                               or one voxel per lane; the column pass before
                               or after the dual update; all give the same
                               bits) over its next plain iterations and keeps
                               the fastest for launch sets of its method, dual
                               storage and grid sizes (to a factor of 2) */
} avionic_engine_config;

/* Engine counters (avionic_engine_counters). */
typedef struct avionic_engine_counts {
  int held;                 /* requests held: queued, running or finished */
  int running;              /* in the launch set */
  int queued;               /* waiting for admission */
  size_t bytes_held, bytes_running, hbm_budget_bytes;
  long long graph_captures; /* iteration graphs captured and instantiated */
  double capture_ms;        /* host time spent capturing them */
  long long launch_tunings; /* launch sets whose variants were timed */
} avionic_engine_counts;

typedef struct avionic_request_desc {
  avionic_method method;
  unsigned width, height, coils, frames;
  const avionic_params *params;  /* NULL = defaults + adapted lambda */
  const float *mask;             /* device, [frames][height][width] or NULL */
  int centered;
  const float *data;             /* device, [frames][coils][height][width] */
  const float *b1;               /* device, [coils][height][width] */
  const float *x0;               /* device, [frames][height][width] or NULL */
  void *stream;                  /* stream that produced the inputs; submit
                                    returns once they are copied */
  /* In-loop PD gap, as PDRecon's SetVerbose / SetDebug / stopPDGap
   * (ictgv2.cpp:462-485): the gap is evaluated after iteration k when
   * verbose and (k < 10 || k % 50 == 0), when debug and k % debugstep == 0,
   * or when params->stopPDGap > 0 and k % 20 == 0; verbose prints it; the
   * request finishes early once a gap falls below stopPDGap.  The gaps
   * (with the regulariser norm and data fidelity) are kept for
   * avionic_engine_get_gap_log.  Each evaluation synchronises the engine. */
  int verbose;
  int debug;
  int debugstep;                 /* <= 0 reads as 1 */
  /* Optional operator of the request (NULL: a Cartesian operator built
   * from mask / centered).  A Cartesian 2D+t or non-Cartesian operator of
   * the request's width, height, coils and frames; it must outlive the
   * request, and requests sharing one operator share its scratch (they run
   * one after the other on the engine's stream).  data is the operator's
   * dense k-space ([frames][coils][M] for non-Cartesian operators). */
  const avionic_operator *op;
} avionic_request_desc;

int avionic_engine_create(const avionic_engine_config *cfg,
                          avionic_engine **engine);
int avionic_engine_destroy(avionic_engine *engine);
/* HBM bytes a request of this shape keeps resident. */
size_t avionic_request_footprint(const avionic_request_desc *desc);
/* Hold a request (copies its inputs) and admit it now or queue it (see
 * avionic_engine_config).  AVIONIC_E_CAPACITY if it does not fit the HBM
 * budget or the queue now; the caller may retry after others finish. */
int avionic_engine_submit(avionic_engine *engine,
                          const avionic_request_desc *desc, long long *id);
/* Admit queued requests that fit, then run up to iterations_per_tick
 * iterations of every running request as one batched launch sequence.
 * Returns the number of requests held. */
int avionic_engine_step(avionic_engine *engine, void *stream);
/* 1 done (result copied to x_out, device [frames][height][width]), 0 running,
 * AVIONIC_E_NOTFOUND unknown.  A finished request is released on retrieval. */
int avionic_engine_poll(avionic_engine *engine, long long id, float *x_out,
                        void *stream);
int avionic_engine_cancel(avionic_engine *engine, long long id);
/* Requests held / bytes in use. */
int avionic_engine_stats(const avionic_engine *engine, int *resident,
                         size_t *bytes_in_use);
int avionic_engine_counters(const avionic_engine *engine, avionic_engine_counts *counts);
/* Average duration in ms of each kernel of one iteration (dual,
 * col_forward, row_dual_update, col_adjoint, primal, step_reg,
 * col_forward_delta, row_norm2, tail) over the first resident batch, each
 * launched `reps` times alone between HIP events.  Advances the resident
 * state: use a dedicated engine.  n: capacity in, count out. */
int avionic_engine_time_kernels(avionic_engine *engine, int reps, double *ms,
                                const char **names, size_t *n);
/* sigma/tau after each iteration run so far (n: capacity in, count out);
 * available until the request is polled or cancelled. */
int avionic_engine_get_step_log(const avionic_engine *engine, long long id,
                                double *sigma, double *tau, size_t *n);
/* PD gaps evaluated so far (see avionic_request_desc), each normalised by the
 * voxel count, with the regulariser norm and the data fidelity
 * lambda/2 ||K x - d|| at the same iterations (the reference's PDGap,
 * ICTGVnorm and DATAfidelity exports, ictgv2.cpp:494-516); any of the
 * arrays may be NULL.  n: capacity in, count out; available until the
 * request is polled or cancelled. */
int avionic_engine_get_gap_log(const avionic_engine *engine, long long id, double *gap,
                               double *norm, double *fidelity, size_t *n);

#ifdef __cplusplus
}
#endif
#endif /* AVIONIC_H_ */
